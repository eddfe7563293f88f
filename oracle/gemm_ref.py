"""FP64 restatement of GemmMainUnitTest.matmul2 and the reference's blocked
FP32 product.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinning (tests/test_oracle.py): numpy FP64 ``a @ b`` (cumm/gemm/main_simu.py:
643-647); the ShuffleAC / ShuffleAB closed forms ``c[c_inds] = a[a_inds] @ b``
and ``c = a[a_inds]^T @ b[b_inds]`` (cumm/gemm/main_real.py:197-229); the
self-cancelling known answer ``A.B + 1 * (-A.B) == 0`` (main_simu.py:720,
806-811); permutation / identity index special cases; Hilbert inputs
(main_real.py:457-475).
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from .spconv_ref import epilogue


def op(x: np.ndarray, trans: bool) -> np.ndarray:
    """cumm's trans flags: trans_a=True means ``a`` is stored [k, m]; trans_b=True
    means ``b`` is stored [n, k] (cumm/gemm/nvrtc_code.py:80-123)."""
    x = np.asarray(x, dtype=np.float64)
    return x.T if trans else x


def matmul2(a, b, c_init, trans_a: bool = False, trans_b: bool = False, shuffle: int = 0,
            a_inds=None, b_inds=None, c_inds=None, d=None, alpha: float = 1.0, beta: float = 0.0,
            act_type: int = 0, act_alpha: float = 0.0) -> np.ndarray:
    """Returns C FP64.

    NoShuffle:  C = act(alpha * op(a) @ op(b) + beta * D)
    ShuffleAC:  C[c_inds[i]] = act(alpha * a[a_inds[i]] @ op(b) + beta * D[c_inds[i]])
                (a_inds empty -> identity; other rows of C untouched)
    ShuffleAB:  C = act(alpha * a[a_inds]^T @ b[b_inds] + beta * D)
    D defaults to C; a 1-D D is a bias broadcast over rows
    (cumm/gemm/nvrtc_code.py:44-46, 126-128; cumm/gemm/kernel.py:293-327).
    """
    c = np.array(c_init, dtype=np.float64, copy=True)
    D = c if d is None else np.asarray(d, dtype=np.float64)
    if shuffle == 1:
        A = np.asarray(a, dtype=np.float64)
        if a_inds is not None:
            A = A[np.asarray(a_inds)]
        acc = A @ op(b, trans_b)
        m = acc.shape[0]
        rows = np.asarray(c_inds)[:m]
        src = None
        if beta != 0.0:
            src = np.broadcast_to(D, acc.shape) if D.ndim == 1 else D[rows]
        c[rows] = epilogue(acc, alpha, beta, src, act_type, act_alpha)
        return c
    if shuffle == 2:
        A = np.asarray(a, dtype=np.float64)[np.asarray(a_inds)]
        B = np.asarray(b, dtype=np.float64)[np.asarray(b_inds)]
        acc = A.T @ B
    else:
        acc = op(a, trans_a) @ op(b, trans_b)
    src = None
    if beta != 0.0:
        src = np.broadcast_to(D, acc.shape) if D.ndim == 1 else D
    return epilogue(acc, alpha, beta, src, act_type, act_alpha)


def py_matmul(a: np.ndarray, b: np.ndarray, acc_type=np.float32, tbk: int = 8) -> np.ndarray:
    """cumm/gemm/main_real.py:484-498: per output, accumulate k in blocks of
    ``tbk`` into separate partials, then sum the partials in order (FP32).
    Vectorised over (i, j); the per-element operation order is unchanged."""
    a = np.asarray(a, dtype=acc_type)
    b = np.asarray(b, dtype=acc_type)
    m, k = a.shape
    n = b.shape[1]
    nparts = k // tbk + 1
    acc = np.zeros((nparts, m, n), dtype=acc_type)
    for nk in range(k):
        acc[nk // tbk] += (a[:, nk:nk + 1] * b[nk:nk + 1, :]).astype(acc_type)
    ret = np.zeros((m, n), dtype=acc_type)
    for l in range(nparts):
        ret += acc[l]
    return ret


def hilbert(m: int, n: int) -> np.ndarray:
    """gen_row_hilbert_mat (cumm/gemm/main_real.py:457-462)."""
    i = np.arange(m)[:, None]
    j = np.arange(n)[None, :]
    return 1.0 / (i + j + 1.0)
