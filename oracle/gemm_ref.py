"""FP64 restatement of GemmMainUnitTest.matmul2 and the reference's blocked
FP32 product.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinning (tests/test_oracle.py): numpy FP64 ``a @ b`` (cumm/gemm/main_simu.py:
643-647); the ShuffleAC / ShuffleAB closed forms ``c[c_inds] = a[a_inds] @ b``
and ``c = a[a_inds]^T @ b[b_inds]`` (cumm/gemm/main_real.py:197-229); the
self-cancelling known answer ``A.B + 1 * (-A.B) == 0`` (main_simu.py:720,
806-811); permutation / identity index special cases; Hilbert inputs
(main_real.py:457-475).  The int8 forms (matmul2_int / int_epilogue): brute
force over explicit Python integer loops on small inputs, equality with the
FP64 form where no value saturates or wraps, and hand-computed saturation,
truncation and wrap-around cases.
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


def _wrap32(x: np.ndarray) -> np.ndarray:
    """Two's-complement int32 wrap of an int64 array."""
    return ((np.asarray(x, dtype=np.int64) + (1 << 31)) % (1 << 32) - (1 << 31)).astype(np.int64)


def int_epilogue(acc, alpha: float, beta: float, src, act_type: int = 0, act_alpha: float = 0.0,
                 out_dtype=np.int8) -> np.ndarray:
    """LinearCombination with int32 accumulator and compute type
    (cumm/gemm/output_op/linear.py:112-140, dtype_comp s32): alpha, beta and
    act_alpha are the float parameters converted to int32 by truncation toward
    zero (the output op's int32 members, linear.py:52-72); X = beta * src, then
    D = alpha * acc + X, in int32 (two's-complement wrap); ReLU max(x, 0) /
    LeakyReLU x >= 0 ? x : x * act_alpha on int32
    (cumm/gemm/gemmmath/__init__.py:176-200); int8 output saturates to
    [-128, 127] (include/tensorview/gemm/math/numeric_convert.h:1091-1150),
    int32 output keeps the value.  Sigmoid has no integer form."""
    if act_type == 2:
        raise ValueError("sigmoid is not defined on int32")
    ia, ib, iact = int(np.trunc(alpha)), int(np.trunc(beta)), int(np.trunc(act_alpha))
    x = _wrap32(ia * np.asarray(acc, dtype=np.int64))
    if ib != 0 and src is not None:
        x = _wrap32(x + _wrap32(ib * np.asarray(src, dtype=np.int64)))
    if act_type == 1:
        x = np.maximum(x, 0)
    elif act_type == 3:
        x = np.where(x >= 0, x, _wrap32(x * iact))
    if np.dtype(out_dtype) == np.int8:
        return np.clip(x, -128, 127).astype(np.int8)
    return x.astype(np.int32)


def matmul2_int(a, b, c_init, trans_a: bool = False, trans_b: bool = False, shuffle: int = 0,
                a_inds=None, b_inds=None, c_inds=None, d=None, alpha: float = 1.0, beta: float = 0.0,
                act_type: int = 0, act_alpha: float = 0.0) -> np.ndarray:
    """int8 GEMMs "s8,s8,s8,s32,s32" / "s8,s8,s32,s32,s32"
    (cumm/gemm/main.py:348-366, 499-523): the shuffle forms of matmul2 above
    with the exact int32 accumulator (int64 products and sums, wrapped to
    int32) and int_epilogue.  The output dtype is c_init's (int8 or int32)."""
    c = np.array(c_init, copy=True)
    out_dtype = c.dtype
    D = c if d is None else np.asarray(d)

    def t(x, tr):
        x = np.asarray(x, dtype=np.int64)
        return x.T if tr else x

    if shuffle == 1:
        A = np.asarray(a, dtype=np.int64)
        if a_inds is not None:
            A = A[np.asarray(a_inds)]
        acc = _wrap32(A @ t(b, trans_b))
        rows = np.asarray(c_inds)[:acc.shape[0]]
        src = None if int(np.trunc(beta)) == 0 else (np.broadcast_to(D, acc.shape) if D.ndim == 1 else D[rows])
        c[rows] = int_epilogue(acc, alpha, beta, src, act_type, act_alpha, out_dtype)
        return c
    if shuffle == 2:
        A = np.asarray(a, dtype=np.int64)[np.asarray(a_inds)]
        B = np.asarray(b, dtype=np.int64)[np.asarray(b_inds)]
        acc = _wrap32(A.T @ B)
    else:
        acc = _wrap32(t(a, trans_a) @ t(b, trans_b))
    src = None if int(np.trunc(beta)) == 0 else (np.broadcast_to(D, acc.shape) if D.ndim == 1 else D)
    return int_epilogue(acc, alpha, beta, src, act_type, act_alpha, out_dtype)


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
