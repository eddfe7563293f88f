"""Pure-Python sparse convolution straight from voxel coordinates.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Independent of any pair
table, mask or argsort: a dict maps coordinates to rows, and every output row
sums ``W[:, f, :] @ x[neighbour]`` over the kernel offsets (kv = ks**3, ks
odd, read from W's shape).  Used on tiny inputs (N <= a few hundred) to pin
the pair builder and the FP64 oracle together.  Offsets follow
``f = (kz*ks + ky)*ks + kx`` with neighbour ``coord + (k - ks//2)``
(cross-correlation), the enumeration of cumm/conv/main_real_spconv.py's
igemm tables; strided layers use ``in = out*stride - padding + k``.
"""
from __future__ import annotations

import itertools

import numpy as np


def _ksize(kv: int) -> int:
    ks = int(round(kv ** (1.0 / 3.0)))
    assert ks ** 3 == kv and ks % 2 == 1, f"kv={kv} is not an odd cube"
    return ks


def conv_forward_coords(in_coords, X, W, out_coords=None, stride: int = 1, padding: int = 1,
                        subm: bool = True) -> np.ndarray:
    X = np.asarray(X, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    K, kv, C = W.shape
    out = np.zeros((len(in_coords if subm else out_coords), K))
    for o, f, i in _neighbours(in_coords, out_coords, stride, padding, subm, _ksize(kv)):
        for k in range(K):
            s = 0.0
            for c in range(C):
                s += W[k, f, c] * X[i, c]
            out[o, k] += s
    return out


def _neighbours(in_coords, out_coords, stride: int, padding: int, subm: bool, ks: int = 3):
    """Yield (o, f, i) for every output row o, offset f and input row i that the
    coordinate rule connects, offsets f = (kz*ks + ky)*ks + kx."""
    lut = {tuple(int(v) for v in c): i for i, c in enumerate(in_coords)}
    outs = in_coords if subm else out_coords
    offs = list(itertools.product(range(ks), range(ks), range(ks)))
    h = ks // 2
    for o, oc in enumerate(outs):
        b = int(oc[0])
        for f, (kz, ky, kx) in enumerate(offs):
            if subm:
                nb = (b, int(oc[1]) + kz - h, int(oc[2]) + ky - h, int(oc[3]) + kx - h)
            else:
                nb = (b, int(oc[1]) * stride - padding + kz, int(oc[2]) * stride - padding + ky,
                      int(oc[3]) * stride - padding + kx)
            i = lut.get(nb)
            if i is not None:
                yield o, f, i


def conv_dgrad_coords(in_coords, dO, W, out_coords=None, stride: int = 1, padding: int = 1,
                      subm: bool = True) -> np.ndarray:
    """dX[i] = sum over (o, f) connected to i of W[:, f, :]^T dO[o], by loops."""
    dO = np.asarray(dO, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    K, kv, C = W.shape
    dx = np.zeros((len(in_coords), C))
    for o, f, i in _neighbours(in_coords, out_coords, stride, padding, subm, _ksize(kv)):
        for c in range(C):
            s = 0.0
            for k in range(K):
                s += W[k, f, c] * dO[o, k]
            dx[i, c] += s
    return dx


def conv_wgrad_coords(in_coords, X, dO, kv: int, out_coords=None, stride: int = 1, padding: int = 1,
                      subm: bool = True) -> np.ndarray:
    """dW[k, f, c] = sum over connected (o, i) of dO[o, k] * X[i, c], by loops."""
    X = np.asarray(X, dtype=np.float64)
    dO = np.asarray(dO, dtype=np.float64)
    K, C = dO.shape[1], X.shape[1]
    dw = np.zeros((K, kv, C))
    for o, f, i in _neighbours(in_coords, out_coords, stride, padding, subm, _ksize(kv)):
        for k in range(K):
            for c in range(C):
                dw[k, f, c] += dO[o, k] * X[i, c]
    return dw
