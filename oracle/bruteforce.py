"""Pure-Python sparse convolution straight from voxel coordinates.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Independent of any pair
table, mask or argsort: a dict maps coordinates to rows, and every output row
sums ``W[:, f, :] @ x[neighbour]`` over the 27 offsets.  Used on tiny inputs
(N <= a few hundred) to pin the pair builder and the FP64 oracle together.
Offsets follow ``f = kz*9 + ky*3 + kx`` with neighbour ``coord + (k - 1)``
(cross-correlation), the enumeration of cumm/conv/main_real_spconv.py's
igemm tables; strided layers use ``in = out*stride - padding + k``.
"""
from __future__ import annotations

import itertools

import numpy as np


def conv_forward_coords(in_coords, X, W, out_coords=None, stride: int = 1, padding: int = 1,
                        subm: bool = True) -> np.ndarray:
    X = np.asarray(X, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    K, kv, C = W.shape
    lut = {tuple(int(v) for v in c): i for i, c in enumerate(in_coords)}
    outs = in_coords if subm else out_coords
    out = np.zeros((len(outs), K))
    offs = list(itertools.product(range(3), range(3), range(3)))
    for o, oc in enumerate(outs):
        b = int(oc[0])
        for f, (kz, ky, kx) in enumerate(offs):
            if subm:
                nb = (b, int(oc[1]) + kz - 1, int(oc[2]) + ky - 1, int(oc[3]) + kx - 1)
            else:
                nb = (b, int(oc[1]) * stride - padding + kz, int(oc[2]) * stride - padding + ky,
                      int(oc[3]) * stride - padding + kx)
            i = lut.get(nb)
            if i is None:
                continue
            for k in range(K):
                s = 0.0
                for c in range(C):
                    s += W[k, f, c] * X[i, c]
                out[o, k] += s
    return out
