"""Synthetic voxel scenes and the host-side pair / mask / argsort builder.

This module plays the role of the *caller* of the sparse-conv path: in the
reference that caller is spconv, which builds the implicit-GEMM pair table,
the per-row kernel-offset masks and the mask argsort before it calls
``implicit_gemm2`` (cumm/conv/main_real_spconv.py:91-147).  spconv is not part
of the reference tree, so the formats below are fixed by how cumm consumes them:

* ``pair[kv, N_out]`` int32, ``-1`` for a missing neighbour
  (cumm/conv/nvrtc_code.py:127 checks ``{kernel_volume, -1}``;
  cumm/conv/sparse_iters.py:629-631 zero-fills ``-1``).
* ``mask[N_out]`` uint32, bit ``f`` set iff ``pair[f][row] != -1``, stored in
  *sorted* order together with ``argsort`` (cumm/conv/main_real_spconv.py:141-143).
* offsets enumerate ``f = kz*9 + ky*3 + kx`` with neighbour ``coord + (k - 1)``
  (cross-correlation, centre ``f = 13``).

Everything here is numpy on the host; it prepares inputs and is not on the
device hot path.
"""
from __future__ import annotations

import dataclasses
from typing import Optional, Sequence, Tuple

import numpy as np

KV3 = 27
GRID_KITTI = (41, 1600, 1408)  # (z, y, x) spatial shape of the benchmark scenes


# --------------------------------------------------------------------------
# scene generation
# --------------------------------------------------------------------------
def _plane_points(rng: np.random.Generator, grid: Sequence[int], m: int,
                  noise: float) -> np.ndarray:
    """``m`` in-grid points on one random plane, plus isotropic N(0, noise) jitter.

    The plane passes through a uniform point of the grid box with a normal drawn
    uniformly on the sphere; points are uniform over the part of the plane that
    lies inside the box: rejection from the rectangle (in plane coordinates)
    that bounds the plane's intersection with the box, widened by 6 noise
    standard deviations so that jittered points from just outside the box keep
    their (negligible) chance of landing inside.
    """
    g = np.asarray(grid, dtype=np.float64)
    normal = rng.normal(size=3)
    normal /= np.linalg.norm(normal)
    helper = np.array([1.0, 0.0, 0.0]) if abs(normal[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(normal, helper)
    u /= np.linalg.norm(u)
    v = np.cross(normal, u)
    center = rng.uniform(0.0, 1.0, size=3) * g
    # plane / box-edge intersections, in plane coordinates (s, t)
    corners = np.array([[i, j, k] for i in (0, 1) for j in (0, 1) for k in (0, 1)], dtype=np.float64) * g
    st_pts = []
    for a in range(8):
        for b in range(a + 1, 8):
            d = corners[b] - corners[a]
            if np.count_nonzero(d) != 1:
                continue  # not an edge
            den = float(normal @ d)
            if abs(den) < 1e-12:
                continue
            lam = float(normal @ (center - corners[a])) / den
            if 0.0 <= lam <= 1.0:
                q = corners[a] + lam * d - center
                st_pts.append((float(q @ u), float(q @ v)))
    st_pts = np.asarray(st_pts)
    pad = 6.0 * noise + 1e-6
    lo = st_pts.min(axis=0) - pad
    hi = st_pts.max(axis=0) + pad
    out = []
    have = 0
    while have < m:
        batch = max(4 * (m - have), 65536)
        st = rng.uniform(lo, hi, size=(batch, 2))
        p = center + st[:, :1] * u + st[:, 1:] * v
        p = p + rng.normal(0.0, noise, size=p.shape)
        ok = np.all((p >= 0.0) & (p < g), axis=1)
        p = p[ok]
        out.append(p)
        have += len(p)
    return np.concatenate(out)[:m]


def plane_scene(n_voxels: int, grid: Sequence[int] = GRID_KITTI, n_planes: int = 8,
                noise: float = 0.5, seed: int = 0, batch_index: int = 0) -> np.ndarray:
    """Exactly ``n_voxels`` unique voxel coordinates ``[N, 4]`` int32 (b, z, y, x).

    Points are drawn on ``n_planes`` random planes spanning the grid, jittered by
    N(0, ``noise``) voxels, floored to voxels and de-duplicated; a seeded subset of
    exactly ``n_voxels`` unique voxels is kept.  Rows come out sorted by
    (z, y, x), the order a hash-free voxelizer that sorts keys would produce.
    """
    rng = np.random.default_rng(seed)
    zyx = np.asarray(grid, dtype=np.int64)
    per_plane = max(1, int(np.ceil(1.25 * n_voxels / n_planes)))
    while True:
        pts = np.concatenate([_plane_points(rng, grid, per_plane, noise) for _ in range(n_planes)])
        vox = np.floor(pts).astype(np.int64)
        vox = np.clip(vox, 0, zyx - 1)
        key = (vox[:, 0] * zyx[1] + vox[:, 1]) * zyx[2] + vox[:, 2]
        key = np.unique(key)
        if len(key) >= n_voxels:
            break
        per_plane = int(per_plane * 1.5) + 1
    keep = np.sort(rng.choice(len(key), size=n_voxels, replace=False))
    key = key[keep]
    z = key // (zyx[1] * zyx[2])
    rem = key % (zyx[1] * zyx[2])
    y = rem // zyx[2]
    x = rem % zyx[2]
    coords = np.stack([np.full_like(z, batch_index), z, y, x], axis=1).astype(np.int32)
    return coords


# --------------------------------------------------------------------------
# pair tables
# --------------------------------------------------------------------------
def _keys(coords: np.ndarray, shape: Sequence[int]) -> np.ndarray:
    c = coords.astype(np.int64)
    s = [int(v) for v in shape]
    return ((c[:, 0] * s[0] + c[:, 1]) * s[1] + c[:, 2]) * s[2] + c[:, 3]


def _lookup(sorted_keys: np.ndarray, order: np.ndarray, query: np.ndarray,
            valid: np.ndarray) -> np.ndarray:
    pos = np.searchsorted(sorted_keys, query)
    pos = np.minimum(pos, len(sorted_keys) - 1) if len(sorted_keys) else pos
    hit = valid.copy()
    if len(sorted_keys):
        hit &= sorted_keys[pos] == query
    else:
        hit[:] = False
    out = np.full(query.shape, -1, dtype=np.int32)
    out[hit] = order[pos[hit]].astype(np.int32)
    return out


def kernel_offsets(ksize: Sequence[int] = (3, 3, 3)) -> np.ndarray:
    """Offsets ``[kv, 3]`` in enumeration order ``f = (kz*ky_n + ky)*kx_n + kx``."""
    kz, ky, kx = np.meshgrid(*[np.arange(k) for k in ksize], indexing="ij")
    return np.stack([kz.ravel(), ky.ravel(), kx.ravel()], axis=1).astype(np.int64)


def subm_pairs(coords: np.ndarray, spatial_shape: Sequence[int],
               ksize: Sequence[int] = (3, 3, 3), dilation: int = 1) -> np.ndarray:
    """Submanifold implicit-GEMM pair table ``[kv, N]``: ``pair[f][i]`` is the
    input row at ``coord_i + dilation*(k_f - centre)`` or ``-1``."""
    n = coords.shape[0]
    offs = kernel_offsets(ksize)
    centre = np.array([k // 2 for k in ksize], dtype=np.int64)
    keys = _keys(coords, spatial_shape)
    order = np.argsort(keys, kind="stable")
    sk = keys[order]
    shp = np.asarray(spatial_shape, dtype=np.int64)
    pair = np.full((len(offs), n), -1, dtype=np.int32)
    c = coords.astype(np.int64)
    for f, o in enumerate(offs):
        nb = c.copy()
        nb[:, 1:] += dilation * (o - centre)
        valid = np.all((nb[:, 1:] >= 0) & (nb[:, 1:] < shp), axis=1)
        pair[f] = _lookup(sk, order, _keys(nb, spatial_shape), valid)
    return pair


def strided_out_coords(coords: np.ndarray, spatial_shape: Sequence[int],
                       ksize: Sequence[int] = (3, 3, 3), stride: int = 2,
                       padding: int = 1, rule: str = "reach") -> Tuple[np.ndarray, Tuple[int, int, int]]:
    """Output voxels of a regular strided sparse conv, sorted by (b, z, y, x).

    ``rule="reach"``: every output position whose window
    ``o*stride - padding + k`` holds at least one active input (the active set
    dilates).  ``rule="down"``: the 2x-downsampled set ``unique(floor(c /
    stride))`` (each input's own cell; with padding 1 and k = 3 every input lies
    in its cell's window).  The output spatial shape is
    ``(S + 2*padding - k) // stride + 1`` either way."""
    shp = np.asarray(spatial_shape, dtype=np.int64)
    ks = np.asarray(ksize, dtype=np.int64)
    out_shape = tuple(int(v) for v in (shp + 2 * padding - ks) // stride + 1)
    c = coords.astype(np.int64)
    if rule == "down":
        oc = np.concatenate([c[:, :1], c[:, 1:] // stride], axis=1)
        oc[:, 1:] = np.minimum(oc[:, 1:], np.asarray(out_shape) - 1)
        _, first = np.unique(_keys(oc, out_shape), return_index=True)
        return oc[first].astype(np.int32), out_shape
    if rule != "reach":
        raise ValueError(f"unknown output rule {rule!r}")
    cands = []
    for o in kernel_offsets(ksize):
        num = c[:, 1:] + padding - o
        ok = np.all((num % stride == 0), axis=1)
        oc = num // stride
        ok &= np.all((oc >= 0) & (oc < np.asarray(out_shape)), axis=1)
        cands.append(np.concatenate([c[ok, :1], oc[ok]], axis=1))
    allc = np.concatenate(cands)
    keys = _keys(allc, out_shape)
    _, first = np.unique(keys, return_index=True)
    out = allc[first]
    return out.astype(np.int32), out_shape


def strided_pairs(in_coords: np.ndarray, in_shape: Sequence[int], out_coords: np.ndarray,
                  ksize: Sequence[int] = (3, 3, 3), stride: int = 2,
                  padding: int = 1) -> Tuple[np.ndarray, np.ndarray]:
    """Regular (strided) conv pair tables.

    ``pair_fwd[f][o]`` = input row at ``o*stride - padding + k_f`` (or -1), shape
    ``[kv, N_out]``; ``pair_bwd[f][i]`` = the output row ``o`` with
    ``pair_fwd[f][o] == i`` (or -1), shape ``[kv, N_in]``.
    """
    offs = kernel_offsets(ksize)
    keys = _keys(in_coords, in_shape)
    order = np.argsort(keys, kind="stable")
    sk = keys[order]
    shp = np.asarray(in_shape, dtype=np.int64)
    n_out = out_coords.shape[0]
    n_in = in_coords.shape[0]
    pf = np.full((len(offs), n_out), -1, dtype=np.int32)
    pb = np.full((len(offs), n_in), -1, dtype=np.int32)
    oc = out_coords.astype(np.int64)
    for f, o in enumerate(offs):
        nb = oc.copy()
        nb[:, 1:] = oc[:, 1:] * stride - padding + o
        valid = np.all((nb[:, 1:] >= 0) & (nb[:, 1:] < shp), axis=1)
        pf[f] = _lookup(sk, order, _keys(nb, in_shape), valid)
        hit = pf[f] >= 0
        pb[f, pf[f][hit]] = np.nonzero(hit)[0].astype(np.int32)
    return pf, pb


# --------------------------------------------------------------------------
# masks and argsort
# --------------------------------------------------------------------------
def row_masks(pair: np.ndarray) -> np.ndarray:
    """``mask[i]`` with bit ``f`` set iff ``pair[f][i] != -1``: uint32 [N] for
    kv <= 32, else the dynamic-mask layout [N, div_up(kv, 32)] with offset f in
    bit f % 32 of word f // 32 (cumm/gemm/blockmma/mma.py:168-199)."""
    kv = pair.shape[0]
    mw = (kv + 31) // 32
    m = np.zeros((pair.shape[1], mw), dtype=np.uint64)
    for f in range(kv):
        m[:, f // 32] |= (pair[f] >= 0).astype(np.uint64) << np.uint64(f % 32)
    m = m.astype(np.uint32)
    return m[:, 0].copy() if kv <= 32 else m


def sort_masks(mask: np.ndarray, split_mask: int = 0xFFFFFFFF) -> Tuple[np.ndarray, np.ndarray]:
    """Stable sort of rows by ``mask & split_mask``.

    Returns ``(mask_sorted, argsort)`` where ``mask_sorted[r] = mask[argsort[r]] &
    split_mask``: the in-place key sort the caller performs before
    ``implicit_gemm2`` (cumm/conv/main_real_spconv.py:141-143; with a split mask,
    each launch only sees its own offsets, main_real_spconv.py:127-143).
    """
    key = (mask.astype(np.uint64) & np.uint64(split_mask)).astype(np.uint32)
    if key.ndim == 1:
        order = np.argsort(key, kind="stable").astype(np.int32)
    else:
        # multi-word masks: stable sort by the words as one number, last word first
        order = np.lexsort([key[:, w] for w in range(key.shape[1])]).astype(np.int32)
    return key[order], order


@dataclasses.dataclass
class SubmProblem:
    """Everything a submanifold layer needs on the host."""
    coords: np.ndarray          # [N, 4] int32
    spatial_shape: Tuple[int, int, int]
    pair_fwd: np.ndarray        # [kv, N] int32
    pair_bwd: np.ndarray        # [kv, N] int32
    mask: np.ndarray            # [N] uint32 (original order); [N, W] for kv > 32
    mask_sorted: np.ndarray     # [N] uint32 (sorted order); [N, W] for kv > 32
    argsort: np.ndarray         # [N] int32

    @property
    def n(self) -> int:
        return self.coords.shape[0]

    @property
    def kv(self) -> int:
        return self.pair_fwd.shape[0]

    @property
    def num_pairs(self) -> int:
        return int(np.count_nonzero(self.pair_fwd >= 0))


def build_subm(coords: np.ndarray, spatial_shape: Sequence[int],
               ksize: Sequence[int] = (3, 3, 3), dilation=1) -> SubmProblem:
    pf = subm_pairs(coords, spatial_shape, ksize, dilation)
    kv = pf.shape[0]
    # subm symmetry: coord_j = coord_i + d_f  <=>  coord_i = coord_j + d_{kv-1-f}
    pb = pf[::-1].copy()
    m = row_masks(pf)
    ms, order = sort_masks(m)
    return SubmProblem(coords, tuple(int(s) for s in spatial_shape), pf, pb, m, ms, order)


@dataclasses.dataclass
class StridedProblem:
    in_coords: np.ndarray
    out_coords: np.ndarray
    in_shape: Tuple[int, int, int]
    out_shape: Tuple[int, int, int]
    pair_fwd: np.ndarray        # [kv, N_out]
    pair_bwd: np.ndarray        # [kv, N_in]
    mask_fwd_sorted: np.ndarray
    argsort_fwd: np.ndarray
    mask_bwd_sorted: np.ndarray
    argsort_bwd: np.ndarray

    @property
    def num_pairs(self) -> int:
        return int(np.count_nonzero(self.pair_fwd >= 0))


def build_strided(in_coords: np.ndarray, in_shape: Sequence[int],
                  ksize: Sequence[int] = (3, 3, 3), stride: int = 2,
                  padding: int = 1, rule: str = "reach") -> StridedProblem:
    out_coords, out_shape = strided_out_coords(in_coords, in_shape, ksize, stride, padding, rule)
    pf, pb = strided_pairs(in_coords, in_shape, out_coords, ksize, stride, padding)
    mf, af = sort_masks(row_masks(pf))
    mb, ab = sort_masks(row_masks(pb))
    return StridedProblem(in_coords, out_coords, tuple(int(s) for s in in_shape), out_shape,
                          pf, pb, mf, af, mb, ab)


def tile_masks(mask_sorted: np.ndarray, tile_m: int, mask_filter: int = 0xFFFFFFFF) -> np.ndarray:
    """Per-tile OR of sorted row masks, ANDed with ``mask_filter``: what the
    forward kernel stores in ``mask_output`` (cumm/conv/kernel.py:843-874)."""
    n = mask_sorted.shape[0]
    t = (n + tile_m - 1) // tile_m
    words = mask_sorted.reshape(n, -1)
    pad = np.zeros((t * tile_m, words.shape[1]), dtype=np.uint32)
    pad[:n] = words
    red = (np.bitwise_or.reduce(pad.reshape(t, tile_m, -1), axis=1) & np.uint32(mask_filter)).astype(np.uint32)
    return red[:, 0].copy() if mask_sorted.ndim == 1 else red


def mac_efficiency(mask_sorted: np.ndarray, tile_m: int, num_pairs: int) -> float:
    """``P / sum_t tile_m * popcount(tile_mask_t)``: fraction of issued MACs that are real."""
    tm = tile_masks(mask_sorted, tile_m)
    pc = np.array([bin(int(v)).count("1") for v in tm.reshape(-1)], dtype=np.int64)
    issued = int(pc.sum()) * tile_m
    return num_pairs / issued if issued else 1.0
