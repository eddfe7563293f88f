"""Writes the golden fixtures in this folder (run from the repository root:
``python tests/golden/make_golden.py``).

The reference ships no golden vectors for this path (its pickled fixtures
``test_spconv.pkl`` / ``test_debug.pkl`` are absent), so these are built from
seeded inputs with checkers that do not share code with the kernels or with
the FP64 mask-walk oracle:

* sparse conv expected values come from the coordinate-level brute force
  (oracle/bruteforce.py: a dict from coordinates to rows and plain loops, no
  pair table, no masks, no argsort);
* pair tables, sorted masks, argsort and per-tile masks are the host builder's
  (cumm_amd.pointcloud), stored so later builders (the GPU pair generator)
  are held to them bit-exact;
* GEMM expected values are numpy FP64 products of the closed forms of
  cumm/gemm/main_real.py:197-229 and main_simu.py:643-647 (seed 12315, the
  reference's GEMM test seed, cumm/gemm/main_simu.py:371).

Every array is stored with ``np.savez`` (no pickles).
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "cumm-rocm_amd"))

from oracle import bruteforce  # noqa: E402
from cumm_amd import pointcloud as pc  # noqa: E402

GRID = (6, 14, 14)


def subm_case():
    coords = pc.plane_scene(260, grid=GRID, seed=21)
    prob = pc.build_subm(coords, GRID)
    rng = np.random.default_rng(12315)
    C, K = 8, 16
    X = rng.uniform(-1, 1, (prob.n, C)).astype(np.float32)
    W = rng.uniform(-1, 1, (K, 27, C)).astype(np.float32)
    dO = rng.uniform(-1, 1, (prob.n, K)).astype(np.float32)
    out = bruteforce.conv_forward_coords(coords, X, W)
    din = bruteforce.conv_dgrad_coords(coords, dO, W)
    dw = bruteforce.conv_wgrad_coords(coords, X, dO, 27)
    np.savez(HERE / "subm_tiny.npz", coords=coords, pair_fwd=prob.pair_fwd, pair_bwd=prob.pair_bwd,
             mask=prob.mask, mask_sorted=prob.mask_sorted, argsort=prob.argsort,
             tile_mask16=pc.tile_masks(prob.mask_sorted, 16), X=X, W=W, dO=dO, out=out, din=din, dw=dw)
    print(f"subm_tiny: N={prob.n} P={prob.num_pairs}")


def strided_case():
    coords = pc.plane_scene(260, grid=GRID, seed=22)
    sp = pc.build_strided(coords, GRID)
    rng = np.random.default_rng(12316)
    C, K = 8, 16
    n_in, n_out = coords.shape[0], sp.out_coords.shape[0]
    X = rng.uniform(-1, 1, (n_in, C)).astype(np.float32)
    W = rng.uniform(-1, 1, (K, 27, C)).astype(np.float32)
    dO = rng.uniform(-1, 1, (n_out, K)).astype(np.float32)
    kw = dict(out_coords=sp.out_coords, stride=2, padding=1, subm=False)
    out = bruteforce.conv_forward_coords(coords, X, W, **kw)
    din = bruteforce.conv_dgrad_coords(coords, dO, W, **kw)
    dw = bruteforce.conv_wgrad_coords(coords, X, dO, 27, **kw)
    np.savez(HERE / "strided_tiny.npz", in_coords=coords, out_coords=sp.out_coords,
             pair_fwd=sp.pair_fwd, pair_bwd=sp.pair_bwd, mask_fwd_sorted=sp.mask_fwd_sorted,
             argsort_fwd=sp.argsort_fwd, mask_bwd_sorted=sp.mask_bwd_sorted, argsort_bwd=sp.argsort_bwd,
             X=X, W=W, dO=dO, out=out, din=din, dw=dw)
    print(f"strided_tiny: N_in={n_in} N_out={n_out} P={sp.num_pairs}")


def gemm_case():
    rng = np.random.default_rng(12315)
    a = rng.uniform(-1, 1, (128, 64)).astype(np.float32)
    b = rng.uniform(-1, 1, (64, 128)).astype(np.float32)
    c = a.astype(np.float64) @ b.astype(np.float64)
    m, n, k = 200, 32, 48
    sa = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    sb = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    a_inds = rng.permutation(m).astype(np.int32)
    c_inds = rng.permutation(m).astype(np.int32)
    c_ac = np.zeros((m, n))
    c_ac[c_inds] = sa.astype(np.float64)[a_inds] @ sb.astype(np.float64)
    rows = 160
    ab_a = rng.uniform(-1, 1, (rows, 32)).astype(np.float32)
    ab_b = rng.uniform(-1, 1, (rows, 48)).astype(np.float32)
    ab_ai = rng.permutation(rows).astype(np.int32)
    ab_bi = rng.permutation(rows).astype(np.int32)
    c_ab = ab_a.astype(np.float64)[ab_ai].T @ ab_b.astype(np.float64)[ab_bi]
    np.savez(HERE / "gemm_small.npz", a=a, b=b, c=c, ac_a=sa, ac_b=sb, ac_a_inds=a_inds, ac_c_inds=c_inds,
             ac_c=c_ac, ab_a=ab_a, ab_b=ab_b, ab_a_inds=ab_ai, ab_b_inds=ab_bi, ab_c=c_ab)
    print("gemm_small: config-1 128x128x64, ShuffleAC 200x32x48, ShuffleAB 160 rows -> 32x48")


if __name__ == "__main__":
    subm_case()
    strided_case()
    gemm_case()
