"""GPU results on the golden fixtures (tests/golden/), whose expected values
come from the coordinate-level brute force and numpy FP64 closed forms.

FP32 kernels against FP64 expected values: relative Frobenius error <= 2e-6.
"""
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import spconv_ref as ref
from spconv_case import conv_params, desp_for, to_dev

from cumm_amd import tensorview as tv
from cumm_amd.conv.main import ConvMainUnitTest
from cumm_amd.gemm.main import GemmMainUnitTest
from cumm_amd.tensorview.gemm import GemmParams

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
TOL = 2e-6


def _conv_case(g, pf, pb, mf, af, mb, ab, reverse):
    X, W, dO = g["X"], g["W"], g["dO"]
    n_in, C = X.shape
    n_out, K = dO.shape
    out = torch.zeros((n_out, K), dtype=torch.float32, device="cuda")
    tiles = (n_out + 15) // 16
    mo = torch.zeros(tiles, dtype=torch.int32, device="cuda")
    p = conv_params(0, desp_for(0, "f32", C, K), to_dev(X, "f32"), to_dev(W, "f32"), out, pf, mf, af,
                    mask_output=mo)
    ConvMainUnitTest.implicit_gemm2(p)
    din = torch.zeros((n_in, C), dtype=torch.float32, device="cuda")
    if reverse:
        p = conv_params(1, desp_for(1, "f32", C, K), din, to_dev(W, "f32"), to_dev(dO, "f32"), pb, mf, af,
                        reverse_mask=True)
    else:
        p = conv_params(1, desp_for(1, "f32", C, K), din, to_dev(W, "f32"), to_dev(dO, "f32"), pb, mb, ab)
    ConvMainUnitTest.implicit_gemm2(p)
    dw = torch.zeros((K, 27, C), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    p = conv_params(2, desp_for(2, "f32", C, K), to_dev(X, "f32"), dw, to_dev(dO, "f32"), pf,
                    mo.cpu().numpy().view(np.uint32), af, mask_width=16)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    for name, got in (("out", out), ("din", din), ("dw", dw)):
        rel = ref.relative_error(got.double().cpu().numpy(), g[name])[1]
        print(f"[golden] {name} rel={rel:.2e}", flush=True)
        assert rel <= TOL, name
    return mo.cpu().numpy().view(np.uint32)


def test_golden_subm():
    print("[golden] subm_tiny", flush=True)
    g = np.load(GOLD / "subm_tiny.npz")
    mo = _conv_case(g, g["pair_fwd"], g["pair_bwd"], g["mask_sorted"], g["argsort"], None, None, True)
    assert np.array_equal(mo, g["tile_mask16"])


def test_golden_strided():
    print("[golden] strided_tiny", flush=True)
    g = np.load(GOLD / "strided_tiny.npz")
    _conv_case(g, g["pair_fwd"], g["pair_bwd"], g["mask_fwd_sorted"], g["argsort_fwd"], g["mask_bwd_sorted"],
               g["argsort_bwd"], False)


def _gemm(desp_shuffle, ta, a, b, c, a_inds=None, b_inds=None, c_inds=None):
    for d in GemmMainUnitTest.get_all_algo_desp():
        if (d.dtype_a == tv.float32 and int(d.shuffle_type) == desp_shuffle and d.trans_a == ta
                and not d.trans_b and not d.trans_c):
            desp = d
            break
    p = GemmParams()
    p.algo_desp = desp
    p.a, p.b, p.c = tv.from_torch(a), tv.from_torch(b), tv.from_torch(c)
    for name, t in (("a_inds", a_inds), ("b_inds", b_inds), ("c_inds", c_inds)):
        if t is not None:
            setattr(p, name, tv.from_torch(torch.from_numpy(t).cuda()))
    GemmMainUnitTest.matmul2(p)
    torch.cuda.synchronize()
    return c.double().cpu().numpy()


def test_golden_gemm():
    print("[golden] gemm_small", flush=True)
    g = np.load(GOLD / "gemm_small.npz")
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()  # noqa: E731
    got = _gemm(0, False, dev(g["a"]), dev(g["b"]), torch.zeros((128, 128), device="cuda"))
    assert ref.relative_error(got, g["c"])[1] <= TOL
    m, n = g["ac_c"].shape
    got = _gemm(1, False, dev(g["ac_a"]), dev(g["ac_b"]), torch.zeros((m, n), device="cuda"),
                a_inds=g["ac_a_inds"], c_inds=g["ac_c_inds"])
    assert ref.relative_error(got, g["ac_c"])[1] <= TOL
    got = _gemm(2, True, dev(g["ab_a"]), dev(g["ab_b"]), torch.zeros(g["ab_c"].shape, device="cuda"),
                a_inds=g["ab_a_inds"], b_inds=g["ab_b_inds"])
    assert ref.relative_error(got, g["ab_c"])[1] <= TOL
