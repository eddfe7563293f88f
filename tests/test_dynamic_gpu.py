"""Dynamic (multi-word) masks on the GPU: 5x5x5 submanifold kernels (kv = 125,
masks of four 32-bit words per row) and 7x7x7 (kv = 343, eleven words, taken
by the kernels in windows of four words), through implicit_gemm2 with the
``dynamic_mask`` kernels.

* forward, dgrad (forward masks reversed in-kernel) and wgrad (reduced
  multi-word masks from the forward's mask_output) against the FP64 oracle,
  whose multi-word handling is pinned to the coordinate brute force on the CPU;
  stated tolerances as for 3x3x3 (relative error vs the oracle on the same
  rounded inputs: f32 5e-6, f16 2e-3, bf16 1e-2), mask_output bit-exact;
* both forward kernels: streamed weights (reduction a multiple of the MFMA
  chunk, <= 128 columns) and the block kernel (other shapes);
* mask_filter acts on bit f % 32 of every word;
* a 3x3x3 layer through a dynamic-mask desp equals the single-word kernels bit
  for bit (one mask word per row);
* masks too short for N * words are rejected; ConvParams.mask_int_count must
  be its default (1) or the word count; verbose prints the kernel name, its
  register count and the time.
"""
import numpy as np
import pytest
import torch

from oracle import spconv_ref as ref
from spconv_case import TORCH_DT, conv_params, desp_for, quantize, subm_case, to_dev

from cumm_amd import pointcloud as pc
from cumm_amd.conv.main import ConvMainUnitTest

pytestmark = pytest.mark.gpu
TOL = {"f32": 5e-6, "f16": 2e-3, "bf16": 1e-2}
GRID = (16, 96, 96)


def _rel(y, y64):
    return ref.relative_error(y, y64)[1]


def _case5(n, C, K, seed, ks=5):
    coords = pc.plane_scene(n, grid=GRID, seed=seed)
    prob = pc.build_subm(coords, GRID, ksize=(ks, ks, ks))
    rng = np.random.default_rng(seed + 77)
    X = rng.standard_normal((prob.n, C)).astype(np.float32)
    W = (rng.standard_normal((K, prob.kv, C)) / np.sqrt(prob.kv * C)).astype(np.float32)
    dO = rng.standard_normal((prob.n, K)).astype(np.float32)
    return prob, X, W, dO


def _fwd(prob, X, W, dt, mask_filter=0xFFFFFFFF):
    C, K = X.shape[1], W.shape[0]
    desp = desp_for(0, dt, C, K, dynamic=True)
    assert desp.dynamic_mask
    out = torch.zeros((prob.n, K), dtype=TORCH_DT[dt], device="cuda")
    tiles = (prob.n + 15) // 16
    mo = torch.zeros((tiles, prob.mask_sorted.shape[1]), dtype=torch.int32, device="cuda")
    p = conv_params(0, desp, to_dev(X, dt), to_dev(W, dt), out, prob.pair_fwd, prob.mask_sorted, prob.argsort,
                    mask_output=mo, mask_filter=mask_filter)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    return out.double().cpu().numpy(), mo.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("C,K,dt,ks", [(32, 32, "f16", 5), (64, 64, "bf16", 5), (64, 128, "f16", 5),
                                       (16, 32, "f32", 5), (48, 40, "f16", 5), (32, 160, "bf16", 5),
                                       (32, 32, "f16", 7), (64, 64, "bf16", 7), (16, 32, "f32", 7),
                                       (48, 40, "f16", 7)])
def test_dynamic_forward_dgrad_wgrad(C, K, dt, ks):
    print(f"[dyn] C={C} K={K} dt={dt} ksize={ks}", flush=True)
    prob, X, W, dO = _case5(2500, C, K, seed=C + K + ks, ks=ks)
    words = (ks ** 3 + 31) // 32
    assert prob.kv == ks ** 3 and prob.mask_sorted.shape == (prob.n, words)
    xq, wq, dq = quantize(X, dt), quantize(W, dt), quantize(dO, dt)
    # forward
    got, mo = _fwd(prob, X, W, dt)
    exp, tm = ref.conv_forward(xq, wq, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16)
    assert np.array_equal(mo, tm)
    assert _rel(got, quantize(exp, dt)) <= TOL[dt]
    # dgrad: forward masks, reversed in the kernel
    ddesp = desp_for(1, dt, C, K, dynamic=True)
    din = torch.zeros((prob.n, C), dtype=TORCH_DT[dt], device="cuda")
    p = conv_params(1, ddesp, din, to_dev(W, dt), to_dev(dO, dt), prob.pair_bwd, prob.mask_sorted,
                    prob.argsort, reverse_mask=True)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_dgrad(dq, wq, prob.pair_bwd, prob.mask_sorted, prob.argsort, 16, reverse_mask=True)
    assert _rel(din.double().cpu().numpy(), quantize(exp, dt)) <= TOL[dt]
    # wgrad over the forward's multi-word mask_output
    wdesp = desp_for(2, dt, C, K, dynamic=True)
    dw = torch.zeros((K, prob.kv, C), dtype=TORCH_DT[dt], device="cuda")
    p = conv_params(2, wdesp, to_dev(X, dt), dw, to_dev(dO, dt), prob.pair_fwd, mo, prob.argsort, mask_width=16,
                    split_k=8)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_wgrad(xq, dq, prob.pair_fwd, mo, prob.argsort, 16)
    assert _rel(dw.double().cpu().numpy(), quantize(exp, dt)) <= 2 * TOL[dt]


@pytest.mark.parametrize("dt", ["f16", "f32"])
def test_two_word_masks_forward_wgrad(dt):
    """5x3x3 kernel (kv = 45: two mask words, at most 64 offsets): the forward's
    two-word tile masks, then wgrad, whose partial products run on the grid
    grouped per XCD (kernels of up to 64 offsets), against the FP64 oracle."""
    print(f"[dyn] 5x3x3 dt={dt}", flush=True)
    coords = pc.plane_scene(3000, grid=GRID, seed=11)
    prob = pc.build_subm(coords, GRID, ksize=(5, 3, 3))
    assert prob.kv == 45 and prob.mask_sorted.shape == (prob.n, 2)
    C, K = 64, 48
    rng = np.random.default_rng(45)
    X = rng.standard_normal((prob.n, C)).astype(np.float32)
    W = (rng.standard_normal((K, prob.kv, C)) / np.sqrt(prob.kv * C)).astype(np.float32)
    dO = rng.standard_normal((prob.n, K)).astype(np.float32)
    xq, wq, dq = quantize(X, dt), quantize(W, dt), quantize(dO, dt)
    got, mo = _fwd(prob, X, W, dt)
    exp, tm = ref.conv_forward(xq, wq, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16)
    assert np.array_equal(mo, tm)
    assert _rel(got, quantize(exp, dt)) <= TOL[dt]
    wdesp = desp_for(2, dt, C, K, dynamic=True)
    dw = torch.zeros((K, prob.kv, C), dtype=TORCH_DT[dt], device="cuda")
    p = conv_params(2, wdesp, to_dev(X, dt), dw, to_dev(dO, dt), prob.pair_fwd, mo, prob.argsort, mask_width=16)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_wgrad(xq, dq, prob.pair_fwd, mo, prob.argsort, 16)
    assert _rel(dw.double().cpu().numpy(), quantize(exp, dt)) <= 2 * TOL[dt]


def test_dynamic_mask_filter_every_word():
    print("[dyn] mask filter", flush=True)
    prob, X, W, dO = _case5(2000, 32, 32, seed=5)
    filt = 0xFFFFFFFF & ~(1 << 3) & ~(1 << 30)
    got, mo = _fwd(prob, X, W, "f16", mask_filter=filt)
    exp, tm = ref.conv_forward(quantize(X, "f16"), quantize(W, "f16"), prob.pair_fwd, prob.mask_sorted,
                               prob.argsort, 16, mask_filter=filt)
    assert np.array_equal(mo, tm) and not (mo & np.uint32((1 << 3) | (1 << 30))).any()
    assert _rel(got, quantize(exp, "f16")) <= TOL["f16"]


def test_dynamic_desp_on_small_kernel_is_bit_identical():
    print("[dyn] 3x3x3 through dynamic desp", flush=True)
    c = subm_case(3000, 64, 64, seed=8)
    prob = c.prob
    outs = []
    for dyn in (False, True):
        desp = desp_for(0, "f16", 64, 64, dynamic=dyn)
        out = torch.zeros((prob.n, 64), dtype=torch.float16, device="cuda")
        p = conv_params(0, desp, to_dev(c.X, "f16"), to_dev(c.W, "f16"), out, prob.pair_fwd, prob.mask_sorted,
                        prob.argsort)
        ConvMainUnitTest.implicit_gemm2(p)
        torch.cuda.synchronize()
        outs.append(out.clone())
    assert torch.equal(outs[0], outs[1])


def test_dynamic_masks_validation():
    prob, X, W, _ = _case5(600, 16, 16, seed=2)
    desp = desp_for(0, "f16", 16, 16, dynamic=True)
    out = torch.zeros((prob.n, 16), dtype=torch.float16, device="cuda")
    short = prob.mask_sorted[:, 0].copy()  # one word per row: too short for kv = 125
    p = conv_params(0, desp, to_dev(X, "f16"), to_dev(W, "f16"), out, prob.pair_fwd, short, prob.argsort)
    with pytest.raises(RuntimeError, match="mask must hold"):
        ConvMainUnitTest.implicit_gemm2(p)
    static = desp_for(0, "f16", 16, 16, dynamic=False)
    p = conv_params(0, static, to_dev(X, "f16"), to_dev(W, "f16"), out, prob.pair_fwd, prob.mask_sorted,
                    prob.argsort)
    with pytest.raises(ValueError, match="dynamic"):
        ConvMainUnitTest.implicit_gemm2(p)


def test_dynamic_mask_int_count_and_verbose(capfd):
    prob, X, W, _ = _case5(800, 16, 16, seed=4, ks=7)
    desp = desp_for(0, "f16", 16, 16, dynamic=True)
    out = torch.zeros((prob.n, 16), dtype=torch.float16, device="cuda")

    def params(mic, verbose=False):
        p = conv_params(0, desp, to_dev(X, "f16"), to_dev(W, "f16"), out, prob.pair_fwd, prob.mask_sorted,
                        prob.argsort)
        p.mask_int_count = mic
        p.verbose = verbose
        return p

    ConvMainUnitTest.implicit_gemm2(params(1))   # the reference's default
    ConvMainUnitTest.implicit_gemm2(params(11))  # div_up(343, 32)
    torch.cuda.synchronize()
    with pytest.raises(ValueError, match="mask_int_count"):
        ConvMainUnitTest.implicit_gemm2(params(4))
    capfd.readouterr()
    ConvMainUnitTest.implicit_gemm2(params(11, verbose=True))
    printed = capfd.readouterr().out
    assert repr(desp) in printed and "kernel num regs:" in printed and "time:" in printed
