"""GPU parity of the edge cases the reference's own harness runs and of every
epilogue variant, against the FP64 oracle, through the C ABI.

* The split-mask two-launch pattern for dgrad and wgrad
  (cumm/conv/main_real_spconv.py:266-290): dgrad with ``reverse_mask`` over the
  forward's split masks, beta = 1 on the second launch; wgrad with
  ``mask_filter`` halves 0b1..1 / 0b1..1 << 14 over each split's
  ``mask_output``, beta = 1 on the second launch.
* ``mask_filter`` other than all ones on every single-word kernel form
  (weight-stationary with 16 / 32 / 64 columns and 1 / 2 / 4 reduction chunks,
  streamed weights, the block kernel): forward (mask_output bit-exact), dgrad
  (reverse_mask) and wgrad (filtered offsets keep their initial values).
* Epilogues ``act(alpha * acc + beta * src)`` with the output or a bias as
  source (cumm/gemm/output_op/linear.py:108-195,
  cumm/gemm/gemmmath/__init__.py:152-222) for 16-bit forward (the path that
  does not stage whole rows), dgrad and wgrad.

Tolerances are relative Frobenius errors against the FP64 oracle on the
quantised inputs, with every intermediate 16-bit result quantised the same
way: f32 5e-6, f16 1e-3 (2e-3 for wgrad), bf16 8e-3.
"""
import numpy as np
import pytest
import torch

from oracle import spconv_ref as ref
from spconv_case import ALL, DT, TORCH_DT, conv_params, desp_for, quantize, subm_case, to_dev

from cumm_amd import pointcloud as pc
from cumm_amd.conv.main import ConvMainUnitTest

pytestmark = pytest.mark.gpu

TOL = {"f32": 5e-6, "f16": 1e-3, "bf16": 8e-3}
LO = 0b11111111111111          # offsets 0..13
HI = 0b1111111111111 << 14     # offsets 14..26


def _rel(y, y64):
    return ref.relative_error(y, y64)[1]


def _desp(op, dt, C, K, bn, out_dt=None):
    """Catalogue entry for (op, dtypes) whose tile_shape[1] is bn."""
    o = DT[out_dt or dt]
    for d in ConvMainUnitTest.get_all_conv_algo_desp():
        if int(d.op_type) != op or d.dynamic_mask or d.is_int8_inference or d.tile_shape[1] != bn:
            continue
        io = (d.dtype_input, d.dtype_weight, d.dtype_output)
        want = {0: (DT[dt], DT[dt], o), 1: (o, DT[dt], DT[dt])}[op]
        if io == want:
            return d
    raise LookupError(f"no desp op={op} dt={dt} bn={bn}")


@pytest.fixture(scope="module")
def case():
    return subm_case(3000, 64, 64, seed=31)


@pytest.mark.parametrize("dt", ["f32", "f16"])
def test_dgrad_split_masks_two_launches(case, dt):
    c = case
    prob = c.prob
    print(f"[dgrad-split] dt={dt}", flush=True)
    rng = np.random.default_rng(41)
    init = quantize(rng.standard_normal((prob.n, c.C)), dt)
    din = to_dev(init, dt)
    desp = desp_for(1, dt, c.C, c.K)
    exp = init
    dq, wq = quantize(c.dO, dt), quantize(c.W, dt)
    for j, split in enumerate([LO, HI]):
        ms, order = pc.sort_masks(prob.mask, split)
        p = conv_params(1, desp, din, to_dev(c.W, dt), to_dev(c.dO, dt), prob.pair_bwd, ms, order,
                        reverse_mask=True, beta=float(j))
        ConvMainUnitTest.implicit_gemm2(p)
        torch.cuda.synchronize()
        exp = quantize(ref.conv_dgrad(dq, wq, prob.pair_bwd, ms, order, 16, reverse_mask=True, beta=float(j),
                                      din_init=exp), dt)
    assert _rel(din.double().cpu().numpy(), exp) <= TOL[dt]
    # the two halves together are the whole dgrad (rows of empty tiles keep init)
    full = ref.conv_dgrad(dq, wq, prob.pair_bwd, prob.mask_sorted, prob.argsort, 16, reverse_mask=True)
    assert _rel(exp, quantize(full, dt)) <= 4 * TOL[dt]


@pytest.mark.parametrize("dt,out_dt", [("f32", "f32"), ("f16", "f16"), ("f16", "f32")])
def test_wgrad_split_filters_two_launches(case, dt, out_dt):
    c = case
    prob = c.prob
    print(f"[wgrad-split] dt={dt} out={out_dt}", flush=True)
    rng = np.random.default_rng(43)
    init = quantize(rng.standard_normal((c.K, prob.kv, c.C)), out_dt)
    dw = to_dev(init, out_dt)
    desp = desp_for(2, dt, c.C, c.K, out_dt)
    xq, dq = quantize(c.X, dt), quantize(c.dO, dt)
    exp = init
    for j, filt in enumerate([LO, HI]):
        ms, order = pc.sort_masks(prob.mask, filt)
        mo = ref.tile_masks(ms, prob.n, 16).astype(np.uint32)  # the split forward's mask_output
        p = conv_params(2, desp, to_dev(c.X, dt), dw, to_dev(c.dO, dt), prob.pair_fwd, mo, order, mask_width=16,
                        mask_filter=filt, beta=float(j))
        ConvMainUnitTest.implicit_gemm2(p)
        torch.cuda.synchronize()
        exp = quantize(ref.conv_wgrad(xq, dq, prob.pair_fwd, mo, order, 16, mask_filter=filt, beta=float(j),
                                      dW_init=exp), out_dt)
    got = dw.double().cpu().numpy()
    assert _rel(got, exp) <= 2 * max(TOL[dt], TOL[out_dt])
    # offsets 0..13 come from the first launch alone: init is gone there
    full = ref.conv_wgrad(xq, dq, prob.pair_fwd, ref.tile_masks(prob.mask_sorted, prob.n, 16).astype(np.uint32),
                          prob.argsort, 16)
    assert _rel(got[:, :14], quantize(full[:, :14], out_dt)) <= 2 * max(TOL[dt], TOL[out_dt])


# (dt, C, K, bn): the kernel each reaches for forward (C -> K) and dgrad (K -> C)
FILTER_CASES = [
    ("f16", 32, 64, 64),    # weight-stationary, one reduction chunk
    ("f16", 64, 64, 16),    # 16-column slices
    ("f16", 64, 64, 32),    # 32-column slices
    ("f16", 64, 64, 64),    # two 32-column slices
    ("f16", 128, 64, 32),   # four reduction chunks, 16-column slices (forward)
    ("f16", 128, 64, 64),   # streamed weights, four reduction chunks (forward)
    ("bf16", 64, 128, 128),  # streamed weights
    ("f32", 64, 64, 64),    # FP32 weight-stationary
    ("f32", 128, 64, 64),   # FP32 block kernel (forward: 8 reduction chunks)
]


@pytest.mark.parametrize("filt", [0x5A5A5A5, LO])
@pytest.mark.parametrize("dt,C,K,bn", FILTER_CASES)
def test_mask_filter_every_kernel(dt, C, K, bn, filt):
    c = subm_case(2000, C, K, seed=C + K + bn)
    prob = c.prob
    print(f"[filter] dt={dt} C={C} K={K} bn={bn} filter={filt:#x}", flush=True)
    xq, wq, dq = quantize(c.X, dt), quantize(c.W, dt), quantize(c.dO, dt)
    tiles = (prob.n + 15) // 16
    # forward
    out = torch.zeros((prob.n, K), dtype=TORCH_DT[dt], device="cuda")
    mo = torch.zeros(tiles, dtype=torch.int32, device="cuda")
    p = conv_params(0, _desp(0, dt, C, K, bn), to_dev(c.X, dt), to_dev(c.W, dt), out, prob.pair_fwd,
                    prob.mask_sorted, prob.argsort, mask_output=mo, mask_filter=filt)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp, tm = ref.conv_forward(xq, wq, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16, mask_filter=filt)
    assert np.array_equal(mo.cpu().numpy().view(np.uint32), tm), "mask_output must be bit-exact"
    assert _rel(out.double().cpu().numpy(), quantize(exp, dt)) <= TOL[dt]
    # dgrad over the reversed forward masks: the filter applies before the reverse
    rng = np.random.default_rng(C * K)
    init = quantize(rng.standard_normal((prob.n, C)), dt)
    din = to_dev(init, dt)
    p = conv_params(1, _desp(1, dt, C, K, bn), din, to_dev(c.W, dt), to_dev(c.dO, dt), prob.pair_bwd,
                    prob.mask_sorted, prob.argsort, reverse_mask=True, mask_filter=filt)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_dgrad(dq, wq, prob.pair_bwd, prob.mask_sorted, prob.argsort, 16, mask_filter=filt,
                         reverse_mask=True, din_init=init)
    assert _rel(din.double().cpu().numpy(), quantize(exp, dt)) <= TOL[dt]
    # wgrad: filtered-out offsets keep their initial values bit for bit
    winit = quantize(rng.standard_normal((K, prob.kv, C)), dt)
    dw = to_dev(winit, dt)
    p = conv_params(2, desp_for(2, dt, C, K), to_dev(c.X, dt), dw, to_dev(c.dO, dt), prob.pair_fwd, tm,
                    prob.argsort, mask_width=16, mask_filter=filt)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    got = dw.double().cpu().numpy()
    exp = ref.conv_wgrad(xq, dq, prob.pair_fwd, tm, prob.argsort, 16, mask_filter=filt, dW_init=winit)
    off = np.array([(filt >> (f % 32)) & 1 for f in range(prob.kv)], dtype=bool)
    assert np.array_equal(got[:, ~off], winit[:, ~off]), "filtered offsets must stay untouched"
    assert _rel(got, quantize(exp, dt)) <= 2 * TOL[dt]


EPI = [(0.75, 1.0, False, 0, 0.0), (0.5, 0.5, True, 1, 0.0), (1.25, 1.0, False, 3, 0.1), (1.0, 2.0, True, 2, 0.0)]


@pytest.mark.parametrize("bn", [16, 64, 128])
@pytest.mark.parametrize("dt", ["f16", "bf16"])
def test_forward_16bit_epilogue_with_source(case, dt, bn):
    """beta != 0 makes the 16-bit forward read its source (the output rows or
    the bias) and write element by element instead of staging whole rows."""
    c = case
    prob = c.prob
    rng = np.random.default_rng(bn)
    init = quantize(rng.standard_normal((prob.n, c.K)), dt)
    bias = quantize(rng.standard_normal(c.K), dt)
    xq, wq = quantize(c.X, dt), quantize(c.W, dt)
    for alpha, beta, use_bias, act, aa in EPI:
        print(f"[fwd-epi16] dt={dt} bn={bn} alpha={alpha} beta={beta} bias={use_bias} act={act}", flush=True)
        out = to_dev(init, dt)
        p = conv_params(0, _desp(0, dt, c.C, c.K, bn), to_dev(c.X, dt), to_dev(c.W, dt), out, prob.pair_fwd,
                        prob.mask_sorted, prob.argsort, alpha=alpha, beta=beta,
                        bias=to_dev(bias, dt) if use_bias else None, act_type=act, act_alpha=aa)
        ConvMainUnitTest.implicit_gemm2(p)
        torch.cuda.synchronize()
        exp, _ = ref.conv_forward(xq, wq, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16, alpha=alpha, beta=beta,
                                  bias=bias if use_bias else None, out_init=init, act_type=act, act_alpha=aa)
        assert _rel(out.double().cpu().numpy(), quantize(exp, dt)) <= TOL[dt]


@pytest.mark.parametrize("bn", [64, 128])
@pytest.mark.parametrize("dt", ["f32", "f16"])
def test_dgrad_epilogue(case, dt, bn):
    c = case
    prob = c.prob
    rng = np.random.default_rng(7 + bn)
    init = quantize(rng.standard_normal((prob.n, c.C)), dt)
    bias = quantize(rng.standard_normal(c.C), dt)
    dq, wq = quantize(c.dO, dt), quantize(c.W, dt)
    for alpha, beta, use_bias, act, aa in EPI:
        print(f"[dgrad-epi] dt={dt} bn={bn} alpha={alpha} beta={beta} bias={use_bias} act={act}", flush=True)
        din = to_dev(init, dt)
        p = conv_params(1, _desp(1, dt, c.C, c.K, bn), din, to_dev(c.W, dt), to_dev(c.dO, dt), prob.pair_bwd,
                        prob.mask_sorted, prob.argsort, reverse_mask=True, alpha=alpha, beta=beta,
                        bias=to_dev(bias, dt) if use_bias else None, act_type=act, act_alpha=aa)
        ConvMainUnitTest.implicit_gemm2(p)
        torch.cuda.synchronize()
        exp = ref.conv_dgrad(dq, wq, prob.pair_bwd, prob.mask_sorted, prob.argsort, 16, reverse_mask=True,
                             alpha=alpha, beta=beta, bias=bias if use_bias else None, din_init=init, act_type=act,
                             act_alpha=aa)
        assert _rel(din.double().cpu().numpy(), quantize(exp, dt)) <= TOL[dt]


@pytest.mark.parametrize("dt,out_dt", [("f32", "f32"), ("f16", "f16"), ("bf16", "f32")])
def test_wgrad_epilogue(case, dt, out_dt):
    """alpha / beta / bias / activation in the split-K reduction's epilogue;
    the bias has one value per column of the [K, kv * C] gradient."""
    c = case
    prob = c.prob
    tm = ref.tile_masks(prob.mask_sorted, prob.n, 16).astype(np.uint32)
    rng = np.random.default_rng(11)
    init = quantize(rng.standard_normal((c.K, prob.kv, c.C)), out_dt)
    bias = quantize(rng.standard_normal(prob.kv * c.C), out_dt)
    xq, dq = quantize(c.X, dt), quantize(c.dO, dt)
    for alpha, beta, use_bias, act, aa in EPI:
        print(f"[wgrad-epi] dt={dt} out={out_dt} alpha={alpha} beta={beta} bias={use_bias} act={act}", flush=True)
        dw = to_dev(init, out_dt)
        p = conv_params(2, desp_for(2, dt, c.C, c.K, out_dt), to_dev(c.X, dt), dw, to_dev(c.dO, dt),
                        prob.pair_fwd, tm, prob.argsort, mask_width=16, alpha=alpha, beta=beta,
                        bias=to_dev(bias, out_dt) if use_bias else None, act_type=act, act_alpha=aa)
        ConvMainUnitTest.implicit_gemm2(p)
        torch.cuda.synchronize()
        exp = ref.conv_wgrad(xq, dq, prob.pair_fwd, tm, prob.argsort, 16, alpha=alpha, beta=beta, dW_init=init,
                             act_type=act, act_alpha=aa, bias=bias if use_bias else None)
        assert _rel(dw.double().cpu().numpy(), quantize(exp, out_dt)) <= 2 * max(TOL[dt], TOL[out_dt])


@pytest.mark.parametrize("K", [64, 128])
@pytest.mark.parametrize("dt", ["f16", "bf16"])
def test_streamed_forward_staged_epilogue(dt, K):
    """beta = 0 with alpha and an activation on the streamed-weights kernel
    (tile_shape[1] = 128): output rows leave through its LDS stage as whole
    16-byte segments (64- and 128-column blocks), tile groups run from the
    last one back; mask_output stays bit-exact."""
    c = subm_case(3000, 64, K, seed=K + 7)
    prob = c.prob
    xq, wq = quantize(c.X, dt), quantize(c.W, dt)
    tiles = (prob.n + 15) // 16
    for alpha, act, aa in ((0.5, 1, 0.0), (1.25, 3, 0.1), (1.0, 2, 0.0)):
        print(f"[stream-epi] dt={dt} K={K} alpha={alpha} act={act}", flush=True)
        out = torch.full((prob.n, K), 7.0, dtype=TORCH_DT[dt], device="cuda")
        mo = torch.zeros(tiles, dtype=torch.int32, device="cuda")
        p = conv_params(0, _desp(0, dt, 64, K, 128), to_dev(c.X, dt), to_dev(c.W, dt), out, prob.pair_fwd,
                        prob.mask_sorted, prob.argsort, mask_output=mo, alpha=alpha, beta=0.0, act_type=act,
                        act_alpha=aa)
        ConvMainUnitTest.implicit_gemm2(p)
        torch.cuda.synchronize()
        exp, tm = ref.conv_forward(xq, wq, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16, alpha=alpha,
                                   act_type=act, act_alpha=aa)
        assert np.array_equal(mo.cpu().numpy().view(np.uint32), tm), "mask_output must be bit-exact"
        assert _rel(out.double().cpu().numpy(), quantize(exp, dt)) <= TOL[dt]
