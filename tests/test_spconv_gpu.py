"""GPU parity of ConvMainUnitTest.implicit_gemm2 (mask-sparse forward, dgrad,
wgrad) against the FP64 oracle, called through the C ABI.

Tolerances (relative Frobenius error vs the FP64 oracle evaluated on the
*quantised* inputs, i.e. only accumulation and output rounding count):
f32 2e-6, f16 1e-3, bf16 8e-3, e4m3 inputs with f16 output 1e-3.
Index outputs (mask_output) are compared bit-exact.
"""
import numpy as np
import pytest
import torch

from oracle import spconv_ref as ref
from spconv_case import (ALL, DT, TORCH_DT, conv_params, desp_for, idx_dev, quantize,
                         run_fwd, subm_case, to_dev)

from cumm_amd import pointcloud as pc
from cumm_amd.conv.main import ConvMainUnitTest

pytestmark = pytest.mark.gpu

TOL = {"f32": 2e-6, "f16": 1e-3, "bf16": 8e-3, "e4m3": 1e-3}


def _rel(y, y64):
    return ref.relative_error(y, y64)[1]


@pytest.fixture(scope="module")
def case64():
    return subm_case(3000, 64, 64, seed=3)


@pytest.mark.parametrize("dt", ["f32", "f16", "bf16"])
def test_forward_parity(case64, dt):
    c = case64
    print(f"[fwd] dt={dt} N={c.prob.n} C={c.C} K={c.K}", flush=True)
    out, mo, desp = run_fwd(c, dt)
    exp, tm = ref.conv_forward(quantize(c.X, dt), quantize(c.W, dt), c.prob.pair_fwd, c.prob.mask_sorted,
                               c.prob.argsort, desp.tile_shape[0])
    exp = quantize(exp, dt)  # output rounding of the reference value
    assert np.array_equal(mo, tm), "mask_output must be bit-exact"
    assert _rel(out, exp) <= TOL[dt]


@pytest.mark.parametrize("C,K", [(16, 32), (32, 64), (64, 128), (128, 128), (48, 80)])
def test_forward_channels(C, K):
    c = subm_case(1000 + C, C, K, seed=C + K)
    print(f"[fwd-ch] C={C} K={K}", flush=True)
    out, mo, desp = run_fwd(c, "f16")
    exp, tm = ref.conv_forward(quantize(c.X, "f16"), quantize(c.W, "f16"), c.prob.pair_fwd,
                               c.prob.mask_sorted, c.prob.argsort, desp.tile_shape[0])
    assert np.array_equal(mo, tm)
    assert _rel(out, quantize(exp, "f16")) <= TOL["f16"]


def test_forward_epilogue_bias_beta_act(case64):
    c = case64
    n, K = c.prob.n, c.K
    rng = np.random.default_rng(7)
    init = rng.standard_normal((n, K)).astype(np.float32)
    bias = rng.standard_normal(K).astype(np.float32)
    desp = desp_for(0, "f32", c.C, K)
    for beta, use_bias, act, aa in [(1.0, False, 0, 0.0), (0.5, True, 1, 0.0), (1.0, False, 3, 0.1),
                                    (2.0, True, 2, 0.0)]:
        print(f"[fwd-epi] beta={beta} bias={use_bias} act={act}", flush=True)
        out = torch.from_numpy(init.copy()).cuda()
        bt = torch.from_numpy(bias).cuda() if use_bias else None
        p = conv_params(0, desp, to_dev(c.X, "f32"), to_dev(c.W, "f32"), out, c.prob.pair_fwd,
                        c.prob.mask_sorted, c.prob.argsort, alpha=0.75, beta=beta, bias=bt,
                        act_type=act, act_alpha=aa)
        ConvMainUnitTest.implicit_gemm2(p)
        torch.cuda.synchronize()
        exp, _ = ref.conv_forward(c.X, c.W, c.prob.pair_fwd, c.prob.mask_sorted, c.prob.argsort,
                                  desp.tile_shape[0], alpha=0.75, beta=beta,
                                  bias=bias if use_bias else None, out_init=init, act_type=act,
                                  act_alpha=aa)
        assert _rel(out.double().cpu().numpy(), exp) <= 5e-6


def test_forward_split_masks_two_launches(case64):
    """The reference's split-mask pattern: offsets 0..13 then 14..26, second
    launch with beta=1 (cumm/conv/main_real_spconv.py:127-143, 275-294)."""
    c = case64
    desp = desp_for(0, "f32", c.C, c.K)
    out = torch.zeros((c.prob.n, c.K), dtype=torch.float32, device="cuda")
    exp = None
    tiles = (c.prob.n + 15) // 16
    for j, split in enumerate([0b11111111111111, 0b1111111111111 << 14]):
        ms, order = pc.sort_masks(c.prob.mask, split)
        mo = torch.zeros(tiles, dtype=torch.int32, device="cuda")
        p = conv_params(0, desp, to_dev(c.X, "f32"), to_dev(c.W, "f32"), out, c.prob.pair_fwd, ms, order,
                        mask_output=mo, beta=float(j))
        ConvMainUnitTest.implicit_gemm2(p)
        torch.cuda.synchronize()
        exp, tm = ref.conv_forward(c.X, c.W, c.prob.pair_fwd, ms, order, 16, beta=float(j), out_init=exp)
        assert np.array_equal(mo.cpu().numpy().view(np.uint32), tm)
    full, _ = ref.conv_forward(c.X, c.W, c.prob.pair_fwd, c.prob.mask_sorted, c.prob.argsort, 16)
    assert _rel(out.double().cpu().numpy(), exp) <= 5e-6
    assert _rel(exp, full) <= 1e-12


@pytest.mark.parametrize("dt", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("reverse", [True, False])
def test_dgrad_parity(case64, dt, reverse):
    c = case64
    prob = c.prob
    print(f"[dgrad] dt={dt} reverse={reverse}", flush=True)
    desp = desp_for(1, dt, c.C, c.K)
    if reverse:
        mask, order = prob.mask_sorted, prob.argsort  # forward masks, reversed in-kernel
    else:
        mask, order = pc.sort_masks(pc.row_masks(prob.pair_bwd))
    din = torch.zeros((prob.n, c.C), dtype=TORCH_DT[dt], device="cuda")
    p = conv_params(1, desp, din, to_dev(c.W, dt), to_dev(c.dO, dt), prob.pair_bwd, mask, order,
                    reverse_mask=reverse)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_dgrad(quantize(c.dO, dt), quantize(c.W, dt), prob.pair_bwd, mask, order, 16,
                         reverse_mask=reverse)
    assert _rel(din.double().cpu().numpy(), quantize(exp, dt)) <= TOL[dt]


@pytest.mark.parametrize("dt,out_dt", [("f32", "f32"), ("f16", "f16"), ("bf16", "bf16"), ("f16", "f32")])
@pytest.mark.parametrize("split_k", [1, 7, 32])
def test_wgrad_parity(case64, dt, out_dt, split_k):
    c = case64
    prob = c.prob
    print(f"[wgrad] dt={dt} out={out_dt} split={split_k}", flush=True)
    _, mo, fdesp = run_fwd(c, "f32")
    desp = desp_for(2, dt, c.C, c.K, out_dt)
    dw = torch.zeros((c.K, prob.kv, c.C), dtype=TORCH_DT[out_dt], device="cuda")
    p = conv_params(2, desp, to_dev(c.X, dt), dw, to_dev(c.dO, dt), prob.pair_fwd, mo, prob.argsort,
                    mask_width=fdesp.tile_shape[0], split_k=split_k)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_wgrad(quantize(c.X, dt), quantize(c.dO, dt), prob.pair_fwd, mo, prob.argsort,
                         fdesp.tile_shape[0])
    got = dw.double().cpu().numpy()
    assert _rel(got, quantize(exp, out_dt)) <= max(TOL[dt], TOL[out_dt]) * 2


def test_wgrad_deterministic(case64):
    c = case64
    prob = c.prob
    _, mo, fdesp = run_fwd(c, "f32")
    desp = desp_for(2, "f16", c.C, c.K)
    res = []
    for _ in range(3):
        dw = torch.zeros((c.K, prob.kv, c.C), dtype=torch.float16, device="cuda")
        p = conv_params(2, desp, to_dev(c.X, "f16"), dw, to_dev(c.dO, "f16"), prob.pair_fwd, mo,
                        prob.argsort, mask_width=16, split_k=16)
        ConvMainUnitTest.implicit_gemm2(p)
        torch.cuda.synchronize()
        res.append(dw.cpu().view(torch.int16).numpy())
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])


def test_ragged_and_tiny():
    for n in (1, 15, 17, 63, 65, 257):
        c = subm_case(n, 32, 32, seed=n, grid=(4, 16, 16))
        print(f"[ragged] N={n}", flush=True)
        out, mo, desp = run_fwd(c, "f32")
        exp, tm = ref.conv_forward(c.X, c.W, c.prob.pair_fwd, c.prob.mask_sorted, c.prob.argsort, 16)
        assert np.array_equal(mo, tm)
        assert _rel(out, exp) <= 5e-6


def _e4m3_quant(x, axis_scale):
    """OCP E4M3 quantisation with absmax scales (448 = largest finite e4m3)."""
    q = torch.from_numpy((x / axis_scale).astype(np.float32)).to(torch.float8_e4m3fn)
    return q, q.to(torch.float64).numpy()


@pytest.mark.parametrize("out_dt", ["f16", "bf16"])
def test_forward_fp8_per_channel_scales(case64, out_dt):
    """FP8-E4M3 features (per-tensor scale) and weights (per-output-channel
    absmax scales), FP32 accumulation, scale[k] = s_x * s_w[k] applied in the
    epilogue.  Parity vs the oracle on the dequantised operands; the error
    against the FP64 result on the unquantised inputs is reported, not gated."""
    c = case64
    prob = c.prob
    s_x = np.abs(c.X).max() / 448.0
    s_w = np.abs(c.W).reshape(c.K, -1).max(axis=1) / 448.0
    xq, xq64 = _e4m3_quant(c.X, s_x)
    wq, wq64 = _e4m3_quant(c.W, s_w[:, None, None])
    scale = (s_x * s_w).astype(np.float32)
    desp = desp_for(0, "e4m3", c.C, c.K, out_dt)
    out = torch.zeros((prob.n, c.K), dtype=TORCH_DT[out_dt], device="cuda")
    p = conv_params(0, desp, xq.cuda(), wq.cuda(), out, prob.pair_fwd, prob.mask_sorted, prob.argsort,
                    scale=torch.from_numpy(scale).cuda())
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    got = out.double().cpu().numpy()
    exp, _ = ref.conv_forward(xq64, wq64, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16, scale=scale)
    assert _rel(got, quantize(exp, out_dt)) <= TOL[out_dt]
    full, _ = ref.conv_forward(c.X, c.W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16)
    mx, rel = ref.relative_error(got, full)
    print(f"[fp8] out={out_dt} error vs FP64 on unquantised inputs: rel={rel:.3e} max_abs={mx:.3e}")
    assert rel < 0.1


def test_dgrad_fp8_per_channel_scales(case64):
    """FP8 dgrad: dout per-tensor scale, weights quantised per input channel c
    (the dgrad output channel), scale[c] = s_do * s_w[c] in the epilogue."""
    c = case64
    prob = c.prob
    s_do = np.abs(c.dO).max() / 448.0
    s_w = np.abs(c.W).max(axis=(0, 1)) / 448.0
    dq, dq64 = _e4m3_quant(c.dO, s_do)
    wq, wq64 = _e4m3_quant(c.W, s_w[None, None, :])
    scale = (s_do * s_w).astype(np.float32)
    desp = desp_for(1, "e4m3", c.C, c.K, "f16")
    din = torch.zeros((prob.n, c.C), dtype=torch.float16, device="cuda")
    p = conv_params(1, desp, din, wq.cuda(), dq.cuda(), prob.pair_bwd, prob.mask_sorted, prob.argsort,
                    reverse_mask=True, scale=torch.from_numpy(scale).cuda())
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_dgrad(dq64, wq64, prob.pair_bwd, prob.mask_sorted, prob.argsort, 16, reverse_mask=True)
    exp = exp * scale[None, :]
    assert _rel(din.double().cpu().numpy(), quantize(exp, "f16")) <= TOL["f16"]


def test_wgrad_fp8_deterministic():
    """FP8 wgrad: dout and features per-tensor scaled, alpha = s_do * s_x, FP32
    dW from the fixed-order tree; two runs bit-identical.  Bar 1e-4 relative
    against FP64 on the dequantised operands: the FP8 MFMA sums its 32
    products before the FP32 accumulate (measured 1.3e-5)."""
    c = subm_case(3000, 64, 64, seed=3)
    prob = c.prob
    _, mo, fdesp = run_fwd(c, "f32")
    s_x = np.abs(c.X).max() / 448.0
    s_do = np.abs(c.dO).max() / 448.0
    xq, xq64 = _e4m3_quant(c.X, s_x)
    dq, dq64 = _e4m3_quant(c.dO, s_do)
    desp = desp_for(2, "e4m3", c.C, c.K, "f32")
    res = []
    for _ in range(2):
        dw = torch.zeros((c.K, prob.kv, c.C), dtype=torch.float32, device="cuda")
        p = conv_params(2, desp, xq.cuda(), dw, dq.cuda(), prob.pair_fwd, mo, prob.argsort, mask_width=16,
                        alpha=float(s_do * s_x))
        ConvMainUnitTest.implicit_gemm2(p)
        torch.cuda.synchronize()
        res.append(dw.cpu().numpy())
    assert np.array_equal(res[0].view(np.int32), res[1].view(np.int32))
    exp = ref.conv_wgrad(xq64, dq64, prob.pair_fwd, mo, prob.argsort, 16) * (s_do * s_x)
    assert _rel(res[0].astype(np.float64), exp) <= 1e-4


@pytest.mark.parametrize("C,K", [(32, 160), (16, 192)])
def test_dgrad_block_kernel_path(C, K):
    """dgrad whose reduction (K) is more than four f16 MFMA chunks runs the
    block kernel, which reads the weights through a transposed copy."""
    c = subm_case(1500, C, K, seed=C * K)
    prob = c.prob
    print(f"[dgrad-block] C={C} K={K}", flush=True)
    desp = desp_for(1, "f16", C, K)
    din = torch.zeros((prob.n, C), dtype=torch.float16, device="cuda")
    p = conv_params(1, desp, din, to_dev(c.W, "f16"), to_dev(c.dO, "f16"), prob.pair_bwd, prob.mask_sorted,
                    prob.argsort, reverse_mask=True)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_dgrad(quantize(c.dO, "f16"), quantize(c.W, "f16"), prob.pair_bwd, prob.mask_sorted,
                         prob.argsort, 16, reverse_mask=True)
    assert _rel(din.double().cpu().numpy(), quantize(exp, "f16")) <= TOL["f16"]


@pytest.mark.parametrize("dt,C,K", [("f32", 4, 16), ("f32", 12, 32), ("f16", 8, 32), ("bf16", 24, 64),
                                    ("f16", 16, 16)])
def test_padded_reduction_forward_dgrad(dt, C, K):
    """Reductions of whole 16-byte rows that are not whole MFMA chunks (4 or
    12 f32 channels, 8 / 16 / 24 16-bit channels) run the weight-stationary
    kernel with the reduction padded by loads that read zero: forward over C
    and dgrad over K, against the FP64 oracle."""
    c = subm_case(1500, C, K, seed=7 * C + K)
    prob = c.prob
    print(f"[padded] dt={dt} C={C} K={K}", flush=True)
    out, mo, desp = run_fwd(c, dt)
    exp, tm = ref.conv_forward(quantize(c.X, dt), quantize(c.W, dt), prob.pair_fwd, prob.mask_sorted,
                               prob.argsort, desp.tile_shape[0])
    assert np.array_equal(mo, tm)
    assert _rel(out, quantize(exp, dt)) <= TOL[dt]
    ddesp = desp_for(1, dt, C, K)
    din = torch.zeros((prob.n, C), dtype=TORCH_DT[dt], device="cuda")
    p = conv_params(1, ddesp, din, to_dev(c.W, dt), to_dev(c.dO, dt), prob.pair_bwd, prob.mask_sorted,
                    prob.argsort, reverse_mask=True)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_dgrad(quantize(c.dO, dt), quantize(c.W, dt), prob.pair_bwd, prob.mask_sorted, prob.argsort,
                         16, reverse_mask=True)
    assert _rel(din.double().cpu().numpy(), quantize(exp, dt)) <= TOL[dt]


@pytest.mark.parametrize("dt", ["f16", "bf16"])
def test_wide_layers_streamed_weights(dt):
    """128 -> 128 channels: the weights do not fit the LDS for all output
    columns, so forward and dgrad run the streamed-weights kernel."""
    c = subm_case(2500, 128, 128, seed=17)
    prob = c.prob
    print(f"[wide] dt={dt}", flush=True)
    out, mo, desp = run_fwd(c, dt)
    exp, tm = ref.conv_forward(quantize(c.X, dt), quantize(c.W, dt), prob.pair_fwd, prob.mask_sorted,
                               prob.argsort, 16)
    assert np.array_equal(mo, tm)
    assert _rel(out, quantize(exp, dt)) <= TOL[dt]
    desp = desp_for(1, dt, c.C, c.K)
    din = torch.zeros((prob.n, c.C), dtype=TORCH_DT[dt], device="cuda")
    p = conv_params(1, desp, din, to_dev(c.W, dt), to_dev(c.dO, dt), prob.pair_bwd, prob.mask_sorted,
                    prob.argsort, reverse_mask=True)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_dgrad(quantize(c.dO, dt), quantize(c.W, dt), prob.pair_bwd, prob.mask_sorted, prob.argsort,
                         16, reverse_mask=True)
    assert _rel(din.double().cpu().numpy(), quantize(exp, dt)) <= TOL[dt]


@pytest.mark.parametrize("n", [2000, 200000])
def test_streamed_equals_weight_stationary(n):
    """The streamed-weights kernel (block_n 128) and the weight-stationary one
    (block_n 64) give the same bits: each row sums its offsets in order.  At
    200 000 rows each weight-stationary block orders a pool of more than 64
    tiles (row data of the later entries loads in the ordering pass)."""
    print(f"[stream-vs-ws] n={n}", flush=True)
    c = subm_case(n, 64, 64, seed=23, grid=(16, 96, 96) if n <= 20000 else (41, 1600, 1408))
    prob = c.prob
    outs = []
    for bn in (64, 128):
        desp = [d for d in ConvMainUnitTest.get_all_conv_algo_desp()
                if int(d.op_type) == 0 and d.dtype_input == DT["f16"] and d.dtype_output == DT["f16"]
                and d.tile_shape[1] == bn][0]
        out = torch.zeros((prob.n, c.K), dtype=torch.float16, device="cuda")
        p = conv_params(0, desp, to_dev(c.X, "f16"), to_dev(c.W, "f16"), out, prob.pair_fwd, prob.mask_sorted,
                        prob.argsort)
        ConvMainUnitTest.implicit_gemm2(p)
        torch.cuda.synchronize()
        outs.append(out.clone())
    assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("C,K,dt", [(128, 128, "f16"), (64, 128, "bf16"), (128, 64, "f32"), (128, 128, "e4m3"),
                                    (256, 256, "f16"), (192, 320, "bf16")])
def test_wgrad_wide_tiles(C, K, dt):
    """wgrad with 128-wide output tiles (each gathered row serves 128 channels);
    256 -> 256 and 192 -> 320 take several output tiles per piece (4 and 6), so
    the per-XCD block map runs with tiles innermost."""
    c = subm_case(2000, C, K, seed=C + 2 * K)
    prob = c.prob
    print(f"[wgrad-wide] C={C} K={K} dt={dt}", flush=True)
    _, mo, fdesp = run_fwd(c, "f32")
    out_dt = "f32" if dt == "e4m3" else dt
    desp = desp_for(2, dt, C, K, out_dt)
    dw = torch.zeros((K, prob.kv, C), dtype=TORCH_DT[out_dt], device="cuda")
    if dt == "e4m3":
        xq, x64 = _e4m3_quant(c.X, 1.0)
        dq, d64 = _e4m3_quant(c.dO, 1.0)
        xin, dout = xq.cuda(), dq.cuda()
    else:
        xin, dout = to_dev(c.X, dt), to_dev(c.dO, dt)
        x64, d64 = quantize(c.X, dt), quantize(c.dO, dt)
    p = conv_params(2, desp, xin, dw, dout, prob.pair_fwd, mo, prob.argsort, mask_width=16)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_wgrad(x64, d64, prob.pair_fwd, mo, prob.argsort, 16)
    tol = 1e-4 if dt == "e4m3" else max(TOL[dt], TOL[out_dt]) * 2
    assert _rel(dw.double().cpu().numpy(), quantize(exp, out_dt)) <= tol


@pytest.fixture(scope="module")
def case_c2():
    # subm 3x3x3, 120 000 voxels on 8 planes in 41 x 1600 x 1408, 64 -> 64
    return subm_case(120000, 64, 64, seed=5, grid=pc.GRID_KITTI)


@pytest.fixture(scope="module")
def case_c3():
    # subm 3x3x3, 200 000 voxels on 8 planes in 41 x 1600 x 1408, 128 -> 128
    return subm_case(200000, 128, 128, seed=6, grid=pc.GRID_KITTI)


def _full_size(c, dt):
    """Forward, dgrad (reversed forward masks) and wgrad (the forward's
    mask_output, mask_width = its tile) against the FP64 oracle, same
    tolerances as the small cases; mask_output bit-exact."""
    prob = c.prob
    print(f"[full-size] dt={dt} N={prob.n} P={prob.num_pairs} C={c.C} K={c.K}", flush=True)
    xq, wq, dq = quantize(c.X, dt), quantize(c.W, dt), quantize(c.dO, dt)
    out, mo, fdesp = run_fwd(c, dt)
    tile = fdesp.tile_shape[0]
    exp, tm = ref.conv_forward(xq, wq, prob.pair_fwd, prob.mask_sorted, prob.argsort, tile)
    assert np.array_equal(mo, tm), "mask_output must be bit-exact"
    assert _rel(out, quantize(exp, dt)) <= TOL[dt]

    din = torch.zeros((prob.n, c.C), dtype=TORCH_DT[dt], device="cuda")
    p = conv_params(1, desp_for(1, dt, c.C, c.K), din, to_dev(c.W, dt), to_dev(c.dO, dt), prob.pair_bwd,
                    prob.mask_sorted, prob.argsort, reverse_mask=True)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_dgrad(dq, wq, prob.pair_bwd, prob.mask_sorted, prob.argsort, 16, reverse_mask=True)
    assert _rel(din.double().cpu().numpy(), quantize(exp, dt)) <= TOL[dt]

    dw = torch.zeros((c.K, prob.kv, c.C), dtype=TORCH_DT[dt], device="cuda")
    p = conv_params(2, desp_for(2, dt, c.C, c.K), to_dev(c.X, dt), dw, to_dev(c.dO, dt), prob.pair_fwd, mo,
                    prob.argsort, mask_width=tile)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_wgrad(xq, dq, prob.pair_fwd, mo, prob.argsort, tile)
    assert _rel(dw.double().cpu().numpy(), quantize(exp, dt)) <= TOL[dt] * 2


@pytest.mark.parametrize("dt", ["f32", "f16", "bf16"])
def test_full_size_forward_dgrad_wgrad(case_c2, dt):
    """The 64-channel benchmark layer at full size (weight-stationary kernel
    for the 16-bit forms)."""
    _full_size(case_c2, dt)


@pytest.mark.parametrize("dt", ["f32", "f16", "bf16"])
def test_full_size_128_channels(case_c3, dt):
    """The 128-channel layer at full size (streamed-weights kernel for the
    16-bit forms, block kernel for FP32)."""
    _full_size(case_c3, dt)


def test_full_size_128_channels_fp8(case_c3):
    """FP8-E4M3 at 200 000 voxels, 128 -> 128, the C3 forms: forward (weights
    per output channel, f16 out), dgrad (weights per input channel, f16 out)
    and wgrad (FP32 dW, alpha = s_do * s_x) against the FP64 oracle on the
    dequantised operands, same bars as the small FP8 cases."""
    c = case_c3
    prob = c.prob
    print(f"[full-size fp8] N={prob.n} P={prob.num_pairs}", flush=True)
    s_x = np.abs(c.X).max() / 448.0
    s_do = np.abs(c.dO).max() / 448.0
    s_wk = np.abs(c.W).reshape(c.K, -1).max(axis=1) / 448.0
    s_wc = np.abs(c.W).max(axis=(0, 1)) / 448.0
    xq, xq64 = _e4m3_quant(c.X, s_x)
    dq, dq64 = _e4m3_quant(c.dO, s_do)
    wkq, wkq64 = _e4m3_quant(c.W, s_wk[:, None, None])
    wcq, wcq64 = _e4m3_quant(c.W, s_wc[None, None, :])

    desp = desp_for(0, "e4m3", c.C, c.K, "f16")
    tile = desp.tile_shape[0]
    out = torch.zeros((prob.n, c.K), dtype=torch.float16, device="cuda")
    mo = torch.zeros(((prob.n + tile - 1) // tile,), dtype=torch.int32, device="cuda")
    scale_f = (s_x * s_wk).astype(np.float32)
    p = conv_params(0, desp, xq.cuda(), wkq.cuda(), out, prob.pair_fwd, prob.mask_sorted, prob.argsort,
                    mask_output=mo, scale=torch.from_numpy(scale_f).cuda())
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp, tm = ref.conv_forward(xq64, wkq64, prob.pair_fwd, prob.mask_sorted, prob.argsort, tile, scale=scale_f)
    mo = mo.cpu().numpy().view(np.uint32)
    assert np.array_equal(mo, tm), "mask_output must be bit-exact"
    assert _rel(out.double().cpu().numpy(), quantize(exp, "f16")) <= TOL["f16"]

    din = torch.zeros((prob.n, c.C), dtype=torch.float16, device="cuda")
    scale_d = (s_do * s_wc).astype(np.float32)
    p = conv_params(1, desp_for(1, "e4m3", c.C, c.K, "f16"), din, wcq.cuda(), dq.cuda(), prob.pair_bwd,
                    prob.mask_sorted, prob.argsort, reverse_mask=True, scale=torch.from_numpy(scale_d).cuda())
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_dgrad(dq64, wcq64, prob.pair_bwd, prob.mask_sorted, prob.argsort, 16, reverse_mask=True)
    assert _rel(din.double().cpu().numpy(), quantize(exp * scale_d[None, :], "f16")) <= TOL["f16"]

    dw = torch.zeros((c.K, prob.kv, c.C), dtype=torch.float32, device="cuda")
    p = conv_params(2, desp_for(2, "e4m3", c.C, c.K, "f32"), xq.cuda(), dw, dq.cuda(), prob.pair_fwd, mo,
                    prob.argsort, mask_width=tile, alpha=float(s_do * s_x))
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.conv_wgrad(xq64, dq64, prob.pair_fwd, mo, prob.argsort, tile) * (s_do * s_x)
    assert _rel(dw.double().cpu().numpy(), exp) <= 1e-4
