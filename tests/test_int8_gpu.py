"""int8-inference sparse-conv forward (``is_int8_inference`` kernels) on the GPU.

int8 features (per-tensor scale) and int8 weights (per-output-channel scales),
int32 accumulation, float epilogue ``act(scale * acc + bias + beta * source)``
clamped and rounded to int8.  The accumulator is exact, so the outputs must
equal the oracle's bit for bit (stated tolerance: 0), for every kernel the
catalogue reaches (weight-stationary column slices 16 / 32 / 64, streamed
weights, the block kernel for reductions that are not a multiple of 64).
"""
import numpy as np
import pytest
import torch

from oracle import spconv_ref as ref
from spconv_case import conv_params, subm_case

from cumm_amd import tensorview as tv
from cumm_amd.conv.main import ConvMainUnitTest

pytestmark = pytest.mark.gpu


def _desp(bn):
    for d in ConvMainUnitTest.get_all_conv_algo_desp():
        if int(d.op_type) == 0 and d.is_int8_inference and d.tile_shape[1] == bn:
            assert (d.dtype_input, d.dtype_weight, d.dtype_output) == (tv.int8, tv.int8, tv.int8)
            return d
    raise AssertionError(f"no int8 inference kernel with tile n {bn}")


def _quantised(c, seed):
    """Symmetric int8 quantisation; scale maps the accumulator to output units."""
    rng = np.random.default_rng(seed)
    s_x = np.abs(c.X).max() / 127.0
    s_w = np.abs(c.W).reshape(c.K, -1).max(axis=1) / 127.0
    xq = np.clip(np.rint(c.X / s_x), -127, 127).astype(np.int8)
    wq = np.clip(np.rint(c.W / s_w[:, None, None]), -127, 127).astype(np.int8)
    full, _ = ref.conv_forward(c.X, c.W, c.prob.pair_fwd, c.prob.mask_sorted, c.prob.argsort, 16)
    s_out = 0.8 * np.abs(full).max() / 127.0  # the largest outputs saturate
    scale = (s_x * s_w / s_out).astype(np.float32)
    bias = (rng.standard_normal(c.K) * 4).astype(np.float32)
    acc, tm = ref.conv_forward(xq, wq, c.prob.pair_fwd, c.prob.mask_sorted, c.prob.argsort, 16)
    return xq, wq, scale, bias, acc.astype(np.int64), tm


def _run(c, desp, xq, wq, scale, bias, out_init=None, **kw):
    prob = c.prob
    out = (torch.from_numpy(out_init).cuda() if out_init is not None
           else torch.zeros((prob.n, c.K), dtype=torch.int8, device="cuda"))
    tiles = (prob.n + 15) // 16
    mo = torch.zeros(max(tiles, 1), dtype=torch.int32, device="cuda")
    p = conv_params(0, desp, torch.from_numpy(xq).cuda(), torch.from_numpy(wq).cuda(), out, prob.pair_fwd,
                    prob.mask_sorted, prob.argsort, mask_output=mo, bias=torch.from_numpy(bias).cuda(),
                    scale=torch.from_numpy(scale).cuda(), **kw)
    return p, out, mo, tiles


@pytest.mark.parametrize("C,K,bn", [(64, 64, 64), (64, 64, 16), (64, 48, 32), (128, 128, 128), (128, 128, 64),
                                    (256, 64, 64), (32, 48, 64), (16, 16, 64)])
def test_int8_inference_bit_exact(C, K, bn):
    print(f"[int8] C={C} K={K} bn={bn}", flush=True)
    c = subm_case(3000, C, K, seed=C + K + bn)
    xq, wq, scale, bias, acc, tm = _quantised(c, seed=C * K)
    p, out, mo, tiles = _run(c, _desp(bn), xq, wq, scale, bias)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.int8_inference_epilogue(acc, scale, bias)
    got = out.cpu().numpy()
    assert np.array_equal(mo.cpu().numpy().view(np.uint32)[:tiles], tm)
    assert np.array_equal(got, exp), f"{np.count_nonzero(got != exp)} of {got.size} differ"
    # the largest-magnitude outputs saturate (at 127 or -128, whichever sign they have)
    assert ((got == 127) | (got == -128)).any() and (np.abs(got.astype(np.int32)) < 100).any()


@pytest.mark.parametrize("bn", [64, 128])
def test_int8_inference_output_add_beta_act(bn):
    print(f"[int8] output_add bn={bn}", flush=True)
    c = subm_case(2500, 64, 64, seed=41)
    xq, wq, scale, bias, acc, _ = _quantised(c, seed=5)
    rng = np.random.default_rng(8)
    add = rng.integers(-128, 128, (c.prob.n, c.K)).astype(np.int8)
    # output_add with beta and ReLU
    p, out, _, _ = _run(c, _desp(bn), xq, wq, scale, bias, beta=0.5, act_type=1)
    add_dev = torch.from_numpy(add).cuda()
    p.output_add = tv.from_torch(add_dev)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.int8_inference_epilogue(acc, scale, bias, beta=0.5, source=add, act_type=1)
    assert np.array_equal(out.cpu().numpy(), exp)
    # no output_add: the output itself is the source (read before it is written)
    init = rng.integers(-128, 128, (c.prob.n, c.K)).astype(np.int8)
    p, out, _, _ = _run(c, _desp(bn), xq, wq, scale, bias, out_init=init, beta=-1.25, act_type=3,
                        act_alpha=0.25)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    exp = ref.int8_inference_epilogue(acc, scale, bias, beta=-1.25, source=init, act_type=3, act_alpha=0.25)
    assert np.array_equal(out.cpu().numpy(), exp)


def test_int8_inference_requires_scale_and_bias():
    c = subm_case(500, 64, 64, seed=3)
    xq, wq, scale, bias, _, _ = _quantised(c, seed=1)
    p, _, _, _ = _run(c, _desp(64), xq, wq, scale, bias)
    p.bias = tv.Tensor()
    with pytest.raises(RuntimeError, match="scale and bias"):
        ConvMainUnitTest.implicit_gemm2(p)
    p.bias = tv.from_torch(torch.zeros(c.K, dtype=torch.float16, device="cuda"))
    with pytest.raises(ValueError, match="float32"):
        ConvMainUnitTest.implicit_gemm2(p)  # bias is in the compute type
