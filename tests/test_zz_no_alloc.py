"""cs_implicit_gemm2_noalloc (the allocation-free entry point) against
cs_implicit_gemm2 through the C ABI: forward, dgrad and wgrad give the same
bits as the allocating entry point, on two calls in a row where the second has
a smaller input (the scratch keeps the first call's contents), including a
wgrad whose default layout exceeds the scratch and runs with fewer split
pieces (same partition at this size, so the same bits); a call that cannot fit
raises ValueError before any launch."""
import numpy as np
import pytest
import torch

from spconv_case import TORCH_DT, conv_params, desp_for, subm_case, to_dev

from cumm_amd import _abi
from cumm_amd.conv.main import ConvMainUnitTest

pytestmark = pytest.mark.gpu


def _run(op, dt, c, noalloc):
    prob = c.prob
    desp = desp_for(op, dt, c.C, c.K)
    if op == 0:
        out = torch.zeros((prob.n, c.K), dtype=TORCH_DT[dt], device="cuda")
        mo = torch.zeros(((prob.n + 15) // 16,), dtype=torch.int32, device="cuda")
        p = conv_params(0, desp, to_dev(c.X, dt), to_dev(c.W, dt), out, prob.pair_fwd, prob.mask_sorted,
                        prob.argsort, mask_output=mo)
        res = [out, mo]
    elif op == 1:
        din = torch.zeros((prob.n, c.C), dtype=TORCH_DT[dt], device="cuda")
        p = conv_params(1, desp, din, to_dev(c.W, dt), to_dev(c.dO, dt), prob.pair_bwd, prob.mask_sorted,
                        prob.argsort, reverse_mask=True)
        res = [din]
    else:
        from oracle import spconv_ref as ref
        tm = ref.tile_masks(prob.mask_sorted, prob.n, 16).astype(np.uint32)
        dw = torch.zeros((c.K, prob.kv, c.C), dtype=TORCH_DT[dt], device="cuda")
        p = conv_params(2, desp, to_dev(c.X, dt), dw, to_dev(c.dO, dt), prob.pair_fwd, tm, prob.argsort,
                        mask_width=16)
        res = [dw]
    if noalloc:
        ConvMainUnitTest.implicit_gemm2_noalloc(p)
    else:
        ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    return [r.clone() for r in res]


@pytest.mark.parametrize("dt,C,K", [("f16", 64, 64), ("f32", 32, 48), ("bf16", 128, 128)])
@pytest.mark.parametrize("op", [0, 1, 2])
def test_noalloc_equals_allocating_entry_point(op, dt, C, K):
    for n in (12000, 3000):  # second call smaller: stale scratch from the first
        print(f"[noalloc] op={op} dt={dt} C={C} K={K} N={n}", flush=True)
        c = subm_case(n, C, K, seed=n + op)
        want = _run(op, dt, c, noalloc=False)
        got = _run(op, dt, c, noalloc=True)
        for a, b in zip(got, want):
            assert torch.equal(a, b), f"op {op} N={n}: no-alloc result differs"


def test_noalloc_rejects_what_does_not_fit():
    # wgrad's pair lists alone need 2 * 27 * N * 4 bytes: more than the scratch at 400k rows
    n = 400_000
    print(f"[noalloc] oversize N={n}", flush=True)
    prob_pairs = torch.full((27, n), -1, dtype=torch.int32, device="cuda")
    desp = desp_for(2, "f16", 16, 16)
    from cumm_amd import tensorview as tv
    from cumm_amd.tensorview.gemm import ConvOpType, ConvParams
    p = ConvParams(3, ConvOpType(2))
    p.conv_algo_desp = desp
    p.input = tv.from_torch(torch.zeros((n, 16), dtype=torch.float16, device="cuda"))
    p.output = tv.from_torch(torch.zeros((n, 16), dtype=torch.float16, device="cuda"))
    p.weight = tv.from_torch(torch.zeros((16, 27, 16), dtype=torch.float16, device="cuda"))
    p.indices = tv.from_torch(prob_pairs)
    p.mask = tv.from_torch(torch.zeros(((n + 15) // 16,), dtype=torch.int32, device="cuda"))
    p.mask_argsort = tv.from_torch(torch.arange(n, dtype=torch.int32, device="cuda"))
    p.mask_width = 16
    with pytest.raises(ValueError, match="allocation-free"):
        ConvMainUnitTest.implicit_gemm2_noalloc(p)
    assert CS_LIMIT() == 64 << 20


def CS_LIMIT():
    import re
    from pathlib import Path
    h = (Path(_abi.__file__).resolve().parents[2] / "include" / "cumm_amd.h").read_text()
    m = re.search(r"#define CS_NOALLOC_SCRATCH_BYTES \((\d+)ll << (\d+)\)", h)
    return int(m.group(1)) << int(m.group(2))
