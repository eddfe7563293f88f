"""Standalone gather / scatter-add (GatherAll / ScatterAll) on the GPU.

* gather is a byte copy: bit-exact against numpy fancy indexing for every
  dtype, with duplicate indices, ragged row widths (4-byte and 16-byte access
  paths), strided rows and zero rows;
* scatter-add of distinct indices is one float32 add per element: bit-exact
  against numpy's ``out[idx] += x``;
* round trip at full size (200 000 rows x 64 channels): scattering the gathered
  rows of a permutation back into zeros restores the input exactly.
"""
import numpy as np
import pytest
import torch

from cumm_amd import tensorview as tv
from cumm_amd.gemm.gather import GatherAll, ScatterAll

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
P = (8, 256, 16, 128)  # a published parameter set, valid for both


def _g(out, inp, idx, param=P):
    GatherAll().gather(tv.from_torch(out), tv.from_torch(inp), tv.from_torch(idx), *param)


def _s(out, inp, idx, param=P):
    ScatterAll().scatter(tv.from_torch(out), tv.from_torch(inp), tv.from_torch(idx), *param)


@pytest.mark.parametrize("dt,c", [(torch.float32, 64), (torch.float32, 3), (torch.float16, 64),
                                  (torch.float16, 6), (torch.bfloat16, 128), (torch.int8, 48),
                                  (torch.int8, 20), (torch.float32, 1)])
def test_gather_bit_exact(dt, c):
    print(f"[gather] {dt} c={c}", flush=True)
    rng = np.random.default_rng(c)
    n, m = 5000, 7000
    src = torch.from_numpy(rng.standard_normal((n, c)).astype(np.float32) * 20).to(dt)
    idx = rng.integers(0, n, m).astype(np.int32)  # duplicates included
    bpa = 16 if (c * src.element_size()) % 16 == 0 else 4
    param = (8, 256, bpa, 128)
    assert GatherAll.supported(bpa, c, int(tv.from_torch(src).dtype))
    out = torch.full((m, c), 7, dtype=dt, device=DEV)
    _g(out, src.to(DEV), torch.from_numpy(idx).to(DEV), param)
    torch.cuda.synchronize()
    exp = src.view(torch.uint8) if dt != torch.int8 else src
    got = out.cpu()
    assert torch.equal(got.view(torch.uint8) if dt != torch.int8 else got,
                       exp[torch.from_numpy(idx).long()])


def test_gather_strided_rows_and_empty():
    print("[gather] strided / empty", flush=True)
    rng = np.random.default_rng(2)
    big = torch.from_numpy(rng.standard_normal((300, 40)).astype(np.float32)).to(DEV)
    src = big[:, 8:24]                     # row stride 40, 16 channels
    out_big = torch.zeros((100, 32), dtype=torch.float32, device=DEV)
    out = out_big[:, :16]
    idx = rng.integers(0, 300, 100).astype(np.int32)
    _g(out, src, torch.from_numpy(idx).to(DEV))
    torch.cuda.synchronize()
    assert torch.equal(out.cpu(), src.cpu()[torch.from_numpy(idx).long()])
    assert torch.count_nonzero(out_big[:, 16:]) == 0  # nothing written past the row
    empty = torch.zeros((0,), dtype=torch.int32, device=DEV)
    _g(out, src, empty)
    _s(out, src[:0], empty)
    torch.cuda.synchronize()


def test_scatter_add_bit_exact():
    print("[scatter] f32", flush=True)
    rng = np.random.default_rng(4)
    for c in (64, 5, 128):
        n_out, m = 9000, 6000
        base = rng.standard_normal((n_out, c)).astype(np.float32)
        x = rng.standard_normal((m, c)).astype(np.float32)
        idx = rng.permutation(n_out)[:m].astype(np.int32)  # distinct targets
        out = torch.from_numpy(base).to(DEV)
        bpa = 16 if c % 4 == 0 else 4
        _s(out, torch.from_numpy(x).to(DEV), torch.from_numpy(idx).to(DEV), (8, 256, bpa, 128))
        torch.cuda.synchronize()
        exp = base.copy()
        exp[idx] += x
        assert np.array_equal(out.cpu().numpy(), exp), c


def test_gather_scatter_round_trip_full_size():
    print("[gather] round trip 200k x 64", flush=True)
    g = torch.Generator(device="cpu").manual_seed(9)
    n, c = 200_000, 64
    x = torch.randn((n, c), generator=g).to(DEV)
    perm = torch.randperm(n, generator=g).to(torch.int32).to(DEV)
    mid = torch.empty((n, c), device=DEV)
    back = torch.zeros((n, c), device=DEV)
    _g(mid, x, perm, (8, 2048, 16, 256))
    _s(back, mid, perm, (8, 2048, 16, 256))
    torch.cuda.synchronize()
    assert torch.equal(back, x)
    assert torch.equal(mid, x[perm.long()])
