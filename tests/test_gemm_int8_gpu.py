"""GPU parity of the int8 matmul2 kernels ("s8,s8,s8,s32,s32" and
"s8,s8,s32,s32,s32", cumm/gemm/main.py:348-366, 499-523) and of the
"f16,f16,f16,f16,f16" entries.

int8: bit-exact against the integer oracle (oracle/gemm_ref.matmul2_int:
exact int32 accumulation, int32 epilogue with truncated alpha / beta /
act_alpha, int8 saturation), for NoShuffle (every trans combination),
ShuffleAC (gathered / scattered rows, n-major b and row-major b, the latter
read in place or, when its row stride is not whole 16-byte units, transposed
first; a bias, a source, ReLU / LeakyReLU) and
ShuffleAB (several split counts: int32 partial sums are exact in any order).
Full-range int8 operands make saturation frequent.  f16 accumulate entries:
relative error <= 1e-3 against FP64 on the f16 inputs.
"""
import numpy as np
import pytest
import torch

from oracle import gemm_ref, spconv_ref as ref
from spconv_case import TORCH_DT, quantize

from cumm_amd import tensorview as tv
from cumm_amd.gemm.main import GemmMainUnitTest
from cumm_amd.tensorview.gemm import GemmParams

pytestmark = pytest.mark.gpu


def _desp(dta, dtc, shuffle, ta, tb, tc=False, dacc=None):
    for d in GemmMainUnitTest.get_all_algo_desp():
        if (d.dtype_a == dta and d.dtype_c == dtc and int(d.shuffle_type) == int(shuffle) and d.trans_a == ta
                and d.trans_b == tb and d.trans_c == tc and (dacc is None or d.dacc == dacc)):
            return d
    raise LookupError((dta, dtc, shuffle, ta, tb, tc, dacc))


def _t(x, dtype):
    return torch.from_numpy(np.ascontiguousarray(x).astype(dtype)).cuda()


def _run(desp, a, b, c, d=None, a_inds=None, b_inds=None, c_inds=None, alpha=1.0, beta=0.0, act=0,
         act_alpha=0.0, split=1):
    p = GemmParams()
    p.algo_desp = desp
    p.a, p.b, p.c = tv.from_torch(a), tv.from_torch(b), tv.from_torch(c)
    if d is not None:
        p.d = tv.from_torch(d)
    for name, t in (("a_inds", a_inds), ("b_inds", b_inds), ("c_inds", c_inds)):
        if t is not None:
            setattr(p, name, tv.from_torch(torch.from_numpy(np.asarray(t, np.int32)).cuda()))
    p.alpha, p.beta, p.act_type, p.act_alpha, p.split_k_slices = alpha, beta, act, act_alpha, split
    GemmMainUnitTest.matmul2(p)
    torch.cuda.synchronize()
    return c.cpu().numpy()


OUT = {"s8": (tv.int8, np.int8), "s32": (tv.int32, np.int32)}


@pytest.mark.parametrize("out", ["s8", "s32"])
@pytest.mark.parametrize("ta,tb,tc", [(False, False, False), (True, False, False), (False, True, False),
                                      (True, True, False), (False, False, True)])
def test_int8_noshuffle(out, ta, tb, tc):
    print(f"[gemm-i8] noshuffle {out} {ta} {tb} {tc}", flush=True)
    rng = np.random.default_rng(7)
    m, n, k = 304, 96, 80
    A = rng.integers(-128, 128, (m, k))
    Bm = rng.integers(-128, 128, (k, n))
    dtc, npc = OUT[out]
    c = torch.zeros((m, n), dtype=torch.int8 if out == "s8" else torch.int32, device="cuda")
    got = _run(_desp(tv.int8, dtc, 0, ta, tb, tc), _t(A.T if ta else A, np.int8), _t(Bm.T if tb else Bm, np.int8), c)
    exp = gemm_ref.matmul2_int(A, Bm, np.zeros((m, n), npc))
    assert got.dtype == npc and np.array_equal(got, exp)
    if out == "s8":
        assert (got == 127).any() and (got == -128).any()  # saturation exercised


@pytest.mark.parametrize("out", ["s8", "s32"])
@pytest.mark.parametrize("tb,n,pad", [(False, 128, 0), (True, 128, 0), (False, 48, 8), (True, 48, 0)])
def test_int8_shuffle_ac(out, tb, n, pad):
    """c[c_inds] = epi(a[a_inds] @ b).  A row-major b whose row stride is not
    a multiple of 16 elements (pad = 8) is transposed into the workspace
    first; with whole 16-byte rows it is read k-major in place."""
    print(f"[gemm-i8] shuffleAC {out} tb={tb} n={n} pad={pad}", flush=True)
    rng = np.random.default_rng(12315)
    m_all, k = 6400 + 37, 128
    m = 5000
    a_inds = rng.permutation(m_all)[:m]
    c_inds = rng.permutation(m_all)[:m]
    a = rng.integers(-128, 128, (m_all, k))
    b = rng.integers(-128, 128, (k, n))
    dtc, npc = OUT[out]
    init = rng.integers(-128, 128, (m_all, n)).astype(npc)
    c = _t(init, npc)
    if tb:
        bdev = _t(b.T, np.int8)
    else:
        bdev = _t(np.pad(b, ((0, 0), (0, pad))), np.int8)[:, :n]
        assert bdev.stride(0) == n + pad
    got = _run(_desp(tv.int8, dtc, 1, False, tb), _t(a, np.int8), bdev, c, a_inds=a_inds, c_inds=c_inds)
    exp = gemm_ref.matmul2_int(a, b, init, shuffle=1, a_inds=a_inds, c_inds=c_inds)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("out", ["s8", "s32"])
@pytest.mark.parametrize("mode", ["bias_relu", "source_leaky", "alpha_trunc"])
def test_int8_shuffle_ac_epilogue(out, mode):
    print(f"[gemm-i8] shuffleAC epilogue {out} {mode}", flush=True)
    rng = np.random.default_rng(99)
    m_all, n, k = 3000, 64, 64
    sel = rng.permutation(m_all)[:2000]
    outrows = rng.permutation(m_all)[:2000]
    a = rng.integers(-128, 128, (m_all, k))
    b = rng.integers(-128, 128, (k, n))
    dtc, npc = OUT[out]
    init = rng.integers(-100, 100, (m_all, n)).astype(npc)
    kw = {}
    d = None
    if mode == "bias_relu":
        d = rng.integers(-100, 100, n).astype(npc)
        kw = dict(alpha=1.0, beta=1.0, act_type=1)
    elif mode == "source_leaky":
        kw = dict(alpha=2.0, beta=-3.0, act_type=3, act_alpha=2.0)
    else:
        kw = dict(alpha=1.7, beta=0.9, act_type=0)  # alpha -> 1, beta -> 0
    c = _t(init, npc)
    got = _run(_desp(tv.int8, dtc, 1, False, False), _t(a, np.int8), _t(b, np.int8), c,
               d=None if d is None else _t(d, npc), a_inds=sel, c_inds=outrows, alpha=kw["alpha"],
               beta=kw["beta"], act=kw["act_type"], act_alpha=kw.get("act_alpha", 0.0))
    exp = gemm_ref.matmul2_int(a, b, init, shuffle=1, a_inds=sel, c_inds=outrows, d=d, **kw)
    assert np.array_equal(got, exp)
    untouched = np.setdiff1d(np.arange(m_all), outrows)
    assert np.array_equal(got[untouched], init[untouched])


@pytest.mark.parametrize("out", ["s8", "s32"])
@pytest.mark.parametrize("split", [1, 5, 32])
def test_int8_shuffle_ab(out, split):
    print(f"[gemm-i8] shuffleAB {out} split={split}", flush=True)
    rng = np.random.default_rng(12315)
    rows, m, n = 6400 + 11, 64, 128
    a_inds = rng.permutation(rows)
    b_inds = rng.permutation(rows)
    a = rng.integers(-128, 128, (rows, m))
    b = rng.integers(-128, 128, (rows, n))
    dtc, npc = OUT[out]
    c = torch.zeros((m, n), dtype=torch.int8 if out == "s8" else torch.int32, device="cuda")
    got = _run(_desp(tv.int8, dtc, 2, True, False), _t(a, np.int8), _t(b, np.int8), c, a_inds=a_inds,
               b_inds=b_inds, split=split)
    exp = gemm_ref.matmul2_int(a, b, np.zeros((m, n), npc), trans_a=True, shuffle=2, a_inds=a_inds, b_inds=b_inds)
    assert np.array_equal(got, exp)


def test_int8_long_reduction_and_refusals():
    """k = 512 (two B slabs per group) stays exact; sigmoid and k not a
    multiple of 16 are refused."""
    print("[gemm-i8] long k / refusals", flush=True)
    rng = np.random.default_rng(5)
    m, n, k = 1000, 48, 512
    a = rng.integers(-128, 128, (m, k))
    b = rng.integers(-128, 128, (k, n))
    c = torch.zeros((m, n), dtype=torch.int32, device="cuda")
    got = _run(_desp(tv.int8, tv.int32, 0, False, False), _t(a, np.int8), _t(b, np.int8), c)
    assert np.array_equal(got, gemm_ref.matmul2_int(a, b, np.zeros((m, n), np.int32)))
    with pytest.raises(ValueError):
        _run(_desp(tv.int8, tv.int8, 0, False, False), _t(a, np.int8), _t(b, np.int8),
             torch.zeros((m, n), dtype=torch.int8, device="cuda"), act=2)
    with pytest.raises(ValueError):
        _run(_desp(tv.int8, tv.int8, 0, False, False), _t(a[:, :40], np.int8), _t(b[:40], np.int8),
             torch.zeros((m, n), dtype=torch.int8, device="cuda"))


@pytest.mark.parametrize("shuffle", [0, 1, 2])
def test_f16_accumulate_entries(shuffle):
    """The "f16,f16,f16,f16,f16" entries (FP32 accumulation inside); ShuffleAB
    runs on the f16 entry with FP32 accumulation (there is no f16-accumulate
    ShuffleAB entry)."""
    print(f"[gemm] f16-acc shuffle={shuffle}", flush=True)
    rng = np.random.default_rng(17)
    d = _desp(tv.float16, tv.float16, shuffle, shuffle == 2, False, dacc=tv.float32 if shuffle == 2 else tv.float16)
    if shuffle == 2:
        rows, m, n = 3000, 64, 96
        ai, bi = rng.permutation(rows), rng.permutation(rows)
        a = rng.uniform(-1, 1, (rows, m))
        b = rng.uniform(-1, 1, (rows, n))
        c = torch.zeros((m, n), dtype=torch.float16, device="cuda")
        got = _run(d, _t(a, np.float16), _t(b, np.float16), c, a_inds=ai, b_inds=bi).astype(np.float64)
        exp = quantize(a, "f16")[ai].T @ quantize(b, "f16")[bi]
    else:
        m, n, k = 2000, 96, 160
        a = rng.uniform(-1, 1, (m, k))
        b = rng.uniform(-1, 1, (k, n))
        ai = rng.permutation(m) if shuffle == 1 else None
        ci = rng.permutation(m) if shuffle == 1 else None
        c = torch.zeros((m, n), dtype=torch.float16, device="cuda")
        got = _run(d, _t(a, np.float16), _t(b, np.float16), c, a_inds=ai, c_inds=ci).astype(np.float64)
        exp = np.zeros((m, n))
        q = quantize(a, "f16") @ quantize(b, "f16")
        if shuffle == 1:
            exp[ci] = q[ai]
        else:
            exp = q
    assert ref.relative_error(got, quantize(exp, "f16"))[1] <= 1e-3
