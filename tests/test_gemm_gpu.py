"""GPU parity of GemmMainUnitTest.matmul2 (NoShuffle / ShuffleAC / ShuffleAB)
against the FP64 oracle and the reference's closed forms
(cumm/gemm/main_real.py:197-229, main_simu.py:643-647, 806-811).

Tolerances: relative Frobenius error vs FP64 on the quantised inputs:
f32 2e-6, f16 1e-3, bf16 8e-3.  The self-cancelling known answer uses the
reference's own criterion max|c| / max|c_ref| < 5e-3 (main_real.py:672-673).
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import gemm_ref, spconv_ref as ref
from spconv_case import TORCH_DT, quantize

from cumm_amd import tensorview as tv
from cumm_amd.gemm.main import GemmMainUnitTest
from cumm_amd.tensorview.gemm import GemmParams, ShuffleStrideType

pytestmark = pytest.mark.gpu
TOL = {"f32": 2e-6, "f16": 1e-3, "bf16": 8e-3}
DTV = {"f32": tv.float32, "f16": tv.float16, "bf16": tv.bfloat16}


def _desp(dt, shuffle, ta, tb, tc=False):
    for d in GemmMainUnitTest.get_all_algo_desp():
        if (d.dtype_a == DTV[dt] and int(d.shuffle_type) == int(shuffle) and d.trans_a == ta
                and d.trans_b == tb and d.trans_c == tc):
            return d
    raise LookupError((dt, shuffle, ta, tb, tc))


def _dev(x, dt):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(TORCH_DT[dt]).cuda()


def _run(desp, a, b, c, d=None, a_inds=None, b_inds=None, c_inds=None, alpha=1.0, beta=0.0, act=0,
         act_alpha=0.0, split=1):
    p = GemmParams()
    p.algo_desp = desp
    p.a, p.b, p.c = tv.from_torch(a), tv.from_torch(b), tv.from_torch(c)
    if d is not None:
        p.d = tv.from_torch(d)
    for name, t in (("a_inds", a_inds), ("b_inds", b_inds), ("c_inds", c_inds)):
        if t is not None:
            setattr(p, name, tv.from_torch(torch.from_numpy(t.astype(np.int32)).cuda()))
    p.alpha, p.beta, p.act_type, p.act_alpha, p.split_k_slices = alpha, beta, act, act_alpha, split
    GemmMainUnitTest.matmul2(p)
    torch.cuda.synchronize()
    return c.double().cpu().numpy()


def test_config1_plain_fp32_gemm():
    """Plain FP32 C = A.B, 128x128x64, U(-1,1) seeded; vs numpy FP64 and the CPU
    emulation of the kernel's accumulation order."""
    print("[gemm] config1", flush=True)
    rng = np.random.default_rng(12315)
    a = rng.uniform(-1, 1, (128, 64)).astype(np.float32)
    b = rng.uniform(-1, 1, (64, 128)).astype(np.float32)
    c = torch.zeros((128, 128), dtype=torch.float32, device="cuda")
    got = _run(_desp("f32", 0, False, False), _dev(a, "f32"), _dev(b, "f32"), c)
    exp = a.astype(np.float64) @ b.astype(np.float64)
    assert ref.relative_error(got, exp)[1] < 2e-7
    lib = Path(__file__).resolve().parent.parent / "oracle" / "_build" / "libgemm_emu.so"
    if lib.exists():
        L = ctypes.CDLL(str(lib))
        L.gemm_emu_f32.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int64] * 3
        bt = np.ascontiguousarray(b.T)
        emu = np.zeros((128, 128), np.float32)
        L.gemm_emu_f32(a.ctypes.data, bt.ctypes.data, emu.ctypes.data, 128, 128, 64)
        assert np.max(np.abs(got - emu)) < 2e-5


@pytest.mark.parametrize("dt", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("ta,tb,tc", [(False, False, False), (True, False, False), (False, True, False),
                                      (True, True, False), (False, False, True)])
def test_noshuffle(dt, ta, tb, tc):
    print(f"[gemm] noshuffle {dt} {ta} {tb} {tc}", flush=True)
    rng = np.random.default_rng(5)
    m, n, k = 304, 96, 72
    A = rng.uniform(-1, 1, (m, k))
    Bm = rng.uniform(-1, 1, (k, n))
    a = A.T if ta else A
    b = Bm.T if tb else Bm
    c = torch.zeros((m, n), dtype=TORCH_DT[dt], device="cuda")
    got = _run(_desp(dt, 0, ta, tb, tc), _dev(a, dt), _dev(b, dt), c)
    exp = quantize(quantize(A, dt) @ quantize(Bm, dt), dt)
    assert ref.relative_error(got, exp)[1] <= TOL[dt]


@pytest.mark.parametrize("dt", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("tb", [False, True])
def test_shuffle_ac_closed_form(dt, tb):
    """c[c_inds] = a[a_inds] @ b with shuffled index permutations
    (cumm/gemm/main_real.py:197-229), m not a multiple of the tile."""
    print(f"[gemm] shuffleAC {dt} tb={tb}", flush=True)
    rng = np.random.default_rng(12315)
    m, n, k = 6400 + 37, 128, 128
    a_inds = rng.permutation(m).astype(np.int32)
    c_inds = rng.permutation(m).astype(np.int32)
    a = rng.uniform(-1, 1, (m, k))
    b = rng.uniform(-1, 1, (k, n))
    c = torch.zeros((m, n), dtype=TORCH_DT[dt], device="cuda")
    got = _run(_desp(dt, 1, False, tb), _dev(a, dt), _dev(b.T if tb else b, dt), c, a_inds=a_inds, c_inds=c_inds)
    exp = np.zeros((m, n))
    exp[c_inds] = quantize(a, dt)[a_inds] @ quantize(b, dt)
    assert ref.relative_error(got, quantize(exp, dt))[1] <= TOL[dt]


def test_shuffle_ac_partial_rows_bias_and_source():
    """a_inds shorter than a (gather a subset), 1-D bias d, rows outside c_inds untouched."""
    rng = np.random.default_rng(3)
    m_all, n, k = 500, 64, 32
    sel = rng.permutation(m_all)[:333].astype(np.int32)
    outrows = rng.permutation(m_all)[:333].astype(np.int32)
    a = rng.uniform(-1, 1, (m_all, k)).astype(np.float32)
    b = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    bias = rng.uniform(-1, 1, n).astype(np.float32)
    init = rng.uniform(-1, 1, (m_all, n)).astype(np.float32)
    c = torch.from_numpy(init.copy()).cuda()
    got = _run(_desp("f32", 1, False, False), _dev(a, "f32"), _dev(b, "f32"), c, d=_dev(bias, "f32"),
               a_inds=sel, c_inds=outrows, alpha=0.5, beta=1.0, act=1)
    exp = gemm_ref.matmul2(a, b, init, shuffle=1, a_inds=sel, c_inds=outrows, d=bias, alpha=0.5, beta=1.0,
                           act_type=1)
    assert ref.relative_error(got, exp)[1] < 2e-6
    untouched = np.setdiff1d(np.arange(m_all), outrows)
    assert np.array_equal(got[untouched], init[untouched].astype(np.float64))


@pytest.mark.parametrize("dt", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("split", [1, 5, 32])
def test_shuffle_ab_closed_form(dt, split):
    """c = a[a_inds]^T @ b[b_inds] (backward-weight form)."""
    print(f"[gemm] shuffleAB {dt} split={split}", flush=True)
    rng = np.random.default_rng(12315)
    rows, m, n = 6400 + 11, 64, 128
    a_inds = rng.permutation(rows).astype(np.int32)
    b_inds = rng.permutation(rows).astype(np.int32)
    a = rng.uniform(-1, 1, (rows, m))
    b = rng.uniform(-1, 1, (rows, n))
    c = torch.zeros((m, n), dtype=TORCH_DT[dt], device="cuda")
    got = _run(_desp(dt, 2, True, False), _dev(a, dt), _dev(b, dt), c, a_inds=a_inds, b_inds=b_inds, split=split)
    exp = quantize(a, dt)[a_inds].T @ quantize(b, dt)[b_inds]
    assert ref.relative_error(got, quantize(exp, dt))[1] <= TOL[dt] * 2


def test_self_cancel_known_answer():
    """c = -A.B then matmul2 with beta = 1 must give ~0 (main_simu.py:720, 806-811)."""
    rng = np.random.default_rng(19260817)
    m, n, k = 316, 240, 656
    a = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    b = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    cref = gemm_ref.py_matmul(a, b)
    c = torch.from_numpy(-cref).cuda()
    got = _run(_desp("f32", 0, False, False), _dev(a, "f32"), _dev(b, "f32"), c, beta=1.0)
    assert np.max(np.abs(got)) / np.max(np.abs(cref)) < 5e-3
    assert np.max(np.abs(got)) < 1e-3
