"""CPU tests of the drop-in boundary: the C ABI library loads and exports what
include/cumm_amd.h declares, the tv / tv.gemm shim behaves like the
reference's classes, kernel names / `supported` / workspace formulas match the
reference's own host code (oracle/_ref, compiled from the reference headers),
and the host-side validation raises the reference's error classes.  No GPU."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest
import torch

from cumm_amd import _abi
from cumm_amd import tensorview as tv
from cumm_amd.conv.main import ConvMainUnitTest, select_conv_algo
from cumm_amd.gemm.main import GemmMainUnitTest
from cumm_amd.tensorview.gemm import (Activation, ConvAlgoDesp, ConvOpType, ConvParams, GemmAlgoDesp,
                                      GemmParams, ShuffleStrideType)

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "cumm_amd.h"


# --------------------------------------------------------------- library / ABI
def test_library_exports_every_declared_symbol():
    decl = set(re.findall(r"\b(cs_[a-z0-9_]+)\s*\(", HEADER.read_text()))
    assert decl, "no declarations parsed"
    L = _abi.lib()
    for name in decl:
        assert hasattr(L, name), f"{name} declared in cumm_amd.h but not exported"
    assert decl == set(_abi.SYMBOLS)
    assert L.cs_abi_version() == _abi.ABI_VERSION


def test_struct_layout_matches_header():
    # sizes computed by the C compiler from the header vs the ctypes mirror
    src = """
#include "cumm_amd.h"
#include <stdio.h>
#include <stddef.h>
int main(){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(cs_tensor), sizeof(cs_gemm_algo_desp),
 sizeof(cs_conv_algo_desp), sizeof(cs_gemm_params), sizeof(cs_conv_params),
 offsetof(cs_conv_params, stream), offsetof(cs_gemm_params, stream));return 0;}
"""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "s.c"
        c.write_text(src)
        exe = Path(d) / "s"
        subprocess.run(["gcc", "-I", str(ROOT / "include"), str(c), "-o", str(exe)], check=True)
        got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    exp = [ctypes.sizeof(_abi.CsTensor), ctypes.sizeof(_abi.CsGemmAlgoDesp), ctypes.sizeof(_abi.CsConvAlgoDesp),
           ctypes.sizeof(_abi.CsGemmParams), ctypes.sizeof(_abi.CsConvParams),
           _abi.CsConvParams.stream.offset, _abi.CsGemmParams.stream.offset]
    assert got == exp


# --------------------------------------------------------------- tv shim
def test_tensor_shim():
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    t = tv.from_numpy(a)
    assert t.shape == [3, 4] and t.dim(1) == 4 and t.ndim == 2 and t.dtype == tv.float32
    assert t.device == -1 and not t.empty() and tv.Tensor().empty()
    assert np.array_equal(t.numpy(), a)
    assert t.view([4, 3]).shape == [4, 3]
    assert t.slice_first_axis(1, 3).shape == [2, 4]
    u = tv.from_numpy(np.array([1, 2, 0xFFFFFFFF], dtype=np.uint32))
    assert u.dtype == tv.uint32 and u.numpy()[2] == 0xFFFFFFFF
    z = tv.zeros([2, 2], tv.bfloat16, -1)
    assert z.dtype == tv.bfloat16 and z.nbytes == 8
    assert tv.dtype_short_str(tv.float_e4m3) == "e4m3" and tv.dtype_short_str(tv.bfloat16) == "bf16"
    # codes of include/tensorview/dtypes.h:39-60
    assert (tv.float32, tv.int32, tv.float16, tv.uint32, tv.bfloat16, tv.float_e4m3) == (0, 1, 7, 10, 12, 14)


def test_params_defaults_and_dtype_propagation():
    p = ConvParams(3, ConvOpType.Forward)
    assert (p.alpha, p.beta, p.mask_width, p.mask_filter, p.reverse_mask, p.split_k_slices) == \
        (1.0, 0.0, -1, 0xFFFFFFFF, False, 1)
    assert p.workspace.empty() and p.bias.empty() and p.act_type == Activation.None_
    g = GemmParams()
    g.a = tv.zeros([4, 4], tv.float16)
    g.b = tv.zeros([4, 4], tv.float16)
    g.c = tv.zeros([4, 4], tv.float32)
    assert (g.algo_desp.dtype_a, g.algo_desp.dtype_b, g.algo_desp.dtype_c) == (tv.float16, tv.float16, tv.float32)
    with pytest.raises(RuntimeError):
        g.d = tv.zeros([4], tv.float16)  # d dtype must equal c's
    with pytest.raises(RuntimeError):
        GemmAlgoDesp().check_valid()


def _ref_lib():
    path = ROOT / "oracle" / "_ref" / "libcumm_ref_host.so"
    if not path.exists():
        pytest.skip("reference host shim not built (needs the reference tree)")
    L = ctypes.CDLL(str(path))
    for n in ("ref_gemm_repr", "ref_conv_repr"):
        getattr(L, n).restype = ctypes.c_int
    return L


def test_kernel_names_match_reference_grammar():
    L = _ref_lib()
    buf = ctypes.create_string_buffer(512)
    for d in ConvMainUnitTest.get_all_conv_algo_desp():
        c = _abi.conv_desp_to_cs(d)
        L.ref_conv_repr(ctypes.byref(c), buf, 512)
        assert repr(d) == buf.value.decode() == _abi.algo_name_conv(c)
    for d in GemmMainUnitTest.get_all_algo_desp():
        c = _abi.desp_to_cs(d)
        L.ref_gemm_repr(ctypes.byref(c), buf, 512)
        assert repr(d) == buf.value.decode() == _abi.algo_name_gemm(c)


def test_supported_workspace_mnk_match_reference():
    L = _ref_lib()
    rng = np.random.default_rng(0)
    out = (ctypes.c_int * 3)()
    for d in ConvMainUnitTest.get_all_conv_algo_desp():
        c = _abi.conv_desp_to_cs(d)
        for _ in range(20):
            N, C, K = (int(v) for v in rng.integers(1, 300, 3))
            mw = int(rng.choice([16, 32, 48, 64]))
            m, n, k = ConvMainUnitTest.extract_mnk(int(d.op_type), N, C, K, 27, 1, 1, True)
            L.ref_implicit_gemm_mnk(int(d.op_type), N, C, K, 27, 1, 1, 1, out)
            assert [m, n, k] == list(out)
            assert d.supported(m, n, k, C, K, mw) == bool(L.ref_conv_supported(ctypes.byref(c), m, n, k, C, K, mw))
            split = int(rng.choice([1, 2, 32])) if (d.split_k_serial or d.split_k_parallel) else 1
            assert d.query_conv_workspace_size(m, n, k, split, 27) == \
                L.ref_conv_workspace(ctypes.byref(c), m, n, k, split, 27)
    for d in GemmMainUnitTest.get_all_algo_desp():
        c = _abi.desp_to_cs(d)
        for _ in range(10):
            m, n, k = (int(v) for v in rng.integers(1, 300, 3))
            assert d.supported(m, n, k) == bool(L.ref_gemm_supported(ctypes.byref(c), m, n, k))
            split = 8 if (d.split_k_serial or d.split_k_parallel) else 1
            assert d.query_workspace_size(m, n, k, split) == L.ref_gemm_workspace(ctypes.byref(c), m, n, k, split)
        if not (d.split_k_serial or d.split_k_parallel):
            with pytest.raises(ValueError):
                d.query_workspace_size(10, 10, 10, 2)  # TV_THROW_INVALID_ARG("not impemented")


def test_gemm_helper_statics():
    a2 = GemmMainUnitTest.align_to_power2
    assert [a2(v) for v in (1, 2, 3, 4, 5, 17, 64, 65)] == [1, 2, 4, 4, 8, 32, 64, 128]
    mnk = GemmMainUnitTest.extract_mnk([100, 32], [32, 64], False, False, False)
    assert mnk == (100, 64, 32)
    assert GemmMainUnitTest.extract_mnk([100, 32], [64, 32], False, True, False) == (100, 64, 32)
    assert GemmMainUnitTest.extract_mnk([100, 32], [32, 64], False, False, False, ShuffleStrideType.ShuffleAC,
                                        [70], [], [70]) == (70, 64, 32)
    assert GemmMainUnitTest.extract_mnk([100, 32], [100, 64], True, False, False, ShuffleStrideType.ShuffleAB,
                                        [70], [70], []) == (32, 64, 70)
    table = {16 | (32 << 20) | (8 << 40): [1], 64 | (64 << 20) | (8 << 40): [2]}
    assert GemmMainUnitTest.simple_select_tile_shape(50, 40, 8, [16, 64], [32, 64], [8], table, False) == [2]
    assert GemmMainUnitTest.simple_select_tile_shape(10, 10, 8, [16, 64], [32, 64], [8], table, False) == [1]


def test_catalogue_selection():
    d = select_conv_algo(0, tv.float16, tv.float16, tv.float16, 64, 64)
    assert d.op_type == ConvOpType.Forward and d.mask_sparse and d.tile_shape[0] == 16
    w = select_conv_algo(2, tv.float16, tv.float32, tv.float16, 64, 64)
    assert w.split_k_parallel and w.dtype_input == tv.float16 and w.dtype_weight == tv.float32
    with pytest.raises(ValueError):
        select_conv_algo(1, tv.int8, tv.int8, tv.int8, 64, 64)  # int8 is forward (inference) only


# --------------------------------------------------------------- validation (host only)
def _host_fwd_params(N=100, C=16, K=32, dt=torch.float16):
    d = select_conv_algo(0, tv.float16, tv.float16, tv.float16, C, K)
    p = ConvParams(3, ConvOpType.Forward)
    p.conv_algo_desp = d
    p.input = tv.from_torch(torch.zeros((N, C), dtype=dt))
    p.weight = tv.from_torch(torch.zeros((K, 27, C), dtype=dt))
    p.output = tv.from_torch(torch.zeros((N, K), dtype=dt))
    p.indices = tv.from_torch(torch.full((27, N), -1, dtype=torch.int32))
    p.mask = tv.from_torch(torch.zeros(N, dtype=torch.int32))
    p.mask_argsort = tv.from_torch(torch.arange(N, dtype=torch.int32))
    return p


def test_validation_errors_match_reference_classes():
    run = ConvMainUnitTest.implicit_gemm2
    p = _host_fwd_params()
    p.input = tv.from_torch(torch.zeros((100, 16, 1), dtype=torch.float16))
    with pytest.raises(RuntimeError):
        run(p)  # input rank (nvrtc_code.py:49)
    p = _host_fwd_params()
    p.output = tv.from_torch(torch.zeros((100, 48), dtype=torch.float16))
    with pytest.raises(RuntimeError):
        run(p)  # K2 != K
    p = _host_fwd_params()
    p.output_add = tv.from_torch(torch.zeros((100, 32), dtype=torch.float16))
    with pytest.raises(ValueError):
        run(p)  # only int8 inference supports output_add
    p = _host_fwd_params()
    p.indices = tv.Tensor()
    with pytest.raises(RuntimeError):
        run(p)
    p = _host_fwd_params()
    p.indices = tv.from_torch(torch.full((26, 100), -1, dtype=torch.int32))
    with pytest.raises(RuntimeError):
        run(p)  # indices must be [kv, -1]
    p = _host_fwd_params()
    p.output = tv.from_torch(torch.zeros((99, 32), dtype=torch.float16))
    with pytest.raises(RuntimeError):
        run(p)  # N == output.dim(0)
    p = _host_fwd_params(C=12)
    with pytest.raises(RuntimeError):
        run(p)  # C % element_per_access (supported())
    p = _host_fwd_params()
    p.conv_algo_desp = p.conv_algo_desp.copy()
    p.conv_algo_desp.tile_shape = [32, 64, 32]
    with pytest.raises(ValueError, match="Can't Found Algorithm"):
        run(p)
    p = _host_fwd_params()
    p.split_k_slices = 4
    with pytest.raises(ValueError):
        run(p)  # the reference's never-firing split-K guard, enforced here
    p = _host_fwd_params(dt=torch.float32)
    with pytest.raises(ValueError):
        run(p)  # tensor dtypes must match the algo
    p = _host_fwd_params()
    with pytest.raises(ValueError, match="GPU"):
        run(p)  # host tensors never reach a kernel
    w = _host_fwd_params()
    w.conv_algo_desp = select_conv_algo(2, tv.float16, tv.float16, tv.float16, 16, 32)
    w.mask_width = 8
    with pytest.raises(RuntimeError):
        run(w)  # wgrad mask_width % tile_shape[2]


def test_gemm_validation_errors():
    g = GemmParams()
    with pytest.raises(RuntimeError):
        GemmMainUnitTest.matmul2(g)  # desp unset
    d = [x for x in GemmMainUnitTest.get_all_algo_desp() if x.dtype_a == tv.float32
         and x.shuffle_type == ShuffleStrideType.ShuffleAC][0]
    g.algo_desp = d
    g.a = tv.from_torch(torch.zeros((10, 8)))
    g.b = tv.from_torch(torch.zeros((8, 4)) if not d.trans_b else torch.zeros((4, 8)))
    g.c = tv.from_torch(torch.zeros((10, 4)))
    with pytest.raises(RuntimeError):
        GemmMainUnitTest.matmul2(g)  # ShuffleAC needs c_inds
    g.c_inds = tv.from_torch(torch.arange(10, dtype=torch.int32))
    g.b = tv.from_torch(torch.zeros((12, 4)) if not d.trans_b else torch.zeros((4, 12)))
    with pytest.raises(ValueError):
        GemmMainUnitTest.matmul2(g)  # k != k2


# --------------------------------------------------------------- gather / scatter
def test_gather_scatter_params_and_supported():
    from cumm_amd.gemm.gather import GatherAll, ScatterAll
    g = GatherAll.get_all_gather_params()
    s = ScatterAll.get_all_scatter_params()
    # a set is published iff num_threads * bytes_per_access <= tile_m * tile_k_bytes
    # (cumm/gemm/gather.py:407, 516); the smallest tile that holds 64 threads x 4 B
    # is 8 x 32 B, the first entry of the reference's tile-k-major enumeration
    assert g[0] == (8, 32, 4, 64)
    assert (1, 256, 4, 64) in g and (1, 256, 8, 64) not in g and (8, 2048, 16, 256) in g
    assert (1, 4, 4, 64) not in g
    assert len(set(g)) == len(g) and len(set(s)) == len(s)
    assert all(nt * b <= tm * tk for tm, tk, b, nt in g + s)
    assert all(nt in (128, 256) and tk >= 64 for _, tk, _, nt in s)
    assert s[0] == (8, 64, 4, 128) and (4, 64, 4, 128) not in s
    # supported: (channel_size * itemsize) % bytes_per_access == 0
    assert GatherAll.supported(16, 4, tv.float32) and not GatherAll.supported(16, 6, tv.float32)
    assert GatherAll.supported(8, 4, tv.float16) and not GatherAll.supported(16, 4, tv.float16)
    assert GatherAll.supported(4, 12, tv.int8) and not GatherAll.supported(8, 12, tv.int8)
    sa = ScatterAll()
    assert sa.supported_scatter(8, 64, 4, 128, 3, tv.float32)
    assert not sa.supported_scatter(8, 64, 4, 128, 3, tv.float16)   # float32 kernels only
    assert not sa.supported_scatter(8, 64, 4, 64, 4, tv.float32)    # unpublished set
    assert not sa.supported_scatter(8, 256, 16, 128, 2, tv.float32)  # 8 B rows vs 16 B access


def test_gather_scatter_validation():
    from cumm_amd.gemm.gather import GatherAll, ScatterAll
    x = tv.from_torch(torch.zeros((10, 8), dtype=torch.float32))
    y = tv.from_torch(torch.zeros((4, 8), dtype=torch.float32))
    idx = tv.from_torch(torch.zeros((4,), dtype=torch.int32))
    with pytest.raises(RuntimeError, match="sutable gather"):
        GatherAll().gather(y, x, idx, 1, 4, 4, 64)
    with pytest.raises(RuntimeError, match="sutable scatter"):
        ScatterAll().scatter(x, y, idx, 8, 64, 4, 64)
    with pytest.raises(RuntimeError, match="not supported"):
        GatherAll().gather(tv.from_torch(torch.zeros((4, 3), dtype=torch.float16)),
                           tv.from_torch(torch.zeros((10, 3), dtype=torch.float16)), idx, 8, 32, 4, 64)
    with pytest.raises(RuntimeError):  # channel mismatch
        GatherAll().gather(tv.from_torch(torch.zeros((4, 4), dtype=torch.float32)), x, idx, 8, 32, 4, 64)
    with pytest.raises(ValueError, match="GPU"):
        GatherAll().gather(y, x, idx, 8, 32, 4, 64)  # host tensors never reach a kernel


def test_int8_inference_catalogue_and_validation():
    run = ConvMainUnitTest.implicit_gemm2
    i8 = [d for d in ConvMainUnitTest.get_all_conv_algo_desp() if d.is_int8_inference]
    assert sorted(d.tile_shape[1] for d in i8) == [16, 32, 64, 128]
    for d in i8:
        assert int(d.op_type) == 0 and (d.dtype_a, d.dtype_b, d.dtype_c) == (tv.int8, tv.int8, tv.int8)
        assert d.dacc == tv.int32 and d.dcomp == tv.float32
        assert repr(d).endswith("_S8") and d.element_per_access_a == 16
    d = select_conv_algo(0, tv.int8, tv.int8, tv.int8, 64, 64)
    assert d.is_int8_inference and d.tile_shape[1] == 64

    def params(**kw):
        N, C, K = 100, 64, 32
        p = ConvParams(3, ConvOpType.Forward)
        p.conv_algo_desp = select_conv_algo(0, tv.int8, tv.int8, tv.int8, C, K)
        p.input = tv.from_torch(torch.zeros((N, C), dtype=torch.int8))
        p.weight = tv.from_torch(torch.zeros((K, 27, C), dtype=torch.int8))
        p.output = tv.from_torch(torch.zeros((N, K), dtype=torch.int8))
        p.indices = tv.from_torch(torch.full((27, N), -1, dtype=torch.int32))
        p.mask = tv.from_torch(torch.zeros(N, dtype=torch.int32))
        p.mask_argsort = tv.from_torch(torch.arange(N, dtype=torch.int32))
        p.bias = tv.from_torch(kw.get("bias", torch.zeros(K)))
        p.scale = tv.from_torch(torch.ones(K))
        if "output_add" in kw:
            p.output_add = tv.from_torch(kw["output_add"])
        return p

    p = params()
    p.scale = tv.Tensor()
    with pytest.raises(RuntimeError, match="scale and bias"):
        run(p)  # nvrtc_code.py:81-82
    with pytest.raises(RuntimeError, match="same dtype and shape"):
        run(params(output_add=torch.zeros((100, 16), dtype=torch.int8)))  # nvrtc_code.py:85-87
    with pytest.raises(RuntimeError, match="same dtype and shape"):
        run(params(output_add=torch.zeros((100, 32), dtype=torch.float16)))
    with pytest.raises(ValueError, match="GPU"):
        run(params(output_add=torch.zeros((100, 32), dtype=torch.int8)))  # otherwise valid, host tensors


def test_dynamic_mask_and_int8_gemm_validation():
    """Host checks of the dynamic-mask word count, mask_int_count and the
    int8 GEMM epilogue (no device work is reached)."""
    run = ConvMainUnitTest.implicit_gemm2
    N, C, K = 100, 16, 16
    p = ConvParams(3, ConvOpType.Forward)
    p.conv_algo_desp = select_conv_algo(0, tv.float16, tv.float16, tv.float16, C, K, True)
    p.input = tv.from_torch(torch.zeros((N, C), dtype=torch.float16))
    p.weight = tv.from_torch(torch.zeros((K, 343, C), dtype=torch.float16))
    p.output = tv.from_torch(torch.zeros((N, K), dtype=torch.float16))
    p.indices = tv.from_torch(torch.full((343, N), -1, dtype=torch.int32))
    p.mask = tv.from_torch(torch.zeros((N, 11), dtype=torch.int32))
    p.mask_argsort = tv.from_torch(torch.arange(N, dtype=torch.int32))
    p.mask_int_count = 3
    with pytest.raises(ValueError, match="mask_int_count"):
        run(p)
    p.mask_int_count = 11
    with pytest.raises(ValueError, match="GPU"):
        run(p)  # the word count checks pass; host tensors stop it
    p.weight = tv.from_torch(torch.zeros((K, 1331, C), dtype=torch.float16))
    p.indices = tv.from_torch(torch.full((1331, N), -1, dtype=torch.int32))
    p.mask_int_count = 1
    with pytest.raises(ValueError, match="up to 1024"):
        run(p)  # 11 x 11 x 11
    d = [x for x in GemmMainUnitTest.get_all_algo_desp() if x.dtype_a == tv.int8 and x.dtype_c == tv.int8
         and x.shuffle_type == ShuffleStrideType.NoShuffle][0]
    g = GemmParams()
    g.algo_desp = d
    g.a = tv.from_torch(torch.zeros((64, 32), dtype=torch.int8))
    g.b = tv.from_torch(torch.zeros((32, 48) if not d.trans_b else (48, 32), dtype=torch.int8))
    g.c = tv.from_torch(torch.zeros((64, 48), dtype=torch.int8))
    g.act_type = Activation.Sigmoid
    with pytest.raises(ValueError, match="sigmoid"):
        GemmMainUnitTest.matmul2(g)
    g.act_type = Activation.None_
    g.alpha = float("inf")
    with pytest.raises(ValueError, match="int32"):
        GemmMainUnitTest.matmul2(g)


def test_pair_geometry_validation():
    """cs_pairs_*_geom refuse bad geometries before any device work."""
    L = _abi.lib()
    nb = ctypes.c_int64(0)

    def geom(k=(3, 3, 3), s=(2, 2, 2), p=(1, 1, 1), d=(1, 1, 1)):
        g = _abi.CsConvGeom()
        for i in range(3):
            g.ksize[i], g.stride[i], g.padding[i], g.dilation[i] = k[i], s[i], p[i], d[i]
        return g

    assert L.cs_pairs_subm_geom_workspace(1000, ctypes.byref(geom(k=(5, 5, 5))), ctypes.byref(nb)) == 0 and nb.value > 0
    assert L.cs_pairs_subm_geom_workspace(1000, ctypes.byref(geom(k=(4, 3, 3))), ctypes.byref(nb)) == 1
    assert b"odd" in L.cs_pairs_last_error()
    assert L.cs_pairs_subm_geom_workspace(1000, ctypes.byref(geom(k=(11, 11, 11))), ctypes.byref(nb)) == 1
    assert b"1024" in L.cs_pairs_last_error()
    assert L.cs_pairs_subm_geom_workspace(1000, ctypes.byref(geom(d=(0, 1, 1))), ctypes.byref(nb)) == 1
    assert L.cs_pairs_strided_geom_workspace(1000, 500, ctypes.byref(geom(k=(2, 2, 2), p=(0, 0, 0))),
                                             ctypes.byref(nb)) == 0
    assert L.cs_pairs_strided_geom_workspace(1000, 500, ctypes.byref(geom(s=(0, 2, 2))), ctypes.byref(nb)) == 1
    assert L.cs_pairs_strided_out_geom_workspace(1000, ctypes.byref(geom(p=(-1, 0, 0))), 0, ctypes.byref(nb)) == 1
    assert L.cs_pairs_subm_geom_workspace(1000, None, ctypes.byref(nb)) == 1


def test_row_masks_and_unsorted_tables_validation():
    """cs_pairs_row_masks checks its arguments before any device work (n = 0
    is a no-op), and the table builders take the forward mask outputs both
    set or both null."""
    L = _abi.lib()
    assert L.cs_pairs_row_masks(None, -1, 27, None, None, 0) == 1
    assert b"bad argument" in L.cs_pairs_last_error()
    assert L.cs_pairs_row_masks(None, 10, 0, None, None, 0) == 1
    assert L.cs_pairs_row_masks(None, 10, 1025, None, None, 0) == 1
    assert L.cs_pairs_row_masks(None, 10, 27, None, None, 0) == 1  # rows but no pair table / mask buffer
    assert L.cs_pairs_row_masks(None, 0, 27, None, None, 0) == 0
    g = _abi.CsConvGeom()
    for i in range(3):
        g.ksize[i], g.stride[i], g.padding[i], g.dilation[i] = 3, 1, 0, 1
    shape = (ctypes.c_int32 * 3)(16, 96, 96)
    # a mask buffer without an argsort (or the reverse) is refused before any
    # buffer is touched: the non-null values below are never dereferenced
    assert L.cs_pairs_subm_geom(None, 10, shape, ctypes.byref(g), None, None, 64, None, None, 0, 0) == 1
    assert b"both set or both null" in L.cs_pairs_last_error()
    assert L.cs_pairs_subm_geom(None, 10, shape, ctypes.byref(g), None, None, None, 64, None, 0, 0) == 1
