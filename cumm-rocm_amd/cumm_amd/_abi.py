"""ctypes binding of libcumm_amd.so (declared in include/cumm_amd.h).

The structures below mirror the header field by field.  Error codes map to the
reference's exception classes: 1 -> ValueError (std::invalid_argument),
2 -> RuntimeError (std::runtime_error).  The library is loaded from this
package directory only; a missing library raises, there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import List, Optional

import torch  # noqa: F401  -- loads the HIP runtime the library binds to

from . import tensorview as tv

LIB_PATH = Path(__file__).resolve().parent / "libcumm_amd.so"
MAX_DIM = 4
ALGO_LEN = 32
ABI_VERSION = 1

# exported symbols, in header order
SYMBOLS = [
    "cs_abi_version", "cs_implicit_gemm2", "cs_implicit_gemm2_noalloc", "cs_matmul2", "cs_conv_workspace_size",
    "cs_gemm_workspace_size", "cs_get_all_conv_algo_desp", "cs_get_all_algo_desp",
    "cs_gemm_algo_name", "cs_conv_algo_name", "cs_device_synchronize",
    "cs_stream_synchronize", "cs_last_error",
    "cs_pairs_subm_workspace", "cs_pairs_subm", "cs_pairs_strided_out_workspace", "cs_pairs_strided_out",
    "cs_pairs_strided_workspace", "cs_pairs_strided", "cs_pairs_last_error",
    "cs_pairs_subm_geom_workspace", "cs_pairs_subm_geom", "cs_pairs_strided_out_geom_workspace",
    "cs_pairs_strided_out_geom", "cs_pairs_strided_geom_workspace", "cs_pairs_strided_geom", "cs_pairs_row_masks",
    "cs_get_all_gather_params", "cs_get_all_scatter_params", "cs_gather_supported", "cs_scatter_supported",
    "cs_gather", "cs_scatter",
]


class CsTensor(C.Structure):
    _fields_ = [
        ("data", C.c_void_p), ("dtype", C.c_int32), ("ndim", C.c_int32),
        ("shape", C.c_int64 * MAX_DIM), ("stride", C.c_int64 * MAX_DIM),
        ("device", C.c_int32), ("reserved", C.c_int32),
    ]


class CsGemmAlgoDesp(C.Structure):
    _fields_ = [
        ("dtype_a", C.c_int32), ("dtype_b", C.c_int32), ("dtype_c", C.c_int32),
        ("trans_a", C.c_int32), ("trans_b", C.c_int32), ("trans_c", C.c_int32),
        ("tile_shape", C.c_int32 * 3), ("warp_tile_shape", C.c_int32 * 3),
        ("num_stage", C.c_int32), ("dacc", C.c_int32), ("dcomp", C.c_int32),
        ("algo", C.c_char * ALGO_LEN), ("tensorop", C.c_int32 * 3),
        ("split_k_serial", C.c_int32), ("split_k_parallel", C.c_int32),
        ("shuffle_type", C.c_int32),
        ("element_per_access_a", C.c_int32), ("element_per_access_b", C.c_int32),
        ("element_per_access_c", C.c_int32), ("access_per_vector", C.c_int32),
        ("is_nvrtc", C.c_int32), ("min_arch", C.c_int32 * 2),
    ]


class CsConvAlgoDesp(C.Structure):
    _fields_ = [
        ("gemm", CsGemmAlgoDesp),
        ("ndim", C.c_int32), ("op_type", C.c_int32), ("iter_algo", C.c_int32),
        ("layout_i", C.c_int32), ("layout_w", C.c_int32), ("layout_o", C.c_int32),
        ("interleave_i", C.c_int32), ("interleave_w", C.c_int32), ("interleave_o", C.c_int32),
        ("mask_sparse", C.c_int32), ("increment_k_first", C.c_int32),
        ("dynamic_mask", C.c_int32), ("is_int8_inference", C.c_int32),
    ]


class CsConvGeom(C.Structure):
    """cs_conv_geom: kernel size, stride, padding, dilation per (z, y, x)."""
    _fields_ = [("ksize", C.c_int32 * 3), ("stride", C.c_int32 * 3), ("padding", C.c_int32 * 3),
                ("dilation", C.c_int32 * 3)]


class CsGemmParams(C.Structure):
    _fields_ = [
        ("algo_desp", CsGemmAlgoDesp),
        ("a", CsTensor), ("b", CsTensor), ("c", CsTensor), ("d", CsTensor),
        ("workspace", CsTensor), ("a_inds", CsTensor), ("b_inds", CsTensor), ("c_inds", CsTensor),
        ("split_k_slices", C.c_int32),
        ("alpha", C.c_float), ("beta", C.c_float), ("act_alpha", C.c_float), ("act_beta", C.c_float),
        ("act_type", C.c_int32), ("stream", C.c_uint64),
    ]


class CsConvParams(C.Structure):
    _fields_ = [
        ("conv_algo_desp", CsConvAlgoDesp),
        ("input", CsTensor), ("weight", CsTensor), ("output", CsTensor),
        ("output_add", CsTensor), ("bias", CsTensor), ("scale", CsTensor),
        ("workspace", CsTensor), ("mask", CsTensor), ("mask_argsort", CsTensor),
        ("indices", CsTensor), ("mask_output", CsTensor),
        ("split_k_slices", C.c_int32),
        ("alpha", C.c_float), ("beta", C.c_float), ("act_alpha", C.c_float), ("act_beta", C.c_float),
        ("act_type", C.c_int32), ("mask_width", C.c_int32), ("mask_filter", C.c_uint32),
        ("reverse_mask", C.c_int32), ("verbose", C.c_int32), ("mask_int_count", C.c_int32),
        ("stream", C.c_uint64),
    ]


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load libcumm_amd.so once; raise if it has not been built.
    ``CUMM_AMD_LIBRARY`` names another build of the same library (the
    host-sanitizer build of ``_build.build_sanitized``)."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ["CUMM_AMD_LIBRARY"]) if os.environ.get("CUMM_AMD_LIBRARY") else LIB_PATH
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it with cumm_amd._build.build() "
                           "(hipcc --offload-arch=gfx950)")
    L = C.CDLL(str(path))
    L.cs_abi_version.restype = C.c_int
    L.cs_implicit_gemm2.argtypes = [C.POINTER(CsConvParams)]
    L.cs_implicit_gemm2.restype = C.c_int
    L.cs_implicit_gemm2_noalloc.argtypes = [C.POINTER(CsConvParams)]
    L.cs_implicit_gemm2_noalloc.restype = C.c_int
    L.cs_matmul2.argtypes = [C.POINTER(CsGemmParams)]
    L.cs_matmul2.restype = C.c_int
    L.cs_conv_workspace_size.argtypes = [C.POINTER(CsConvParams), C.POINTER(C.c_int64)]
    L.cs_conv_workspace_size.restype = C.c_int
    L.cs_gemm_workspace_size.argtypes = [C.POINTER(CsGemmParams), C.POINTER(C.c_int64)]
    L.cs_gemm_workspace_size.restype = C.c_int
    L.cs_get_all_conv_algo_desp.argtypes = [C.POINTER(CsConvAlgoDesp), C.c_int]
    L.cs_get_all_conv_algo_desp.restype = C.c_int
    L.cs_get_all_algo_desp.argtypes = [C.POINTER(CsGemmAlgoDesp), C.c_int]
    L.cs_get_all_algo_desp.restype = C.c_int
    L.cs_gemm_algo_name.argtypes = [C.POINTER(CsGemmAlgoDesp), C.c_char_p, C.c_int]
    L.cs_gemm_algo_name.restype = C.c_int
    L.cs_conv_algo_name.argtypes = [C.POINTER(CsConvAlgoDesp), C.c_char_p, C.c_int]
    L.cs_conv_algo_name.restype = C.c_int
    L.cs_device_synchronize.restype = C.c_int
    L.cs_stream_synchronize.argtypes = [C.c_uint64]
    L.cs_stream_synchronize.restype = C.c_int
    L.cs_last_error.restype = C.c_char_p
    P, I64, I32, U64 = C.c_void_p, C.c_int64, C.c_int32, C.c_uint64
    L.cs_pairs_subm_workspace.argtypes = [I64, C.POINTER(I64)]
    L.cs_pairs_subm.argtypes = [P, I64, P, P, P, P, P, P, I64, U64]
    L.cs_pairs_strided_out_workspace.argtypes = [I64, I32, C.POINTER(I64)]
    L.cs_pairs_strided_out.argtypes = [P, I64, P, I32, P, I64, C.POINTER(I64), P, I64, U64]
    L.cs_pairs_strided_workspace.argtypes = [I64, I64, C.POINTER(I64)]
    L.cs_pairs_strided.argtypes = [P, I64, P, P, I64, P, P, P, P, P, P, P, I64, U64]
    G = C.POINTER(CsConvGeom)
    L.cs_pairs_subm_geom_workspace.argtypes = [I64, G, C.POINTER(I64)]
    L.cs_pairs_subm_geom.argtypes = [P, I64, P, G, P, P, P, P, P, I64, U64]
    L.cs_pairs_strided_out_geom_workspace.argtypes = [I64, G, I32, C.POINTER(I64)]
    L.cs_pairs_strided_out_geom.argtypes = [P, I64, P, G, I32, P, I64, C.POINTER(I64), P, I64, U64]
    L.cs_pairs_strided_geom_workspace.argtypes = [I64, I64, G, C.POINTER(I64)]
    L.cs_pairs_strided_geom.argtypes = [P, I64, P, G, P, I64, P, P, P, P, P, P, P, I64, U64]
    L.cs_pairs_row_masks.argtypes = [P, I64, I32, P, P, U64]
    for name in ("cs_pairs_subm_workspace", "cs_pairs_subm", "cs_pairs_strided_out_workspace",
                 "cs_pairs_strided_out", "cs_pairs_strided_workspace", "cs_pairs_strided",
                 "cs_pairs_subm_geom_workspace", "cs_pairs_subm_geom", "cs_pairs_strided_out_geom_workspace",
                 "cs_pairs_strided_out_geom", "cs_pairs_strided_geom_workspace", "cs_pairs_strided_geom",
                 "cs_pairs_row_masks"):
        getattr(L, name).restype = C.c_int
    L.cs_pairs_last_error.restype = C.c_char_p
    L.cs_get_all_gather_params.argtypes = [P, C.c_int]
    L.cs_get_all_scatter_params.argtypes = [P, C.c_int]
    L.cs_gather_supported.argtypes = [I32, I32, I32]
    L.cs_scatter_supported.argtypes = [I64, I64, I64, I64, I32, I32]
    T = C.POINTER(CsTensor)
    L.cs_gather.argtypes = [T, T, T, I64, I64, I64, I64, U64]
    L.cs_scatter.argtypes = [T, T, T, I64, I64, I64, I64, U64]
    for name in ("cs_get_all_gather_params", "cs_get_all_scatter_params", "cs_gather_supported",
                 "cs_scatter_supported", "cs_gather", "cs_scatter"):
        getattr(L, name).restype = C.c_int
    if L.cs_abi_version() != ABI_VERSION:
        raise RuntimeError("libcumm_amd.so ABI version mismatch")
    _lib = L
    return L


def check(rc: int):
    if rc == 0:
        return
    msg = lib().cs_last_error().decode(errors="replace")
    if rc == 1:
        raise ValueError(msg)
    raise RuntimeError(msg)


# -------------------------------------------------------------- conversions
def to_cs_tensor(t: tv.Tensor) -> CsTensor:
    """The borrowed view as a cs_tensor.  The struct is cached on the
    tv.Tensor and rebuilt whenever the torch view's pointer, shape or strides
    differ from the cached ones (a torch tensor's dtype and device never
    change in place).  Callers copy the returned struct and never modify it."""
    if t is None or t.empty():
        return _EMPTY
    tt = t.torch()
    key = (tt.data_ptr(), tt.shape, tt.stride())
    hit = t._abi
    if hit is not None and hit[0] == key:
        return hit[1]
    ct = _make_cs_tensor(t, tt)
    t._abi = (key, ct)
    return ct


_EMPTY = CsTensor()


def _make_cs_tensor(t: tv.Tensor, tt) -> CsTensor:
    ct = CsTensor()
    if tt.dim() > MAX_DIM:
        raise ValueError(f"rank {tt.dim()} > {MAX_DIM}")
    ct.data = tt.data_ptr() if tt.numel() > 0 else None
    if tt.numel() == 0:
        # an allocated tensor with zero elements is not "empty" in tensorview;
        # keep a non-null sentinel so shape checks still apply
        ct.data = C.c_void_p(16)
    ct.dtype = t.dtype
    ct.ndim = tt.dim()
    for i in range(tt.dim()):
        ct.shape[i] = tt.shape[i]
        ct.stride[i] = tt.stride(i)
    ct.device = t.device
    return ct


def _desp_lists(d):
    # list fields can change in place without passing through __setattr__
    return (tuple(d.tile_shape), tuple(d.warp_tile_shape), tuple(d.tensorop), tuple(d.min_arch))


def desp_to_cs(d) -> CsGemmAlgoDesp:
    """GemmAlgoDesp -> struct, cached on the desp until a field changes."""
    key = _desp_lists(d)
    hit = getattr(d, "_abi_gemm", None)
    if hit is not None and hit[0] == key:
        return hit[1]
    c = _make_desp(d)
    object.__setattr__(d, "_abi_gemm", (key, c))
    return c


def _make_desp(d) -> CsGemmAlgoDesp:
    c = CsGemmAlgoDesp()
    c.dtype_a, c.dtype_b, c.dtype_c = int(d.dtype_a), int(d.dtype_b), int(d.dtype_c)
    c.trans_a, c.trans_b, c.trans_c = int(d.trans_a_), int(d.trans_b_), int(d.trans_c_)
    for i in range(3):
        c.tile_shape[i] = int(d.tile_shape[i])
        c.warp_tile_shape[i] = int(d.warp_tile_shape[i])
        c.tensorop[i] = int(d.tensorop[i])
    c.num_stage = int(d.num_stage)
    c.dacc, c.dcomp = int(d.dacc), int(d.dcomp)
    c.algo = d.algo.encode()[:ALGO_LEN - 1]
    c.split_k_serial, c.split_k_parallel = int(d.split_k_serial_), int(d.split_k_parallel_)
    c.shuffle_type = int(d.shuffle_type)
    c.element_per_access_a = int(d.element_per_access_a)
    c.element_per_access_b = int(d.element_per_access_b)
    c.element_per_access_c = int(d.element_per_access_c)
    c.access_per_vector = int(d.access_per_vector)
    c.is_nvrtc = int(bool(d.is_nvrtc))
    c.min_arch[0], c.min_arch[1] = int(d.min_arch[0]), int(d.min_arch[1])
    return c


def conv_desp_to_cs(d) -> CsConvAlgoDesp:
    """ConvAlgoDesp -> struct, cached on the desp until a field changes."""
    key = _desp_lists(d)
    hit = getattr(d, "_abi_conv", None)
    if hit is not None and hit[0] == key:
        return hit[1]
    c = _make_conv_desp(d)
    object.__setattr__(d, "_abi_conv", (key, c))
    return c


def _make_conv_desp(d) -> CsConvAlgoDesp:
    c = CsConvAlgoDesp()
    c.gemm = _make_desp(d)
    c.ndim, c.op_type, c.iter_algo = int(d.ndim), int(d.op_type), int(d.iter_algo)
    c.layout_i, c.layout_w, c.layout_o = int(d.layout_i), int(d.layout_w), int(d.layout_o)
    c.interleave_i, c.interleave_w, c.interleave_o = int(d.interleave_i), int(d.interleave_w), int(d.interleave_o)
    c.mask_sparse = int(bool(d.mask_sparse))
    c.increment_k_first = int(bool(d.increment_k_first))
    c.dynamic_mask = int(bool(d.dynamic_mask))
    c.is_int8_inference = int(bool(d.is_int8_inference))
    return c


def cs_to_desp(c: CsGemmAlgoDesp, d=None):
    d = d if d is not None else tv.gemm.GemmAlgoDesp()
    d.dtype_a, d.dtype_b, d.dtype_c = c.dtype_a, c.dtype_b, c.dtype_c
    d.trans_a_, d.trans_b_, d.trans_c_ = c.trans_a, c.trans_b, c.trans_c
    d.tile_shape = list(c.tile_shape)
    d.warp_tile_shape = list(c.warp_tile_shape)
    d.num_stage = c.num_stage
    d.dacc, d.dcomp = c.dacc, c.dcomp
    d.algo = c.algo.decode()
    d.tensorop = list(c.tensorop)
    d.split_k_serial_, d.split_k_parallel_ = c.split_k_serial, c.split_k_parallel
    d.shuffle_type = tv.gemm.ShuffleStrideType(c.shuffle_type)
    d.element_per_access_a = c.element_per_access_a
    d.element_per_access_b = c.element_per_access_b
    d.element_per_access_c = c.element_per_access_c
    d.access_per_vector = c.access_per_vector
    d.is_nvrtc = bool(c.is_nvrtc)
    d.min_arch = (c.min_arch[0], c.min_arch[1])
    return d


def cs_to_conv_desp(c: CsConvAlgoDesp):
    d = tv.gemm.ConvAlgoDesp(c.ndim, tv.gemm.ConvOpType(c.op_type))
    cs_to_desp(c.gemm, d)
    d.iter_algo = tv.gemm.ConvIterAlgo(c.iter_algo)
    d.layout_i = tv.gemm.ConvLayoutType(c.layout_i)
    d.layout_w = tv.gemm.ConvLayoutType(c.layout_w)
    d.layout_o = tv.gemm.ConvLayoutType(c.layout_o)
    d.interleave_i, d.interleave_w, d.interleave_o = c.interleave_i, c.interleave_w, c.interleave_o
    d.mask_sparse = bool(c.mask_sparse)
    d.increment_k_first = bool(c.increment_k_first)
    d.dynamic_mask = bool(c.dynamic_mask)
    d.is_int8_inference = bool(c.is_int8_inference)
    return d


def conv_params_to_cs(p) -> CsConvParams:
    c = CsConvParams()
    c.conv_algo_desp = conv_desp_to_cs(p.conv_algo_desp)
    for name in ("input", "weight", "output", "output_add", "bias", "scale", "workspace",
                 "mask", "mask_argsort", "indices", "mask_output"):
        setattr(c, name, to_cs_tensor(getattr(p, name)))
    c.split_k_slices = int(p.split_k_slices)
    c.alpha, c.beta = float(p.alpha), float(p.beta)
    c.act_alpha, c.act_beta = float(p.act_alpha), float(p.act_beta)
    c.act_type = int(p.act_type)
    c.mask_width = int(p.mask_width)
    c.mask_filter = int(p.mask_filter) & 0xFFFFFFFF
    c.reverse_mask = int(bool(p.reverse_mask))
    c.verbose = int(bool(p.verbose))
    c.mask_int_count = int(p.mask_int_count)
    c.stream = int(p.stream)
    return c


def gemm_params_to_cs(p) -> CsGemmParams:
    c = CsGemmParams()
    c.algo_desp = desp_to_cs(p.algo_desp)
    for name in ("a", "b", "c", "d", "workspace", "a_inds", "b_inds", "c_inds"):
        setattr(c, name, to_cs_tensor(getattr(p, name)))
    c.split_k_slices = int(p.split_k_slices)
    c.alpha, c.beta = float(p.alpha), float(p.beta)
    c.act_alpha, c.act_beta = float(p.act_alpha), float(p.act_beta)
    c.act_type = int(p.act_type)
    c.stream = int(p.stream)
    return c


def algo_name_gemm(c: CsGemmAlgoDesp) -> str:
    buf = C.create_string_buffer(512)
    n = lib().cs_gemm_algo_name(C.byref(c), buf, 512)
    if n < 0:
        check(-n)
    return buf.value.decode()


def algo_name_conv(c: CsConvAlgoDesp) -> str:
    buf = C.create_string_buffer(512)
    n = lib().cs_conv_algo_name(C.byref(c), buf, 512)
    if n < 0:
        check(-n)
    return buf.value.decode()
