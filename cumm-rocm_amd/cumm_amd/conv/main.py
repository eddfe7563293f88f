"""``ConvMainUnitTest``: the sparse-conv entry point of the boundary.

Same call surface as the reference's generated pybind class
(cumm/conv/main.py:242-1179): ``implicit_gemm2(params)`` runs the mask-sparse
implicit GEMM (forward, dgrad, wgrad) on ``params.stream``; companions
``get_all_conv_algo_desp`` and ``extract_mnk``.  Work is done by hand-written
gfx950 kernels in libcumm_amd.so through the C ABI of include/cumm_amd.h.
"""
from __future__ import annotations

import copy

import ctypes as C
from typing import List

from .. import _abi
from .. import tensorview as tv
from ..tensorview.gemm import ConvAlgoDesp, ConvParams


class ConvMainUnitTest:
    @staticmethod
    def implicit_gemm2(params: ConvParams) -> None:
        """cumm/conv/main.py:483-745.  Asynchronous on ``params.stream``; outputs
        are written in place into ``output`` (forward), ``input`` (dgrad) or
        ``weight`` (wgrad)."""
        if params.nvrtc_params.cumodule is not None:
            raise ValueError("runtime-compiled kernels are not supported; use a desp from "
                             "get_all_conv_algo_desp()")
        cp = _abi.conv_params_to_cs(params)
        L = _abi.lib()
        if params.timer.enable:
            name = repr(params.conv_algo_desp)
            with params.timer.record(name, params.stream):
                _abi.check(L.cs_implicit_gemm2(C.byref(cp)))
        else:
            _abi.check(L.cs_implicit_gemm2(C.byref(cp)))

    @staticmethod
    def implicit_gemm2_noalloc(params: ConvParams) -> None:
        """``implicit_gemm2`` without any device allocation (no counterpart in
        the reference): the workspace is the library's per-GPU scratch array
        (``params.workspace`` is ignored).  Calls that overlap on one GPU race;
        a call needing more than the scratch raises ValueError."""
        cp = _abi.conv_params_to_cs(params)
        _abi.check(_abi.lib().cs_implicit_gemm2_noalloc(C.byref(cp)))

    @staticmethod
    def query_workspace_size(params: ConvParams) -> int:
        """Bytes of workspace this build's kernels need for ``params`` (a caller
        that passes an empty workspace gets one allocated per call)."""
        cp = _abi.conv_params_to_cs(params)
        out = C.c_int64(0)
        _abi.check(_abi.lib().cs_conv_workspace_size(C.byref(cp), C.byref(out)))
        return int(out.value)

    @staticmethod
    def get_all_conv_algo_desp() -> List[ConvAlgoDesp]:
        """cumm/conv/main.py:983-1050."""
        L = _abi.lib()
        n = L.cs_get_all_conv_algo_desp(None, 0)
        arr = (_abi.CsConvAlgoDesp * n)()
        L.cs_get_all_conv_algo_desp(arr, n)
        return [_abi.cs_to_conv_desp(arr[i]) for i in range(n)]

    @staticmethod
    def extract_mnk(op_type: int, N: int, C: int, K: int, kernel_volume: int,
                    in_prod: int, out_prod: int, mask_sparse: bool) -> List[int]:
        """cumm/conv/main.py:467-481 -> implicit_gemm_mnk
        (include/tensorview/gemm/core/utils.h:62-92)."""
        op = int(op_type)
        if mask_sparse:
            table = {0: [N, K, C * kernel_volume], 1: [N, C, K * kernel_volume],
                     2: [K, C * kernel_volume, N]}
        else:
            table = {0: [N * out_prod, K, C * kernel_volume], 1: [N * in_prod, C, K * kernel_volume],
                     2: [K, C * kernel_volume, N * out_prod]}
        return table.get(op, [])


_CATALOGUE: List[ConvAlgoDesp] = []
_SELECTED: dict = {}  # (op, dtypes, dynamic_mask) -> catalogue entry


def select_conv_algo(op_type: int, dtype_input: int, dtype_weight: int, dtype_output: int,
                     C: int, K: int, dynamic_mask: bool = False) -> ConvAlgoDesp:
    """Pick this build's kernel for a conv layer, the way a caller filters
    ``get_all_conv_algo_desp()`` by op, dtypes and mask kind (64 output
    columns per wave); ``dynamic_mask`` selects the multi-word-mask kernels
    that kernel volumes above 32 need.  Returns a fresh copy; the catalogue is
    read from the library once."""
    if not _CATALOGUE:
        _CATALOGUE.extend(ConvMainUnitTest.get_all_conv_algo_desp())
    op = int(op_type)
    key = (op, int(dtype_input), int(dtype_weight), int(dtype_output), bool(dynamic_mask))
    hit = _SELECTED.get(key)
    if hit is not None:
        return copy.deepcopy(hit)
    bn = 64
    best = None
    for d in _CATALOGUE:
        if int(d.op_type) != op or bool(d.dynamic_mask) != bool(dynamic_mask):
            continue
        if (d.dtype_input, d.dtype_weight, d.dtype_output) != (dtype_input, dtype_weight, dtype_output):
            continue
        if op == 2 or d.tile_shape[1] == bn:
            best = d
            break
        best = best or d
    if best is None:
        raise ValueError(f"no kernel for op {op} with dtypes input={tv.dtype_short_str(dtype_input)} "
                         f"weight={tv.dtype_short_str(dtype_weight)} output={tv.dtype_short_str(dtype_output)}")
    _SELECTED[key] = best
    return copy.deepcopy(best)
