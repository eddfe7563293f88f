"""``GemmMainUnitTest``: the gather/scatter GEMM entry point of the boundary.

Same call surface as cumm/gemm/main.py:528-1335: ``matmul2(params)`` for
NoShuffle, ShuffleAC (gather A rows, scatter C rows) and ShuffleAB (gather A
and B rows, reduce over them), plus ``get_all_algo_desp``, ``extract_mnk``,
``align_to_power2``, ``device_synchronize``, ``stream_synchronize`` and
``simple_select_tile_shape``.  Kernels run in libcumm_amd.so (C ABI in
include/cumm_amd.h).
"""
from __future__ import annotations

import bisect
import ctypes as C
from typing import Dict, List, Sequence, Tuple

from .. import _abi
from .. import tensorview as tv
from ..tensorview.gemm import GemmAlgoDesp, GemmParams, ShuffleStrideType


class GemmMainUnitTest:
    @staticmethod
    def matmul2(params: GemmParams) -> None:
        """cumm/gemm/main.py:1176-1335; asynchronous on ``params.stream``."""
        if params.nvrtc_params.cumodule is not None:
            raise ValueError("runtime-compiled kernels are not supported; use a desp from "
                             "get_all_algo_desp()")
        params.check_valid()
        cp = _abi.gemm_params_to_cs(params)
        L = _abi.lib()
        if params.timer.enable:
            with params.timer.record(repr(params.algo_desp), params.stream):
                _abi.check(L.cs_matmul2(C.byref(cp)))
        else:
            _abi.check(L.cs_matmul2(C.byref(cp)))

    @staticmethod
    def query_workspace_size(params: GemmParams) -> int:
        cp = _abi.gemm_params_to_cs(params)
        out = C.c_int64(0)
        _abi.check(_abi.lib().cs_gemm_workspace_size(C.byref(cp), C.byref(out)))
        return int(out.value)

    @staticmethod
    def get_all_algo_desp() -> List[GemmAlgoDesp]:
        """cumm/gemm/main.py:914-965."""
        L = _abi.lib()
        n = L.cs_get_all_algo_desp(None, 0)
        arr = (_abi.CsGemmAlgoDesp * n)()
        L.cs_get_all_algo_desp(arr, n)
        return [_abi.cs_to_desp(arr[i]) for i in range(n)]

    @staticmethod
    def extract_mnk(a_shape: Sequence[int], b_shape: Sequence[int], trans_a: bool, trans_b: bool,
                    trans_c: bool, shuffle_type: int = ShuffleStrideType.NoShuffle,
                    a_inds_shape: Sequence[int] = (), b_inds_shape: Sequence[int] = (),
                    c_inds_shape: Sequence[int] = ()) -> Tuple[int, int, int]:
        """cumm/gemm/main.py:967-1026."""
        a_shape, b_shape = list(a_shape), list(b_shape)
        if trans_c:
            trans_a, trans_b = not trans_a, not trans_b
            trans_a, trans_b = trans_b, trans_a
            a_shape, b_shape = b_shape, a_shape
        st = int(shuffle_type)
        if st == ShuffleStrideType.ShuffleAC:
            if trans_a:
                raise RuntimeError("a of shuffle AB must be row major")
            if len(c_inds_shape) == 0:
                raise RuntimeError("c_inds must not empty")
            m = a_inds_shape[0] if len(a_inds_shape) else a_shape[0]
            k = a_shape[int(not trans_a)]
            n = b_shape[int(not trans_b)]
        elif st == ShuffleStrideType.ShuffleAB:
            if not len(a_inds_shape) or not len(b_inds_shape):
                raise RuntimeError("a_inds and c_inds must not empty")
            if not (trans_a and not trans_b):
                raise RuntimeError("shuffle AB must be nt, i.e. backward weight")
            m = a_shape[int(trans_a)]
            k = a_inds_shape[0]
            n = b_shape[int(not trans_b)]
        else:
            m = a_shape[int(trans_a)]
            k = a_shape[int(not trans_a)]
            n = b_shape[int(not trans_b)]
        return int(m), int(n), int(k)

    @staticmethod
    def align_to_power2(val: int) -> int:
        """cumm/gemm/main.py:1028-1048: smallest power of two >= val (val > 0)."""
        r = 0
        num_1_bit = 1 if val & 1 else 0
        while True:
            val >>= 1
            if not val:
                break
            r += 1
            if val & 1:
                num_1_bit += 1
        return 1 << r if num_1_bit == 1 else 1 << (r + 1)

    @staticmethod
    def device_synchronize() -> None:
        _abi.check(_abi.lib().cs_device_synchronize())

    @staticmethod
    def stream_synchronize(stream: int) -> None:
        _abi.check(_abi.lib().cs_stream_synchronize(int(stream)))

    @staticmethod
    def simple_select_tile_shape(m: int, n: int, k: int, tile_ms: Sequence[int], tile_ns: Sequence[int],
                                 tile_ks: Sequence[int], tile_shape_to_algos: Dict[int, List[int]],
                                 large_k_first: bool) -> List[int]:
        """cumm/gemm/main.py:1072-1134: walk down from the smallest tile >= (m, n, k)
        and return the algo list of the first tile key ``tm | tn << 20 | tk << 40`` present."""
        def start(lst, v):
            i = bisect.bisect_left(lst, v)
            return min(i, len(lst) - 1)

        im, in_, ik = start(tile_ms, m), start(tile_ns, n), start(tile_ks, k)
        if large_k_first:
            order = ((a, b, c) for c in range(ik, -1, -1) for b in range(in_, -1, -1) for a in range(im, -1, -1))
        else:
            order = ((a, b, c) for a in range(im, -1, -1) for b in range(in_, -1, -1) for c in range(ik, -1, -1))
        for a, b, c in order:
            key = int(tile_ms[a]) | (int(tile_ns[b]) << 20) | (int(tile_ks[c]) << 40)
            if key in tile_shape_to_algos:
                return list(tile_shape_to_algos[key])
        return []
