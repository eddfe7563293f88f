"""``GatherAll`` / ``ScatterAll``: standalone row gather and scatter-add.

The glue either side of ``matmul2`` in spconv's gather-GEMM-scatter ("native")
algorithm, with the call surface of cumm/gemm/gather.py:391-618:

* ``GatherAll().gather(output, input, indices, tile_m, tile_k_bytes,
  bytes_per_access, num_threads, stream)``: ``output[i] = input[indices[i]]``
  for ``i < indices.dim(0)``, any dtype;
* ``ScatterAll().scatter(...)``: ``output[indices[i]] += input[i]``, float32,
  indices of one call distinct;
* ``get_all_gather_params`` / ``get_all_scatter_params``: the published
  ``(tile_m, tile_k_bytes, bytes_per_access, num_threads)`` sets; ``supported``
  / ``supported_scatter``; ``stream_synchronize``.

The tile parameters must be a published set (RuntimeError "can't find a
sutable gather/scatter" otherwise, like the reference); the HIP kernels in
libcumm_amd.so choose their own launch shape for the row width.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Tuple

from .. import _abi
from .. import tensorview as tv

Param = Tuple[int, int, int, int]


def _params(fn) -> List[Param]:
    L = _abi.lib()
    n = getattr(L, fn)(None, 0)
    buf = (C.c_int32 * (4 * n))()
    getattr(L, fn)(buf, n)
    return [tuple(buf[4 * i:4 * i + 4]) for i in range(n)]


def _sync(stream: int) -> None:
    _abi.check(_abi.lib().cs_stream_synchronize(int(stream)))


class GatherAll:
    """cumm/gemm/gather.py:391-496."""

    @staticmethod
    def get_all_gather_params() -> List[Param]:
        return _params("cs_get_all_gather_params")

    @staticmethod
    def supported(bytes_per_access: int, channel_size: int, dtype: int) -> bool:
        return bool(_abi.lib().cs_gather_supported(int(bytes_per_access), int(channel_size), int(dtype)))

    @staticmethod
    def stream_synchronize(stream: int = 0) -> None:
        _sync(stream)

    def gather(self, output: tv.Tensor, input: tv.Tensor, indices: tv.Tensor, tile_m: int, tile_k_bytes: int,
               bytes_per_access: int, num_threads: int, stream: int = 0) -> None:
        o, i, x = _abi.to_cs_tensor(output), _abi.to_cs_tensor(input), _abi.to_cs_tensor(indices)
        _abi.check(_abi.lib().cs_gather(C.byref(o), C.byref(i), C.byref(x), int(tile_m), int(tile_k_bytes),
                                        int(bytes_per_access), int(num_threads), int(stream)))


class ScatterAll:
    """cumm/gemm/gather.py:499-618."""

    @staticmethod
    def get_all_scatter_params() -> List[Param]:
        return _params("cs_get_all_scatter_params")

    def supported_scatter(self, tile_m: int, tile_k_bytes: int, bytes_per_access: int, num_threads: int,
                          channel_size: int, dtype: int) -> bool:
        return bool(_abi.lib().cs_scatter_supported(int(tile_m), int(tile_k_bytes), int(bytes_per_access),
                                                    int(num_threads), int(channel_size), int(dtype)))

    @staticmethod
    def stream_synchronize(stream: int = 0) -> None:
        _sync(stream)

    def scatter(self, output: tv.Tensor, input: tv.Tensor, indices: tv.Tensor, tile_m: int, tile_k_bytes: int,
                bytes_per_access: int, num_threads: int, stream: int = 0) -> None:
        o, i, x = _abi.to_cs_tensor(output), _abi.to_cs_tensor(input), _abi.to_cs_tensor(indices)
        _abi.check(_abi.lib().cs_scatter(C.byref(o), C.byref(i), C.byref(x), int(tile_m), int(tile_k_bytes),
                                         int(bytes_per_access), int(num_threads), int(stream)))
