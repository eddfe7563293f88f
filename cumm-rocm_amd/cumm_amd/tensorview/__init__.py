"""``tv``: the tensor face of the boundary, backed by PyTorch-ROCm storage.

Mirrors the parts of ``cumm.tensorview`` that callers of ``matmul2`` /
``implicit_gemm2`` use (cumm/tensorview/__init__.py; typed surface at
cumm/tensorview_bind_anno.pyi:586-928): a ``Tensor`` with ``dim``/``shape``/
``dtype``/``numpy``/``cpu``/``cuda``/``zero_``/``view``/``clone``/
``slice_first_axis``, the ``tv.DType`` codes of include/tensorview/dtypes.h:39-60,
and the constructors ``from_numpy``, ``zeros``, ``empty``, ``full``.  A Tensor
borrows a torch tensor (``from_torch``) exactly like
``torch_tensor_to_tv`` borrows one (cumm/inliner/__init__.py:185-218).
Device ``-1`` is the host, ``>= 0`` a GPU ordinal.
"""
from __future__ import annotations

import time
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

# tv::DType codes (include/tensorview/dtypes.h:39-60)
float32 = 0
int32 = 1
int16 = 2
int8 = 3
float64 = 4
bool_ = 5
uint8 = 6
float16 = 7
int64 = 8
uint16 = 9
uint32 = 10
uint64 = 11
bfloat16 = 12
tf32 = 13
float_e4m3 = 14
float_e5m2 = 15
unknown = -1

_TORCH = {
    float32: torch.float32, int32: torch.int32, int16: torch.int16, int8: torch.int8,
    float64: torch.float64, bool_: torch.bool, uint8: torch.uint8, float16: torch.float16,
    int64: torch.int64, uint16: torch.uint16, uint32: torch.uint32, uint64: torch.uint64,
    bfloat16: torch.bfloat16, float_e4m3: torch.float8_e4m3fn, float_e5m2: torch.float8_e5m2,
}
_FROM_TORCH = {v: k for k, v in _TORCH.items()}
_NP = {
    np.dtype(np.float32): float32, np.dtype(np.int32): int32, np.dtype(np.int16): int16,
    np.dtype(np.int8): int8, np.dtype(np.float64): float64, np.dtype(np.bool_): bool_,
    np.dtype(np.uint8): uint8, np.dtype(np.float16): float16, np.dtype(np.int64): int64,
    np.dtype(np.uint16): uint16, np.dtype(np.uint32): uint32, np.dtype(np.uint64): uint64,
}
_SHORT = {
    bool_: "b1", float32: "f32", int8: "s8", int16: "s16", int32: "s32", float64: "f64",
    int64: "s64", uint8: "u8", uint16: "u16", uint32: "u32", uint64: "u64", float16: "f16",
    float_e4m3: "e4m3", float_e5m2: "e5m2", tf32: "tf32", bfloat16: "bf16",
}
_BITS = {
    bool_: 1, float32: 32, int8: 8, int16: 16, int32: 32, float64: 64, int64: 64, uint8: 8,
    uint16: 16, uint32: 32, uint64: 64, float16: 16, float_e4m3: 8, float_e5m2: 8, tf32: 32,
    bfloat16: 16,
}


def dtype_short_str(dtype: int) -> str:
    """tv::dtype_short_str (include/tensorview/dtypes.h:217-267)."""
    return _SHORT.get(int(dtype), "unknown")


def bit_size(dtype: int) -> int:
    return _BITS[int(dtype)]


def torch_dtype(dtype: int) -> torch.dtype:
    return _TORCH[int(dtype)]


def tv_dtype(dt: Union[torch.dtype, np.dtype, int]) -> int:
    if isinstance(dt, int):
        return dt
    if isinstance(dt, torch.dtype):
        return _FROM_TORCH[dt]
    return _NP[np.dtype(dt)]


def _device_of(t: torch.Tensor) -> int:
    return -1 if t.device.type == "cpu" else (t.device.index if t.device.index is not None else 0)


def _torch_device(device: int) -> torch.device:
    return torch.device("cpu") if device < 0 else torch.device("cuda", device)


class Tensor:
    """A borrowed view of torch storage with tensorview's method names.

    ``Tensor()`` is the empty tensor (``tv::Tensor()``), used for optional
    parameters such as ``workspace`` or ``bias``.
    """

    __slots__ = ("_t", "_abi")

    def __init__(self, t: Optional[torch.Tensor] = None):
        self._t = t
        self._abi = None  # (view key, C-ABI struct) cached by the binding

    # -- construction ------------------------------------------------------
    @staticmethod
    def from_torch(t: torch.Tensor) -> "Tensor":
        if t.dtype not in _FROM_TORCH:
            raise ValueError(f"unsupported dtype {t.dtype}")
        return Tensor(t)

    def torch(self) -> torch.Tensor:
        if self._t is None:
            raise ValueError("empty tensor")
        return self._t

    # -- tensorview API ----------------------------------------------------
    def empty(self) -> bool:
        return self._t is None

    @property
    def shape(self) -> List[int]:
        return [] if self._t is None else list(self._t.shape)

    @property
    def ndim(self) -> int:
        return 0 if self._t is None else self._t.dim()

    @property
    def dtype(self) -> int:
        return unknown if self._t is None else _FROM_TORCH[self._t.dtype]

    @property
    def device(self) -> int:
        return -1 if self._t is None else _device_of(self._t)

    @property
    def size(self) -> int:
        return 0 if self._t is None else self._t.numel()

    @property
    def itemsize(self) -> int:
        return 0 if self._t is None else self._t.element_size()

    @property
    def nbytes(self) -> int:
        return self.size * self.itemsize

    def dim(self, i: int) -> int:
        return int(self._t.shape[i])

    def stride(self, i: Optional[int] = None):
        if i is None:
            return list(self._t.stride())
        return int(self._t.stride(i))

    @property
    def strides(self) -> List[int]:
        return [] if self._t is None else list(self._t.stride())

    def is_contiguous(self) -> bool:
        return self._t is None or self._t.is_contiguous()

    def data_ptr(self) -> int:
        return 0 if self._t is None else self._t.data_ptr()

    def byte_pointer(self) -> int:
        return self.data_ptr()

    def numpy(self) -> np.ndarray:
        t = self._t.detach()
        if t.device.type != "cpu":
            raise ValueError("numpy() requires a host tensor, call cpu() first")
        if t.dtype == torch.bfloat16:
            return t.view(torch.int16).numpy().view(np.uint16)
        if t.dtype in (torch.float8_e4m3fn, torch.float8_e5m2):
            return t.view(torch.uint8).numpy()
        return t.numpy()

    def numpy_view(self) -> np.ndarray:
        return self.numpy()

    def cpu(self, ctx=None) -> "Tensor":
        return Tensor(self._t.cpu())

    def cuda(self, ctx=None) -> "Tensor":
        return Tensor(self._t.cuda())

    def clone(self, pinned: bool = False, use_cpu_copy: bool = False, ctx=None) -> "Tensor":
        return Tensor(self._t.clone())

    def view(self, shape: Sequence[int]) -> "Tensor":
        return Tensor(self._t.view(*shape))

    def zero_(self, ctx=None) -> "Tensor":
        self._t.zero_()
        return self

    def fill_(self, value) -> "Tensor":
        self._t.fill_(value)
        return self

    def slice_first_axis(self, start: int, end: int) -> "Tensor":
        return Tensor(self._t[start:end])

    def __getitem__(self, idx) -> "Tensor":
        return Tensor(self._t[idx])

    def __repr__(self) -> str:
        if self._t is None:
            return "Tensor(empty)"
        return f"Tensor(shape={self.shape}, dtype={dtype_short_str(self.dtype)}, device={self.device})"


def from_torch(t: torch.Tensor) -> Tensor:
    return Tensor.from_torch(t)


def from_numpy(arr: np.ndarray) -> Tensor:
    """Host tensor holding a copy of ``arr`` (uint32 supported)."""
    a = np.ascontiguousarray(arr)
    return Tensor(torch.from_numpy(a.copy()))


def zeros(shape: Sequence[int], dtype: int = float32, device: int = -1) -> Tensor:
    return Tensor(torch.zeros(list(shape), dtype=torch_dtype(dtype), device=_torch_device(device)))


def empty(shape: Sequence[int], dtype: int = float32, device: int = -1) -> Tensor:
    return Tensor(torch.empty(list(shape), dtype=torch_dtype(dtype), device=_torch_device(device)))


def full(shape: Sequence[int], value, dtype: int = float32, device: int = -1) -> Tensor:
    return Tensor(torch.full(list(shape), value, dtype=torch_dtype(dtype), device=_torch_device(device)))


class CUDAKernelTimer:
    """Per-name device-event timer (tv::CUDAKernelTimer,
    include/tensorview/profile/cuda_profiler.h:320-470).  When enabled, every
    ``matmul2`` / ``implicit_gemm2`` call records an event pair on its stream
    under the algo name; durations (ms) are read after a sync."""

    def __init__(self, enable: bool = False):
        self.enable = enable
        self._pairs: Dict[str, List[Tuple[object, object]]] = {}

    def record(self, name: str, stream: int):
        if not self.enable:
            return _NullScope()
        return _EventScope(self, name, stream)

    def get_all_pair_duration(self) -> Dict[str, float]:
        torch.cuda.synchronize()
        return {k: sum(a.elapsed_time(b) for a, b in v) for k, v in self._pairs.items()}


class _NullScope:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


class _EventScope:
    def __init__(self, timer: CUDAKernelTimer, name: str, stream: int):
        self.timer, self.name, self.stream = timer, name, stream

    def __enter__(self):
        s = torch.cuda.ExternalStream(self.stream) if self.stream else torch.cuda.default_stream()
        self.s = s
        self.a = torch.cuda.Event(enable_timing=True)
        self.a.record(s)
        return self

    def __exit__(self, *exc):
        b = torch.cuda.Event(enable_timing=True)
        b.record(self.s)
        self.timer._pairs.setdefault(self.name, []).append((self.a, b))
        return False


from . import gemm  # noqa: E402  (tv.gemm namespace)
