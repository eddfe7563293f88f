"""``tv.gemm``: parameter and kernel-descriptor classes of the boundary.

Field names, defaults and behaviour follow
include/tensorview/gemm/core/params.h:24-428 and their pybind definitions
(cumm/tensorview_bind.py:174-472).  Enum values follow
include/tensorview/gemm/core/constants.h:20-54.
"""
from __future__ import annotations

import copy as _copy
import enum
from typing import List, Optional, Tuple

from . import Tensor, CUDAKernelTimer, dtype_short_str, unknown


class ConvOpType(enum.IntEnum):
    Forward = 0
    BackwardInput = 1
    BackwardWeight = 2


class ConvMode(enum.IntEnum):
    Convolution = 0
    CrossCorrelation = 1


class ConvIterAlgo(enum.IntEnum):
    Analytic = 0
    Optimized = 1


class ConvLayoutType(enum.IntEnum):
    ChannelFirst = 0
    ChannelLast = 1
    SpatialFirst = 2


class ShuffleStrideType(enum.IntEnum):
    NoShuffle = 0
    ShuffleAC = 1
    ShuffleAB = 2


class Activation(enum.IntEnum):
    None_ = 0
    ReLU = 1
    Sigmoid = 2
    LeakyReLU = 3


class NVRTCParams:
    """Placeholder for the reference's runtime-compiled kernel handle
    (cumm/tensorview_bind.py:217-235).  This build is ahead-of-time compiled;
    a non-empty ``cumodule`` is rejected."""

    def __init__(self):
        self.cumodule = None
        self.kernel_name = ""
        self.init_kernel_name = ""
        self.constant_name = ""
        self.param_size = -1
        self.param_storage = Tensor()
        self.param_storage_cpu = Tensor()
        self.num_threads = -1
        self.smem_size = -1
        self.mode = 0


class GemmAlgoDesp:
    """tv::gemm::GemmAlgoDesp (params.h:24-164)."""

    def __setattr__(self, name, value):
        # a field change drops the C-ABI structs the binding cached for this desp
        object.__setattr__(self, "_abi_gemm", None)
        object.__setattr__(self, "_abi_conv", None)
        object.__setattr__(self, name, value)

    def __deepcopy__(self, memo):
        # fields are scalars, enums, tuples and short int lists; the cached
        # C-ABI structs are not copied (the copy rebuilds its own)
        c = self.__class__.__new__(self.__class__)
        d = {k: (list(v) if isinstance(v, list) else v) for k, v in self.__dict__.items()}
        d["_abi_gemm"] = None
        d["_abi_conv"] = None
        c.__dict__.update(d)
        memo[id(self)] = c
        return c

    def __init__(self):
        self.dtype_a = unknown
        self.dtype_b = unknown
        self.dtype_c = unknown
        self.trans_a_ = -1
        self.trans_b_ = -1
        self.trans_c_ = -1
        self.tile_shape = [-1, -1, -1]
        self.warp_tile_shape = [-1, -1, -1]
        self.num_stage = -1
        self.dacc = unknown
        self.dcomp = unknown
        self.algo = ""
        self.tensorop = [-1, -1, -1]
        self.split_k_serial_ = 0
        self.split_k_parallel_ = 0
        self.shuffle_type = ShuffleStrideType.NoShuffle
        self.element_per_access_a = -1
        self.element_per_access_b = -1
        self.element_per_access_c = -1
        self.access_per_vector = 1
        self.is_nvrtc = False
        self.min_arch = (0, 0)

    # bool views of the int flags (pybind def_property)
    trans_a = property(lambda s: s.trans_a_ == 1, lambda s, v: setattr(s, "trans_a_", 1 if v else 0))
    trans_b = property(lambda s: s.trans_b_ == 1, lambda s, v: setattr(s, "trans_b_", 1 if v else 0))
    trans_c = property(lambda s: s.trans_c_ == 1, lambda s, v: setattr(s, "trans_c_", 1 if v else 0))
    split_k_serial = property(lambda s: s.split_k_serial_ == 1,
                              lambda s, v: setattr(s, "split_k_serial_", 1 if v else 0))
    split_k_parallel = property(lambda s: s.split_k_parallel_ == 1,
                                lambda s, v: setattr(s, "split_k_parallel_", 1 if v else 0))

    def check_valid(self):
        """GemmAlgoDesp::check_valid (params.h:84-108); RuntimeError like TV_ASSERT_RT_ERR."""
        if self.trans_a_ == -1 or self.trans_b_ == -1 or self.trans_c_ == -1 or not self.algo:
            raise RuntimeError("trans_a, trans_b, trans_c and algo must be set")
        for i in range(3):
            if not (self.tile_shape[i] > 0 and self.warp_tile_shape[i] > 0):
                raise RuntimeError(f"tile_shape and warp_tile_shape must be set, but they are "
                                   f"{self.tile_shape} {self.warp_tile_shape}")
        if self.algo not in ("Simt", "SimtDP4A", "SimtDP2A"):
            for i in range(3):
                if not self.tensorop[i] > 0:
                    raise RuntimeError(f"tensorop must be set, but they are {self.tensorop}")
        if unknown in (self.dtype_a, self.dtype_b, self.dtype_c):
            raise RuntimeError("dacc and dcomp must be set to valid value")
        if unknown in (self.dacc, self.dcomp):
            raise RuntimeError("dacc and dcomp must be set to valid value")
        if not self.num_stage > 0:
            raise RuntimeError("num_stage must larger than zero")

    def __repr__(self) -> str:
        """Kernel name grammar of params.h:56-78."""
        self.check_valid()
        s = (f"{self.algo}_{dtype_short_str(self.dtype_a)}{dtype_short_str(self.dtype_b)}"
             f"{dtype_short_str(self.dtype_c)}{dtype_short_str(self.dacc)}{dtype_short_str(self.dcomp)}")
        s += ("n" if self.trans_a else "t") + ("n" if self.trans_b else "t") + ("n" if self.trans_c else "t")
        ts, ws = self.tile_shape, self.warp_tile_shape
        s += f"_m{ts[0]}n{ts[1]}k{ts[2]}m{ws[0]}n{ws[1]}k{ws[2]}A{self.access_per_vector}"
        if self.tensorop[0] != -1:
            s += f"T{self.tensorop[0]}{self.tensorop[1]}{self.tensorop[2]}"
        s += f"_{self.num_stage}{1 if self.split_k_serial else 0}{1 if self.split_k_parallel else 0}"
        if int(self.shuffle_type) != 0:
            s += f"_S{int(self.shuffle_type)}"
        return s

    def copy(self):
        return _copy.deepcopy(self)

    def query_workspace_size(self, m: int, n: int, k: int, split_k_slices: int) -> int:
        """params.h:116-132 (split-K workspaces)."""
        if split_k_slices <= 1:
            return 0
        if self.split_k_serial:
            return 4 * -(-m // self.tile_shape[0]) * -(-n // self.tile_shape[1])
        if self.split_k_parallel:
            return 4 * m * n * split_k_slices
        raise ValueError("not impemented")

    def supported(self, m: int, n: int, k: int) -> bool:
        """params.h:133-149."""
        lda = m if self.trans_a else k
        ldb = k if self.trans_b else n
        ldc = m if self.trans_c else n
        res = True
        if self.element_per_access_a > 0:
            res &= lda % self.element_per_access_a == 0
        if self.element_per_access_b > 0:
            res &= ldb % self.element_per_access_b == 0
        if self.element_per_access_c > 0:
            res &= ldc % self.element_per_access_c == 0
        return res

    def supported_ldx(self, lda: int, ldb: int, ldc: int) -> bool:
        """params.h:150-163."""
        res = True
        if self.element_per_access_a > 0:
            res &= lda % self.element_per_access_a == 0
        if self.element_per_access_b > 0:
            res &= ldb % self.element_per_access_b == 0
        if self.element_per_access_c > 0:
            res &= ldc % self.element_per_access_c == 0
        return res


class ConvAlgoDesp(GemmAlgoDesp):
    """tv::gemm::ConvAlgoDesp (params.h:166-327)."""

    def __init__(self, ndim: int, op_type: ConvOpType):
        super().__init__()
        self.ndim = ndim
        self.op_type = ConvOpType(op_type)
        self.iter_algo = ConvIterAlgo.Optimized
        self.layout_i = ConvLayoutType.ChannelLast
        self.layout_w = ConvLayoutType.ChannelLast
        self.layout_o = ConvLayoutType.ChannelLast
        self.interleave_i = 1
        self.interleave_w = 1
        self.interleave_o = 1
        self.mask_sparse = False
        self.increment_k_first = False
        self.dynamic_mask = False
        self.is_int8_inference = False

    @property
    def conv2gemm_inds(self) -> List[int]:
        return ConvAlgoDesp.conv_iwo_012_to_abc(self.op_type)

    @property
    def gemm2conv_inds(self) -> List[int]:
        return ConvAlgoDesp.gemm_abc_012_to_iwo(self.op_type)

    @staticmethod
    def conv_iwo_012_to_abc(op_type) -> List[int]:
        """params.h:225-237."""
        return {0: [0, 1, 2], 1: [2, 1, 0], 2: [1, 2, 0]}[int(op_type)]

    @staticmethod
    def gemm_abc_012_to_iwo(op_type) -> List[int]:
        """params.h:238-250."""
        return {0: [0, 1, 2], 1: [2, 1, 0], 2: [2, 0, 1]}[int(op_type)]

    @property
    def dtype_input(self) -> int:
        return [self.dtype_a, self.dtype_b, self.dtype_c][self.conv2gemm_inds[0]]

    @property
    def dtype_weight(self) -> int:
        return [self.dtype_a, self.dtype_b, self.dtype_c][self.conv2gemm_inds[1]]

    @property
    def dtype_output(self) -> int:
        return [self.dtype_a, self.dtype_b, self.dtype_c][self.conv2gemm_inds[2]]

    def __repr__(self) -> str:
        """params.h:191-224."""
        s = GemmAlgoDesp.__repr__(self)
        s += f"_C{self.ndim}{int(self.op_type)}{int(self.iter_algo)}"

        def lay(layout, il):
            r = "F" if int(layout) == 0 else "L"
            return r + (str(il) if il > 1 else "")

        s += lay(self.layout_i, self.interleave_i) + lay(self.layout_w, self.interleave_w) + \
            lay(self.layout_o, self.interleave_o)
        if self.mask_sparse:
            s += "_" + ("SK" if self.increment_k_first else "SF")
        if self.dynamic_mask:
            s += "D"
        if self.is_int8_inference:
            s += "_S8"
        return s

    def supported(self, m: int, n: int, k: int, C: int, K: int, mask_width: int) -> bool:
        """params.h:266-285."""
        res = GemmAlgoDesp.supported(self, m, n, k)
        if self.mask_sparse:
            if self.op_type == ConvOpType.Forward:
                res &= C % self.element_per_access_a == 0
            elif self.op_type == ConvOpType.BackwardInput:
                res &= K % self.element_per_access_a == 0
            else:
                res &= mask_width % self.tile_shape[2] == 0
                res &= K % self.element_per_access_a == 0
                res &= C % self.element_per_access_b == 0
        return res

    def query_conv_workspace_size(self, m: int, n: int, k: int, split_k_slices: int, kv: int) -> int:
        """params.h:286-306, reference formula.  The byte count this build's
        kernels need for given tensors is ``ConvMainUnitTest.query_workspace_size``."""
        if not self.mask_sparse:
            return self.query_workspace_size(m, n, k, split_k_slices)
        if split_k_slices <= 1:
            return 0
        if self.split_k_serial:
            if self.op_type == ConvOpType.BackwardWeight:
                grid_n = -(-(n // kv) // self.tile_shape[1]) * kv
            else:
                grid_n = -(-n // self.tile_shape[1])
            return 4 * -(-m // self.tile_shape[0]) * grid_n
        if self.split_k_parallel:
            return 4 * m * n * split_k_slices
        raise ValueError("not impemented")

    def supported_ldx_conv(self, ldi: int, ldw: int, ldo: int) -> bool:
        """params.h:308-326."""
        epas = [self.element_per_access_a, self.element_per_access_b, self.element_per_access_c]
        inds = self.conv2gemm_inds
        res = True
        for ld, ind in zip((ldi, ldw, ldo), inds):
            if epas[ind] > 0:
                res &= ld % epas[ind] == 0
        return res


class GemmParams:
    """tv::gemm::GemmParams (params.h:329-390).  Setting ``a``/``b``/``c`` also
    sets the desp dtypes; ``d`` must match ``c``'s dtype."""

    def __init__(self, timer: Optional[CUDAKernelTimer] = None):
        self.algo_desp = GemmAlgoDesp()
        self._a = Tensor()
        self._b = Tensor()
        self._c = Tensor()
        self._d = Tensor()
        self.split_k_slices = 1
        self.workspace = Tensor()
        self.a_inds = Tensor()
        self.b_inds = Tensor()
        self.c_inds = Tensor()
        self.alpha = 1.0
        self.beta = 0.0
        self.act_alpha = 0.0
        self.act_beta = 0.0
        self.act_type = Activation.None_
        self.stream = 0
        self.timer = timer if timer is not None else CUDAKernelTimer(False)
        self.nvrtc_params = NVRTCParams()

    @property
    def a(self) -> Tensor:
        return self._a

    @a.setter
    def a(self, val: Tensor):
        self._a = val
        self.algo_desp.dtype_a = val.dtype

    @property
    def b(self) -> Tensor:
        return self._b

    @b.setter
    def b(self, val: Tensor):
        self._b = val
        self.algo_desp.dtype_b = val.dtype

    @property
    def c(self) -> Tensor:
        return self._c

    @c.setter
    def c(self, val: Tensor):
        self._c = val
        self.algo_desp.dtype_c = val.dtype

    @property
    def d(self) -> Tensor:
        return self._d

    @d.setter
    def d(self, val: Tensor):
        if self._c.dtype != val.dtype:
            raise RuntimeError("d dtype must equal to c")
        self._d = val

    def check_valid(self):
        """params.h:351-364."""
        self.algo_desp.check_valid()
        if self._a.empty() or self._b.empty() or self._c.empty():
            raise RuntimeError("a,b,c must not empty")
        if self.algo_desp.shuffle_type == ShuffleStrideType.ShuffleAC:
            if self.c_inds.empty():
                raise RuntimeError("a_inds,c_inds tensor must not empty")
        elif self.algo_desp.shuffle_type == ShuffleStrideType.ShuffleAB:
            if self.a_inds.empty() or self.b_inds.empty():
                raise RuntimeError("a_inds,b_inds tensor must not empty")


class ConvParams:
    """tv::gemm::ConvParams (params.h:392-428).  ``act_alpha``/``act_beta`` are
    uninitialised in C++; here they start at 0."""

    def __init__(self, ndim: int, op_type: ConvOpType, timer: Optional[CUDAKernelTimer] = None):
        self.conv_algo_desp = ConvAlgoDesp(ndim, op_type)
        self.input = Tensor()
        self.weight = Tensor()
        self.output = Tensor()
        self.output_add = Tensor()
        self.bias = Tensor()
        self.scale = Tensor()
        self.split_k_slices = 1
        self.padding: List[int] = []
        self.stride: List[int] = []
        self.dilation: List[int] = []
        self.alpha = 1.0
        self.beta = 0.0
        self.act_alpha = 0.0
        self.act_beta = 0.0
        self.act_type = Activation.None_
        self.mask_width = -1
        self.mask_filter = 0xFFFFFFFF
        self.reverse_mask = False
        self.verbose = False
        self.mask_int_count = 1
        self.timer = timer if timer is not None else CUDAKernelTimer(False)
        self.workspace = Tensor()
        self.mask = Tensor()
        self.mask_argsort = Tensor()
        self.indices = Tensor()
        self.mask_output = Tensor()
        self.stream = 0
        self.nvrtc_params = NVRTCParams()
