"""Shared helpers for the sparse-conv tests: build a layer's host inputs, run
ConvMainUnitTest.implicit_gemm2 through the C ABI, quantise like the device."""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np
import torch

from cumm_amd import pointcloud as pc
from cumm_amd import tensorview as tv
from cumm_amd.conv.main import ConvMainUnitTest, select_conv_algo
from cumm_amd.tensorview.gemm import ConvOpType, ConvParams

DT = {"f32": tv.float32, "f16": tv.float16, "bf16": tv.bfloat16, "e4m3": tv.float_e4m3}
TORCH_DT = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16,
            "e4m3": torch.float8_e4m3fn}
ALL = 0xFFFFFFFF


def quantize(x: np.ndarray, dt: str) -> np.ndarray:
    """Round an FP64/FP32 array to ``dt`` (RNE, via torch) and return FP64."""
    t = torch.from_numpy(np.asarray(x, dtype=np.float32)).to(TORCH_DT[dt])
    return t.to(torch.float64).numpy()


def to_dev(x: np.ndarray, dt: str) -> torch.Tensor:
    return torch.from_numpy(np.asarray(x, dtype=np.float32)).to(TORCH_DT[dt]).cuda()


def idx_dev(x: np.ndarray) -> torch.Tensor:
    x = np.ascontiguousarray(x)
    if x.dtype == np.uint32:
        return torch.from_numpy(x.view(np.int32)).cuda()
    return torch.from_numpy(x.astype(np.int32)).cuda()


@dataclasses.dataclass
class SubmCase:
    prob: pc.SubmProblem
    C: int
    K: int
    X: np.ndarray        # [N, C] fp32
    W: np.ndarray        # [K, kv, C] fp32
    dO: np.ndarray       # [N, K] fp32


def subm_case(n: int, C: int, K: int, seed: int = 0, grid=(16, 96, 96)) -> SubmCase:
    coords = pc.plane_scene(n, grid=grid, seed=seed)
    prob = pc.build_subm(coords, grid)
    rng = np.random.default_rng(seed + 1000)
    X = rng.standard_normal((n, C)).astype(np.float32)
    W = (rng.standard_normal((K, prob.kv, C)) / np.sqrt(prob.kv * C)).astype(np.float32)
    dO = rng.standard_normal((n, K)).astype(np.float32)
    return SubmCase(prob, C, K, X, W, dO)


def conv_params(op: int, desp, inp: torch.Tensor, weight: torch.Tensor, out: torch.Tensor,
                indices: np.ndarray, mask: np.ndarray, argsort: np.ndarray,
                mask_output: Optional[torch.Tensor] = None, mask_width: int = -1,
                mask_filter: int = ALL, reverse_mask: bool = False, alpha: float = 1.0,
                beta: float = 0.0, bias: Optional[torch.Tensor] = None,
                scale: Optional[torch.Tensor] = None, split_k: int = 1, act_type: int = 0,
                act_alpha: float = 0.0, stream: int = 0) -> ConvParams:
    p = ConvParams(3, ConvOpType(op))
    p.conv_algo_desp = desp
    p.input = tv.from_torch(inp)
    p.weight = tv.from_torch(weight)
    p.output = tv.from_torch(out)
    p.indices = tv.from_torch(idx_dev(indices))
    p.mask = tv.from_torch(idx_dev(mask))
    p.mask_argsort = tv.from_torch(idx_dev(argsort))
    if mask_output is not None:
        p.mask_output = tv.from_torch(mask_output)
    p.mask_width = mask_width
    p.mask_filter = mask_filter
    p.reverse_mask = reverse_mask
    p.alpha = alpha
    p.beta = beta
    if bias is not None:
        p.bias = tv.from_torch(bias)
    if scale is not None:
        p.scale = tv.from_torch(scale)
    p.split_k_slices = split_k
    p.act_type = act_type
    p.act_alpha = act_alpha
    p.stream = stream
    return p


def desp_for(op: int, dt: str, C: int, K: int, out_dt: Optional[str] = None, dynamic: bool = False):
    o = DT[out_dt or dt]
    if op == 0:
        return select_conv_algo(0, DT[dt], DT[dt], o, C, K, dynamic)
    if op == 1:
        return select_conv_algo(1, o, DT[dt], DT[dt], C, K, dynamic)
    return select_conv_algo(2, DT[dt], o, DT[dt], C, K, dynamic)


def run_fwd(case: SubmCase, dt: str, out_dt: Optional[str] = None, **kw):
    """Forward through the boundary; returns (out fp64 numpy, mask_output numpy)."""
    prob = case.prob
    desp = desp_for(0, dt, case.C, case.K, out_dt)
    od = out_dt or dt
    x = to_dev(case.X, dt)
    w = to_dev(case.W, dt)
    out = torch.zeros((prob.n, case.K), dtype=TORCH_DT[od], device="cuda")
    tiles = (prob.n + desp.tile_shape[0] - 1) // desp.tile_shape[0]
    mo = torch.zeros(max(tiles, 1), dtype=torch.int32, device="cuda")
    p = conv_params(0, desp, x, w, out, prob.pair_fwd, prob.mask_sorted, prob.argsort, mask_output=mo, **kw)
    ConvMainUnitTest.implicit_gemm2(p)
    torch.cuda.synchronize()
    return out.double().cpu().numpy(), mo.cpu().numpy().view(np.uint32)[:tiles], desp
