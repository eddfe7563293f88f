"""Autotuner for implicit_gemm2 layers.

For one layer (a fully built ``ConvParams`` on the caller's own device
buffers) the tuner times every catalogue kernel that can run it — same op,
same dtypes, ``supported()`` for its channels — and, for wgrad, a ladder of
split-K piece caps, with HIP events on the layer's stream.  The fastest
(kernel, split) is cached under a shape key (op, dtypes, C, K, kv, log2 of the
row count, mask width), so layers of similar size reuse one measurement.

Outputs the tuning runs overwrite (forward output and mask_output, dgrad
input gradient, wgrad weight gradient) are restored afterwards, so tuning a
layer whose epilogue accumulates (beta != 0) leaves its buffers as they were.

The reference's own selection helper is the heuristic
``simple_select_tile_shape`` (cumm/gemm/main.py:1072-1134); this tuner
measures instead, which is what the batched and serving forms use.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Tuple

import torch

from . import tensorview as tv
from .conv.main import ConvMainUnitTest
from .tensorview.gemm import ConvAlgoDesp, ConvOpType, ConvParams

WGRAD_SPLITS = (16, 32, 64, 128)


@dataclasses.dataclass
class TuneResult:
    desp: ConvAlgoDesp
    split_k: int
    us: float
    tried: List[Tuple[str, int, float]]


def _dims(p: ConvParams) -> Tuple[int, int, int, int]:
    """(rows walked by the kernel, C, K, kv) of a layer."""
    K, kv, C = p.weight.shape
    op = int(p.conv_algo_desp.op_type)
    rows = p.output.shape[0] if op != 1 else p.input.shape[0]
    return rows, C, K, kv


def shape_key(p: ConvParams) -> tuple:
    d = p.conv_algo_desp
    rows, C, K, kv = _dims(p)
    return (int(d.op_type), int(d.dtype_input), int(d.dtype_weight), int(d.dtype_output), C, K, kv,
            int(math.log2(max(rows, 1))), int(p.mask_width))


def candidates(p: ConvParams) -> List[ConvAlgoDesp]:
    d0 = p.conv_algo_desp
    rows, C, K, kv = _dims(p)
    out = []
    for d in ConvMainUnitTest.get_all_conv_algo_desp():
        if int(d.op_type) != int(d0.op_type):
            continue
        if (d.dtype_input, d.dtype_weight, d.dtype_output) != (d0.dtype_input, d0.dtype_weight, d0.dtype_output):
            continue
        # same mask kind and int8 form as the caller's desp (the masks' layout depends on it)
        if bool(d.dynamic_mask) != bool(d0.dynamic_mask) or bool(d.is_int8_inference) != bool(d0.is_int8_inference):
            continue
        m, n, k = ConvMainUnitTest.extract_mnk(int(d.op_type), rows, C, K, kv, -1, -1, True)
        if not d.supported(m, n, k, C, K, p.mask_width):
            continue
        out.append(d)
    return out


class ConvTuner:
    def __init__(self, warmup: int = 2, iters: int = 5):
        self.warmup = warmup
        self.iters = iters
        self.cache: Dict[tuple, TuneResult] = {}

    def _written(self, p: ConvParams) -> List[torch.Tensor]:
        op = int(p.conv_algo_desp.op_type)
        ts = [p.output if op == 0 else p.input if op == 1 else p.weight]
        if op == 0 and not p.mask_output.empty():
            ts.append(p.mask_output)
        return [t.torch() for t in ts]

    def _time(self, p: ConvParams, stream: torch.cuda.Stream) -> float:
        for _ in range(self.warmup):
            ConvMainUnitTest.implicit_gemm2(p)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(self.iters):
            ConvMainUnitTest.implicit_gemm2(p)
        b.record(stream)
        b.synchronize()
        return a.elapsed_time(b) * 1e3 / self.iters

    def tune(self, p: ConvParams, stream: Optional[torch.cuda.Stream] = None) -> TuneResult:
        """Returns (and caches) the fastest kernel / split for ``p``; leaves
        ``p`` set to it, with a workspace of the right size."""
        key = shape_key(p)
        if key in self.cache:
            self.apply(p, self.cache[key])
            return self.cache[key]
        # default: the stream the params already run on; clones, timing and
        # restores are ordered on it, and p.stream is left as it was
        if stream is None:
            stream = torch.cuda.ExternalStream(p.stream) if p.stream else torch.cuda.current_stream()
        old_stream = p.stream
        p.stream = stream.cuda_stream
        with torch.cuda.stream(stream):
            saved = [t.clone() for t in self._written(p)]
        op = int(p.conv_algo_desp.op_type)
        splits = WGRAD_SPLITS if op == ConvOpType.BackwardWeight else (1,)
        best = None
        tried = []
        for d in candidates(p):
            for sk in splits:
                p.conv_algo_desp = d
                p.split_k_slices = sk
                self._workspace(p)
                us = self._time(p, stream)
                tried.append((repr(d), sk, us))
                if best is None or us < best[2]:
                    best = (d, sk, us)
        with torch.cuda.stream(stream):
            for t, s in zip(self._written(p), saved):
                t.copy_(s)
        torch.cuda.synchronize()
        p.stream = old_stream
        res = TuneResult(best[0], best[1], best[2], tried)
        self.cache[key] = res
        self.apply(p, res)
        return res

    @staticmethod
    def _workspace(p: ConvParams) -> None:
        ws = ConvMainUnitTest.query_workspace_size(p)
        have = 0 if p.workspace.empty() else p.workspace.torch().numel()
        if ws > have:
            p.workspace = tv.from_torch(torch.empty((ws,), dtype=torch.uint8, device="cuda"))

    def apply(self, p: ConvParams, r: TuneResult) -> None:
        p.conv_algo_desp = r.desp
        p.split_k_slices = r.split_k
        self._workspace(p)
