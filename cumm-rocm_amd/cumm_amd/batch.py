"""Batched sparse-conv stacks: many point clouds in one launch per layer and op.

A batch of clouds is one sparse tensor whose coordinates carry the cloud index
in column 0, so pair tables never connect two clouds and every
``implicit_gemm2`` call covers all clouds at once: the forward / dgrad kernel
is persistent (one block per CU, waves take 16-row tiles of the whole batch
round-robin), and the wgrad kernel's pieces span the whole batch.  A
``ConvStack`` keeps features, weights, pair tables, masks, gradients and
workspaces device-resident and prebuilds one ``ConvParams`` per layer and op,
so a training step is a fixed launch sequence that can be captured once into a
HIP graph (``torch.cuda.CUDAGraph``) and replayed.

Per-row results do not depend on which other rows share a tile: each output
row sums its offsets in ascending offset order and a missing neighbour adds an
exact zero, so a cloud's forward and dgrad rows are bit-identical whether it
runs alone or inside a batch.  Wgrad sums rows over pieces that depend on the
batch, so it matches within FP32 rounding.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import pairs
from . import pointcloud as pc
from . import tensorview as tv
from .conv.main import ConvMainUnitTest, select_conv_algo
from .tensorview.gemm import ConvOpType, ConvParams

TORCH_DT = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}
TV_DT = {"f32": tv.float32, "f16": tv.float16, "bf16": tv.bfloat16}


# ----------------------------------------------------------------- clouds
def cloud_sizes(n: int, lo: int, hi: int, seed: int) -> List[int]:
    """``n`` voxel counts log-uniform in [lo, hi] (seeded)."""
    rng = np.random.default_rng(seed)
    return [int(v) for v in np.exp(rng.uniform(np.log(lo), np.log(hi), size=n)).round()]


def make_clouds(sizes: Sequence[int], grid=pc.GRID_KITTI, seed: int = 0) -> List[np.ndarray]:
    """One plane scene per size, cloud index ``b`` in column 0."""
    return [pc.plane_scene(s, grid=grid, seed=seed + 7919 * b, batch_index=b) for b, s in enumerate(sizes)]


# ----------------------------------------------------------------- tables
@dataclasses.dataclass
class LayerTables:
    """Host pair tables of one layer over a (possibly batched) coordinate set."""
    kind: str                      # "subm" or "strided"
    in_coords: np.ndarray
    out_coords: np.ndarray
    in_shape: Tuple[int, int, int]
    out_shape: Tuple[int, int, int]
    pair_fwd: np.ndarray           # [kv, N_out]
    pair_bwd: np.ndarray           # [kv, N_in]
    mask_fwd: np.ndarray           # sorted per-row masks of the output rows
    argsort_fwd: np.ndarray
    mask_bwd: Optional[np.ndarray]  # strided only (subm dgrad reverses the forward masks)
    argsort_bwd: Optional[np.ndarray]

    @property
    def n_in(self) -> int:
        return self.in_coords.shape[0]

    @property
    def n_out(self) -> int:
        return self.out_coords.shape[0]

    @property
    def num_pairs(self) -> int:
        return int(np.count_nonzero(self.pair_fwd >= 0))


def build_tables(kind, coords: np.ndarray, shape: Sequence[int], rule: str = "down") -> LayerTables:
    """Host pair tables of one layer; ``kind`` goes through
    ``pairs.layer_spec`` ("subm" / "strided" are 3x3x3, strided stride 2,
    padding 1).  Strided outputs follow ``rule`` (``down``: unique(floor(c /
    stride)), the downsampling of the batched and serving stacks; ``reach``:
    every reached position)."""
    spec = pairs.layer_spec(kind)
    shape = tuple(int(v) for v in shape)
    if spec.kind == "subm":
        prob = pc.build_subm(coords, shape, spec.ksize, dilation=np.asarray(spec.dilation))
        return LayerTables("subm", coords, coords, shape, shape, prob.pair_fwd, prob.pair_bwd,
                           prob.mask_sorted, prob.argsort, None, None)
    sp = pc.build_strided(coords, shape, spec.ksize, np.asarray(spec.stride), np.asarray(spec.padding), rule=rule)
    return LayerTables("strided", coords, sp.out_coords, shape, sp.out_shape, sp.pair_fwd, sp.pair_bwd,
                       sp.mask_fwd_sorted, sp.argsort_fwd, sp.mask_bwd_sorted, sp.argsort_bwd)


def build_stack_tables(kinds: Sequence, coords: np.ndarray, shape: Sequence[int]) -> List[LayerTables]:
    out = []
    for k in kinds:
        t = build_tables(k, coords, shape)
        out.append(t)
        coords, shape = t.out_coords, t.out_shape
    return out


# ----------------------------------------------------------------- device stack
def _dev_idx(a, dev) -> torch.Tensor:
    """int32 device copy of a host table (device tables pass through)."""
    if isinstance(a, torch.Tensor):
        return a.to(dev).contiguous()
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a.astype(np.int32, copy=False)).to(dev)


class DeviceLayer:
    """One conv layer's device-resident tables, weights, buffers and prebuilt
    params (forward, dgrad, wgrad) for one dtype.  ``t`` is a host
    ``LayerTables`` (copied to the device) or a ``pairs.DeviceTables`` built on
    the device (used in place)."""

    def __init__(self, t: LayerTables, C: int, K: int, dt: str, W, stream: int, dev="cuda", train: bool = True):
        self.t, self.C, self.K, self.dt, self.train = t, C, K, dt, train
        td = TORCH_DT[dt]
        kv = t.pair_fwd.shape[0]
        self.kv = kv
        self.pf = _dev_idx(t.pair_fwd, dev)
        if train and t.pair_bwd is None:
            raise ValueError("training needs backward pair tables (build them with backward=True)")
        self.pb = _dev_idx(t.pair_bwd, dev) if t.pair_bwd is not None else None
        self.mf = _dev_idx(t.mask_fwd, dev)
        self.af = _dev_idx(t.argsort_fwd, dev)
        self.mb = _dev_idx(t.mask_bwd, dev) if t.mask_bwd is not None else None
        self.ab = _dev_idx(t.argsort_bwd, dev) if t.argsort_bwd is not None else None
        if isinstance(W, torch.Tensor):  # device weights already in the layer dtype are shared
            self.w = W if (W.dtype == td and W.device.type == "cuda") else W.to(dev, td)
        else:
            self.w = torch.from_numpy(np.ascontiguousarray(W, dtype=np.float32)).to(dev, td)
        self.out = torch.empty((t.n_out, K), dtype=td, device=dev)
        if train:
            self.dw = torch.zeros((K, kv, C), dtype=td, device=dev)
            self.dout = torch.zeros((t.n_out, K), dtype=td, device=dev)
            self.din = torch.zeros((t.n_in, C), dtype=td, device=dev)
        else:
            self.dw = self.dout = self.din = None
        # kernel volumes above 32 take the dynamic (multi-word) mask kernels
        self.dynamic = kv > 32
        self.mw = (kv + 31) // 32
        # the forward writes every tile's mask; zero-filled only where wgrad reads it
        alloc = torch.zeros if train else torch.empty
        self.mask_out = alloc((((t.n_out + 15) // 16 + 1) * self.mw,), dtype=torch.int32, device=dev)
        self.stream = stream
        self.x = None
        self.params = {}

    def bind(self, x: torch.Tensor, dout: Optional[torch.Tensor] = None) -> None:
        """Input features (and the gradient of the output, when it comes from
        outside the stack); builds the params and their workspaces."""
        self.x = x
        if dout is not None:
            self.dout = dout
        tvdt = TV_DT[self.dt]
        C, K = self.C, self.K

        def mk(op: int) -> ConvParams:
            p = ConvParams(3, ConvOpType(op))
            p.conv_algo_desp = select_conv_algo(op, tvdt, tvdt, tvdt, C, K, self.dynamic)
            p.mask_argsort = tv.from_torch(self.af)
            p.stream = self.stream
            if op == 0:
                p.input, p.weight, p.output = tv.from_torch(self.x), tv.from_torch(self.w), tv.from_torch(self.out)
                p.indices, p.mask = tv.from_torch(self.pf), tv.from_torch(self.mf)
                p.mask_output = tv.from_torch(self.mask_out)
            elif op == 1:
                p.input, p.weight, p.output = tv.from_torch(self.din), tv.from_torch(self.w), tv.from_torch(self.dout)
                p.indices = tv.from_torch(self.pb)
                if self.mb is None:  # subm: reverse the forward masks
                    p.mask = tv.from_torch(self.mf)
                    p.reverse_mask = True
                else:
                    p.mask, p.mask_argsort = tv.from_torch(self.mb), tv.from_torch(self.ab)
            else:
                p.input, p.weight, p.output = tv.from_torch(self.x), tv.from_torch(self.dw), tv.from_torch(self.dout)
                p.indices, p.mask = tv.from_torch(self.pf), tv.from_torch(self.mask_out)
                p.mask_width = 16
            ws = ConvMainUnitTest.query_workspace_size(p)
            if ws > 0:
                p.workspace = tv.from_torch(torch.empty((ws,), dtype=torch.uint8, device=self.out.device))
            return p

        self.params = {"fwd": mk(0), "dgrad": mk(1), "wgrad": mk(2)} if self.train else {"fwd": mk(0)}

    def run(self, op: str) -> None:
        ConvMainUnitTest.implicit_gemm2(self.params[op])

    def set_stream(self, stream: int) -> None:
        self.stream = stream
        for p in self.params.values():
            p.stream = stream


class ConvStack:
    """Layers chained output -> input; ``step`` runs forward through all layers
    then, when training, the backward: wgrad of every layer and dgrad of every
    layer but the first (whose input needs no gradient)."""

    def __init__(self, tables: List[LayerTables], channels: Sequence[int], dt: str, weights: Sequence[np.ndarray],
                 stream: int, train: bool = True):
        self.layers = [DeviceLayer(t, channels[i], channels[i + 1], dt, weights[i], stream, train=train)
                       for i, t in enumerate(tables)]
        self.train = train
        self.graph = None

    def bind(self, x: torch.Tensor, dout: Optional[torch.Tensor] = None) -> None:
        L = self.layers
        # layer i reads layer i-1's output; its output gradient is what layer
        # i+1's dgrad writes (the last layer's comes from outside)
        for i, l in enumerate(L):
            if not self.train:
                l.bind(x if i == 0 else L[i - 1].out)
                continue
            l.bind(x if i == 0 else L[i - 1].out, dout if i == len(L) - 1 else L[i + 1].din)

    def ops(self) -> List[Tuple[DeviceLayer, str]]:
        seq = [(l, "fwd") for l in self.layers]
        if self.train:
            for i in reversed(range(len(self.layers))):
                l = self.layers[i]
                if i > 0:
                    seq.append((l, "dgrad"))
                seq.append((l, "wgrad"))
        return seq

    def step(self) -> None:
        for l, op in self.ops():
            l.run(op)

    def flops(self) -> int:
        f = 0
        for l, op in self.ops():
            f += 2 * l.t.num_pairs * l.C * l.K
        return f

    def tune(self, tuner) -> list:
        res = []
        for l, op in self.ops():
            r = tuner.tune(l.params[op])
            res.append((op, l.C, l.K, repr(r.desp), r.split_k, r.us))
        return res

    def capture(self) -> "torch.cuda.CUDAGraph":
        """Records one step into a HIP graph (warm the kernels up first: the
        first call of each kernel queries the device)."""
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            s = torch.cuda.current_stream().cuda_stream
            old = [l.stream for l in self.layers]
            for l in self.layers:
                l.set_stream(s)
            self.step()
        for l, o in zip(self.layers, old):
            l.set_stream(o)
        self.graph = g
        return g

    def replay(self) -> None:
        self.graph.replay()
