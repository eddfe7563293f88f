"""Device-side pair tables (the caller's side of implicit_gemm2, built on the
GPU so features and tables stay resident), for any kernel geometry up to
kv = 1024: submanifold layers of odd kernel size and dilation, strided layers
of any kernel size, stride and padding.

``subm_tables`` and ``strided_tables`` return ``DeviceTables`` whose fields
match ``batch.LayerTables`` (pair_fwd / pair_bwd int32 [kv, N], sorted masks
[N] or [N, W] for kv > 32, and argsorts), bit-identical to the host builder
in ``pointcloud``: the same neighbour rule, offsets f = (kz*K1 + ky)*K2 + kx,
stable mask sort (multi-word masks by the words read as one number, last
word most significant).  Output coordinates of strided layers come out
sorted by (b, z, y, x), as ``np.unique`` on keys gives them.  A layer is
named by ``layer_spec``: "subm" / "strided" (3x3x3; strided stride 2,
padding 1), or ("subm", ksize[, dilation]) / ("strided", ksize[, stride[,
padding]]).  Kernels: csrc/cs_pairs.hip.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import dataclasses
from typing import Optional, Sequence, Tuple

import torch

from . import _abi

RULES = {"down": 0, "reach": 1}


def _triple(v) -> Tuple[int, int, int]:
    if isinstance(v, (int, float)):
        return (int(v),) * 3
    t = tuple(int(x) for x in v)
    assert len(t) == 3
    return t


@dataclasses.dataclass(frozen=True)
class LayerSpec:
    """One layer's kind and kernel geometry (per z, y, x)."""
    kind: str                                  # "subm" or "strided"
    ksize: Tuple[int, int, int] = (3, 3, 3)
    stride: Tuple[int, int, int] = (2, 2, 2)   # strided layers
    padding: Tuple[int, int, int] = (1, 1, 1)  # strided layers
    dilation: Tuple[int, int, int] = (1, 1, 1)  # submanifold layers

    @property
    def kv(self) -> int:
        return self.ksize[0] * self.ksize[1] * self.ksize[2]

    def cs(self) -> "_abi.CsConvGeom":
        g = _abi.CsConvGeom()
        for i in range(3):
            g.ksize[i], g.stride[i] = self.ksize[i], self.stride[i]
            g.padding[i], g.dilation[i] = self.padding[i], self.dilation[i]
        return g

    def out_shape(self, shape: Sequence[int]) -> Tuple[int, int, int]:
        if self.kind == "subm":
            return tuple(int(v) for v in shape)
        return tuple((int(s) + 2 * p - k) // st + 1 for s, p, k, st in zip(shape, self.padding, self.ksize,
                                                                         self.stride))


def layer_spec(kind) -> LayerSpec:
    """"subm" / "strided", a LayerSpec, or ("subm", ksize[, dilation]) /
    ("strided", ksize[, stride[, padding]])."""
    if isinstance(kind, LayerSpec):
        return kind
    if isinstance(kind, str):
        if kind not in ("subm", "strided"):
            raise ValueError(f"unknown layer kind {kind!r}")
        return LayerSpec(kind)
    name, *rest = kind
    if name == "subm":
        return LayerSpec("subm", _triple(rest[0]) if rest else (3, 3, 3),
                         dilation=_triple(rest[1]) if len(rest) > 1 else (1, 1, 1))
    if name == "strided":
        return LayerSpec("strided", _triple(rest[0]) if rest else (3, 3, 3),
                         _triple(rest[1]) if len(rest) > 1 else (2, 2, 2),
                         _triple(rest[2]) if len(rest) > 2 else (1, 1, 1))
    raise ValueError(f"unknown layer kind {name!r}")


@dataclasses.dataclass
class DeviceTables:
    kind: str
    in_coords: torch.Tensor        # [N_in, 4] int32
    out_coords: torch.Tensor       # [N_out, 4] int32
    in_shape: Tuple[int, int, int]
    out_shape: Tuple[int, int, int]
    pair_fwd: torch.Tensor         # [kv, N_out] int32
    pair_bwd: Optional[torch.Tensor]  # [kv, N_in] int32 (None: forward-only tables)
    mask_fwd: torch.Tensor         # [N_out] (or [N_out, W]) int32 view of uint32 masks, sorted
    argsort_fwd: torch.Tensor      # (or, sort_masks=False: masks in row order, identity argsort)
    mask_bwd: Optional[torch.Tensor]
    argsort_bwd: Optional[torch.Tensor]

    @property
    def n_in(self) -> int:
        return self.in_coords.shape[0]

    @property
    def n_out(self) -> int:
        return self.out_coords.shape[0]

    @property
    def num_pairs(self) -> int:
        return int((self.pair_fwd >= 0).sum().item())


def _check(rc: int) -> None:
    if rc == 0:
        return
    msg = _abi.lib().cs_pairs_last_error().decode(errors="replace")
    raise (ValueError if rc == 1 else RuntimeError)(msg)


def _shape_ptr(shape) -> C.Array:
    return (C.c_int32 * 3)(*[int(v) for v in shape])


def _ptr(t: Optional[torch.Tensor]):
    return t.data_ptr() if t is not None else None


def _ws(nbytes: int, dev) -> torch.Tensor:
    return torch.empty((max(int(nbytes), 1),), dtype=torch.uint8, device=dev)


def _on(stream: Optional[int]):
    """(stream handle, context): every buffer the kernels write, and their
    scratch, is allocated inside the context, i.e. on the stream the kernels
    are queued on.  The caching allocator then reuses a freed scratch block
    only in that stream's order, so freeing it on return is safe while the
    kernels still run; outputs used on another stream need the caller's sync."""
    cur = torch.cuda.current_stream()
    if stream is None or stream == cur.cuda_stream:
        return cur.cuda_stream, contextlib.nullcontext()
    return stream, torch.cuda.stream(torch.cuda.ExternalStream(stream))


def _mask_buf(n: int, kv: int, dev) -> torch.Tensor:
    w = (kv + 31) // 32
    return torch.empty((n,) if w == 1 else (n, w), dtype=torch.int32, device=dev)


def _row_masks(pf: torch.Tensor, ms: torch.Tensor, am: torch.Tensor, s: int) -> None:
    """Masks of pair table pf in row order and the identity argsort (queued
    on stream s)."""
    kv, n = pf.shape
    _check(_abi.lib().cs_pairs_row_masks(pf.data_ptr(), n, kv, ms.data_ptr(), am.data_ptr(), s))


def subm_tables(coords: torch.Tensor, shape: Sequence[int], stream: Optional[int] = None,
                backward: bool = True, ksize=(3, 3, 3), dilation=(1, 1, 1), sort_masks: bool = True) -> DeviceTables:
    """``backward=False`` builds forward-only tables (no pair_bwd), all that
    inference needs.  ``sort_masks=False`` (forward-only tables) leaves the
    masks in row order with the identity as argsort: implicit_gemm2 gives every
    row the same result over unsorted masks (the tile grouping only changes how
    dense the tile masks are), and the mask sort's time is saved."""
    assert sort_masks or not backward, "unsorted masks are for forward-only tables"
    L = _abi.lib()
    spec = LayerSpec("subm", _triple(ksize), dilation=_triple(dilation))
    g = spec.cs()
    coords = coords.contiguous()
    assert coords.dtype == torch.int32 and coords.dim() == 2 and coords.shape[1] == 4 and coords.is_cuda
    n = coords.shape[0]
    kv = spec.kv
    dev = coords.device
    s, ctx = _on(stream)
    with ctx:
        pf = torch.empty((kv, n), dtype=torch.int32, device=dev)
        pb = torch.empty((kv, n), dtype=torch.int32, device=dev) if backward else None
        ms = _mask_buf(n, kv, dev)
        am = torch.empty((n,), dtype=torch.int32, device=dev)
        nb = C.c_int64(0)
        _check(L.cs_pairs_subm_geom_workspace(n, C.byref(g), C.byref(nb)))
        ws = _ws(nb.value, dev)
        _check(L.cs_pairs_subm_geom(coords.data_ptr(), n, _shape_ptr(shape), C.byref(g), pf.data_ptr(), _ptr(pb),
                                    ms.data_ptr() if sort_masks else None, am.data_ptr() if sort_masks else None,
                                    ws.data_ptr(), ws.numel(), s))
        if not sort_masks:
            _row_masks(pf, ms, am, s)
    shp = tuple(int(v) for v in shape)
    return DeviceTables("subm", coords, coords, shp, shp, pf, pb, ms, am, None, None)


def strided_out_coords(coords: torch.Tensor, shape: Sequence[int], rule: str = "down",
                       stream: Optional[int] = None, ksize=(3, 3, 3), stride=(2, 2, 2),
                       padding=(1, 1, 1)) -> Tuple[torch.Tensor, Tuple[int, int, int]]:
    L = _abi.lib()
    spec = LayerSpec("strided", _triple(ksize), _triple(stride), _triple(padding))
    g = spec.cs()
    coords = coords.contiguous()
    n = coords.shape[0]
    dev = coords.device
    s, ctx = _on(stream)
    r = RULES[rule]
    cap = n if r == 0 else spec.kv * n
    with ctx:
        out = torch.empty((max(cap, 1), 4), dtype=torch.int32, device=dev)
        nb = C.c_int64(0)
        _check(L.cs_pairs_strided_out_geom_workspace(n, C.byref(g), r, C.byref(nb)))
        ws = _ws(nb.value, dev)
        n_out = C.c_int64(0)
        _check(L.cs_pairs_strided_out_geom(coords.data_ptr(), n, _shape_ptr(shape), C.byref(g), r, out.data_ptr(),
                                           cap, C.byref(n_out), ws.data_ptr(), ws.numel(), s))
        oc = out[:n_out.value]  # written by a kernel queued on stream s
        if r == 1:
            # the kv*n-row buffer is mostly unused: keep a tight copy, ordered on the same stream
            oc = oc.clone()
    return oc, spec.out_shape(shape)


def strided_tables(coords: torch.Tensor, shape: Sequence[int], rule: str = "down",
                   stream: Optional[int] = None, backward: bool = True, ksize=(3, 3, 3), stride=(2, 2, 2),
                   padding=(1, 1, 1), sort_masks: bool = True) -> DeviceTables:
    """``backward=False`` skips pair_bwd and the backward masks / argsort;
    ``sort_masks=False`` as for subm_tables."""
    assert sort_masks or not backward, "unsorted masks are for forward-only tables"
    L = _abi.lib()
    spec = LayerSpec("strided", _triple(ksize), _triple(stride), _triple(padding))
    g = spec.cs()
    coords = coords.contiguous()
    dev = coords.device
    s, ctx = _on(stream)
    oc, out_shape = strided_out_coords(coords, shape, rule, s, spec.ksize, spec.stride, spec.padding)
    n_in, n_out = coords.shape[0], oc.shape[0]
    kv = spec.kv
    with ctx:
        pf = torch.empty((kv, n_out), dtype=torch.int32, device=dev)
        pb = torch.empty((kv, n_in), dtype=torch.int32, device=dev) if backward else None
        mf = _mask_buf(n_out, kv, dev)
        af = torch.empty((n_out,), dtype=torch.int32, device=dev)
        mb = _mask_buf(n_in, kv, dev) if backward else None
        ab = torch.empty((n_in,), dtype=torch.int32, device=dev) if backward else None
        nb = C.c_int64(0)
        _check(L.cs_pairs_strided_geom_workspace(n_in, n_out, C.byref(g), C.byref(nb)))
        ws = _ws(nb.value, dev)
        _check(L.cs_pairs_strided_geom(coords.data_ptr(), n_in, _shape_ptr(shape), C.byref(g), oc.data_ptr(), n_out,
                                       pf.data_ptr(), _ptr(pb), mf.data_ptr() if sort_masks else None,
                                       af.data_ptr() if sort_masks else None, _ptr(mb), _ptr(ab),
                                       ws.data_ptr(), ws.numel(), s))
        if not sort_masks:
            _row_masks(pf, mf, af, s)
    return DeviceTables("strided", coords, oc, tuple(int(v) for v in shape), out_shape, pf, pb, mf, af, mb, ab)


def stack_tables(kinds: Sequence, coords: torch.Tensor, shape: Sequence[int], rule: str = "down",
                 stream: Optional[int] = None, backward: bool = True, sort_masks: bool = True):
    """Tables of a chain of layers; each entry of ``kinds`` goes through
    ``layer_spec``."""
    out = []
    for k in kinds:
        sp = layer_spec(k)
        if sp.kind == "subm":
            t = subm_tables(coords, shape, stream, backward, sp.ksize, sp.dilation, sort_masks)
        else:
            t = strided_tables(coords, shape, rule, stream, backward, sp.ksize, sp.stride, sp.padding, sort_masks)
        out.append(t)
        coords, shape = t.out_coords, t.out_shape
    return out
