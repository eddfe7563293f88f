"""Device-side pair tables for 3x3x3 layers (the caller's side of
implicit_gemm2, built on the GPU so features and tables stay resident).

``subm_tables`` and ``strided_tables`` return ``DeviceTables`` whose fields
match ``batch.LayerTables`` (pair_fwd / pair_bwd int32 [27, N], sorted masks
and argsorts), bit-identical to the host builder in ``pointcloud``: the same
neighbour rule, offsets f = kz*9 + ky*3 + kx, stable mask sort.  Output
coordinates of strided layers come out sorted by (b, z, y, x), as
``np.unique`` on keys gives them.  Kernels: csrc/cs_pairs.hip.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import dataclasses
from typing import Optional, Sequence, Tuple

import torch

from . import _abi

RULES = {"down": 0, "reach": 1}


@dataclasses.dataclass
class DeviceTables:
    kind: str
    in_coords: torch.Tensor        # [N_in, 4] int32
    out_coords: torch.Tensor       # [N_out, 4] int32
    in_shape: Tuple[int, int, int]
    out_shape: Tuple[int, int, int]
    pair_fwd: torch.Tensor         # [27, N_out] int32
    pair_bwd: Optional[torch.Tensor]  # [27, N_in] int32 (None: forward-only tables)
    mask_fwd: torch.Tensor         # [N_out] int32 view of uint32 masks, sorted
    argsort_fwd: torch.Tensor
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


def subm_tables(coords: torch.Tensor, shape: Sequence[int], stream: Optional[int] = None,
                backward: bool = True) -> DeviceTables:
    """``backward=False`` builds forward-only tables (no pair_bwd), all that
    inference needs."""
    L = _abi.lib()
    coords = coords.contiguous()
    assert coords.dtype == torch.int32 and coords.dim() == 2 and coords.shape[1] == 4 and coords.is_cuda
    n = coords.shape[0]
    dev = coords.device
    s, ctx = _on(stream)
    with ctx:
        pf = torch.empty((27, n), dtype=torch.int32, device=dev)
        pb = torch.empty((27, n), dtype=torch.int32, device=dev) if backward else None
        ms = torch.empty((n,), dtype=torch.int32, device=dev)
        am = torch.empty((n,), dtype=torch.int32, device=dev)
        nb = C.c_int64(0)
        _check(L.cs_pairs_subm_workspace(n, C.byref(nb)))
        ws = _ws(nb.value, dev)
        _check(L.cs_pairs_subm(coords.data_ptr(), n, _shape_ptr(shape), pf.data_ptr(), _ptr(pb), ms.data_ptr(),
                               am.data_ptr(), ws.data_ptr(), ws.numel(), s))
    shp = tuple(int(v) for v in shape)
    return DeviceTables("subm", coords, coords, shp, shp, pf, pb, ms, am, None, None)


def strided_out_coords(coords: torch.Tensor, shape: Sequence[int], rule: str = "down",
                       stream: Optional[int] = None) -> Tuple[torch.Tensor, Tuple[int, int, int]]:
    L = _abi.lib()
    coords = coords.contiguous()
    n = coords.shape[0]
    dev = coords.device
    s, ctx = _on(stream)
    r = RULES[rule]
    cap = n if r == 0 else 27 * n
    with ctx:
        out = torch.empty((max(cap, 1), 4), dtype=torch.int32, device=dev)
        nb = C.c_int64(0)
        _check(L.cs_pairs_strided_out_workspace(n, r, C.byref(nb)))
        ws = _ws(nb.value, dev)
        n_out = C.c_int64(0)
        _check(L.cs_pairs_strided_out(coords.data_ptr(), n, _shape_ptr(shape), r, out.data_ptr(), cap,
                                      C.byref(n_out), ws.data_ptr(), ws.numel(), s))
        oc = out[:n_out.value]  # written by a kernel queued on stream s
        if r == 1:
            # the 27n-row buffer is mostly unused: keep a tight copy, ordered on the same stream
            oc = oc.clone()
    out_shape = tuple((int(v) - 1) // 2 + 1 for v in shape)
    return oc, out_shape


def strided_tables(coords: torch.Tensor, shape: Sequence[int], rule: str = "down",
                   stream: Optional[int] = None, backward: bool = True) -> DeviceTables:
    """``backward=False`` skips pair_bwd and the backward masks / argsort."""
    L = _abi.lib()
    coords = coords.contiguous()
    dev = coords.device
    s, ctx = _on(stream)
    oc, out_shape = strided_out_coords(coords, shape, rule, s)
    n_in, n_out = coords.shape[0], oc.shape[0]
    with ctx:
        pf = torch.empty((27, n_out), dtype=torch.int32, device=dev)
        pb = torch.empty((27, n_in), dtype=torch.int32, device=dev) if backward else None
        mf = torch.empty((n_out,), dtype=torch.int32, device=dev)
        af = torch.empty((n_out,), dtype=torch.int32, device=dev)
        mb = torch.empty((n_in,), dtype=torch.int32, device=dev) if backward else None
        ab = torch.empty((n_in,), dtype=torch.int32, device=dev) if backward else None
        nb = C.c_int64(0)
        _check(L.cs_pairs_strided_workspace(n_in, n_out, C.byref(nb)))
        ws = _ws(nb.value, dev)
        _check(L.cs_pairs_strided(coords.data_ptr(), n_in, _shape_ptr(shape), oc.data_ptr(), n_out, pf.data_ptr(),
                                  _ptr(pb), mf.data_ptr(), af.data_ptr(), _ptr(mb), _ptr(ab),
                                  ws.data_ptr(), ws.numel(), s))
    return DeviceTables("strided", coords, oc, tuple(int(v) for v in shape), out_shape, pf, pb, mf, af, mb, ab)


def stack_tables(kinds: Sequence[str], coords: torch.Tensor, shape: Sequence[int], rule: str = "down",
                 stream: Optional[int] = None, backward: bool = True):
    out = []
    for k in kinds:
        t = (subm_tables(coords, shape, stream, backward) if k == "subm"
             else strided_tables(coords, shape, rule, stream, backward))
        out.append(t)
        coords, shape = t.out_coords, t.out_shape
    return out
