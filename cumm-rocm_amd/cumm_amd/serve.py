"""Continuous-batching serving loop for sparse-conv stacks.

Requests are voxel frames (int32 coordinates + float features on the host)
that arrive over time.  The server

* admits a waiting frame when a slot is free and its estimated device memory
  fits the HBM budget still unreserved (frames larger than the whole budget
  are refused at submission);
* forms a batch from every admitted frame not yet started (up to
  ``max_batch`` frames) as soon as a stream is idle, so frames that arrive
  while others are on the GPU join the next batch instead of waiting for the
  queue to drain;
* runs each batch on its own HIP stream from a small pool: upload, device pair
  tables for every layer (cumm_amd.pairs), then one implicit_gemm2 launch per
  layer over all frames of the batch; features, tables and activations stay
  on the device.  Each stream has a launch thread of its own: the host work of
  a batch (staging, table building, whose strided layers wait on their stream
  for the output count, parameter set-up) blocks only that stream's thread,
  so with two streams a batch formed while another is being launched or is
  running starts at once and the two overlap on the device;
* completes frames by polling the batch's event, releasing their memory
  reservation and slot, and records latency = completion - arrival.

A frame's output rows are bit-identical alone and inside any batch: pair
tables never connect two frames (the frame index is coordinate column 0), and
the forward kernels sum each row's offsets in a fixed order whatever rows
share its tile.
"""
from __future__ import annotations

import collections
import dataclasses
import queue
import threading
import time
from typing import Deque, Dict, List, Optional, Sequence

import numpy as np
import torch

from . import batch as B
from . import pairs

GRID = (41, 1600, 1408)


@dataclasses.dataclass
class Frame:
    fid: int
    coords: np.ndarray       # [N, 4] int32, column 0 ignored (set per batch)
    feats: np.ndarray        # [N, C0] float32
    arrival: float = 0.0     # seconds on the server clock
    start: float = 0.0
    done: float = 0.0
    batch_size: int = 0
    out: Optional[torch.Tensor] = None
    out_coords: Optional[torch.Tensor] = None


@dataclasses.dataclass
class _Batch:
    frames: List[Frame]
    stream: torch.cuda.Stream
    event: Optional[torch.cuda.Event]
    keep: list               # device objects alive until the batch completes
    reserved: int
    launched: threading.Event = dataclasses.field(default_factory=threading.Event)
    error: Optional[BaseException] = None


class SparseConvServer:
    def __init__(self, kinds: Sequence, channels: Sequence[int], weights: Sequence[np.ndarray], dt: str = "bf16",
                 hbm_budget: int = 8 << 30, max_inflight: int = 64, max_batch: int = 16, n_streams: int = 1,
                 grid=GRID, keep_outputs: bool = True):
        assert len(kinds) + 1 == len(channels) == len(weights) + 1
        self.kinds, self.channels, self.dt = tuple(kinds), tuple(channels), dt
        self.weights = [torch.from_numpy(np.ascontiguousarray(w, dtype=np.float32)).cuda().to(B.TORCH_DT[dt])
                        for w in weights]
        self.budget = int(hbm_budget)
        self.reserved = 0
        self.max_inflight, self.max_batch = max_inflight, max_batch
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]
        # the weights were converted on the current stream; the pool streams
        # (created non-blocking) must not read them before that finishes
        for st in self.streams:
            st.wait_stream(torch.cuda.current_stream())
        self.busy: Dict[int, Optional[_Batch]] = {i: None for i in range(n_streams)}
        self.grid = grid
        self.keep_outputs = keep_outputs
        self.waiting: Deque[Frame] = collections.deque()
        self.admitted: Deque[Frame] = collections.deque()
        self.inflight = 0
        self.completed: List[Frame] = []
        self.t0 = time.perf_counter()
        # one launch thread per stream (started on first use)
        self._jobs: List["queue.Queue[Optional[_Batch]]"] = [queue.Queue() for _ in range(n_streams)]
        self._workers: List[Optional[threading.Thread]] = [None] * n_streams

    def close(self) -> None:
        """Stops the launch threads (after the work they were given)."""
        for si, w in enumerate(self._workers):
            if w is not None:
                self._jobs[si].put(None)
                w.join()
                self._workers[si] = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ memory model
    def frame_bytes(self, n: int) -> int:
        """Upper estimate of one frame's device footprint: coordinates,
        features and activations of every layer, pair tables and masks of
        every layer (a strided layer's outputs are bounded by its inputs for
        the downsampling rule), pair-builder scratch and conv workspaces."""
        es = 2 if self.dt in ("f16", "bf16") else 4
        per_row = 16 + 4 * self.channels[0] + es * self.channels[0]
        for c in self.channels[1:]:
            per_row += es * c + 16
        for k in self.kinds:
            kv = pairs.layer_spec(k).kv
            per_row += kv * 4 * 2 + 8 * 4 * ((kv + 31) // 32) + 64
        return int(per_row * n * 1.25) + (4 << 20)

    def now(self) -> float:
        return time.perf_counter() - self.t0

    # ------------------------------------------------------------ requests
    def submit(self, f: Frame) -> None:
        if self.frame_bytes(f.coords.shape[0]) > self.budget:
            raise ValueError(f"frame {f.fid} needs more than the whole HBM budget")
        self.waiting.append(f)

    def _admit(self) -> None:
        while self.waiting and self.inflight < self.max_inflight:
            f = self.waiting[0]
            need = self.frame_bytes(f.coords.shape[0])
            if self.reserved + need > self.budget:
                break
            self.waiting.popleft()
            self.reserved += need
            self.inflight += 1
            self.admitted.append(f)

    def _idle_stream(self) -> Optional[int]:
        for i, b in self.busy.items():
            if b is None:
                return i
        return None

    def _launch(self, si: int) -> None:
        """Forms a batch from the admitted frames and hands it to stream si's
        launch thread."""
        frames = []
        while self.admitted and len(frames) < self.max_batch:
            frames.append(self.admitted.popleft())
        now = self.now()
        for f in frames:
            f.start = now
            f.batch_size = len(frames)
        b = _Batch(frames, self.streams[si], None, [], sum(self.frame_bytes(f.coords.shape[0]) for f in frames))
        self.busy[si] = b
        if self._workers[si] is None:
            w = threading.Thread(target=self._worker, args=(si,), daemon=True, name=f"serve-launch-{si}")
            self._workers[si] = w
            w.start()
        self._jobs[si].put(b)

    def _worker(self, si: int) -> None:
        while True:
            b = self._jobs[si].get()
            if b is None:
                return
            try:
                self._enqueue(si, b)
            except BaseException as e:  # reported by the scheduler thread
                b.error = e
            b.launched.set()

    def _enqueue(self, si: int, b: _Batch) -> None:
        """Host work and device launches of one batch (runs on the stream's
        launch thread)."""
        frames = b.frames
        s = self.streams[si]
        n_all = sum(f.coords.shape[0] for f in frames)
        c0 = frames[0].feats.shape[1]
        with torch.cuda.stream(s):
            # each frame's host arrays go up as they are: a copy from pageable
            # memory returns once the runtime has read the source (it measured
            # faster than a host copy into pinned staging first); cloud index
            # k in column 0
            c_dev = torch.empty((n_all, 4), dtype=torch.int32, device="cuda")
            x32 = torch.empty((n_all, c0), dtype=torch.float32, device="cuda")
            off = 0
            for k, f in enumerate(frames):
                n = f.coords.shape[0]
                c_dev[off:off + n].copy_(torch.from_numpy(np.ascontiguousarray(f.coords, dtype=np.int32)),
                                         non_blocking=True)
                c_dev[off:off + n, 0] = k
                x32[off:off + n].copy_(torch.from_numpy(np.ascontiguousarray(f.feats, dtype=np.float32)),
                                      non_blocking=True)
                off += n
            x_dev = x32.to(B.TORCH_DT[self.dt])
            # forward-only tables with masks left in row order: the mask sort
            # would cost more than the denser tiles it saves the convolutions
            tabs = pairs.stack_tables(self.kinds, c_dev, self.grid, stream=s.cuda_stream, backward=False,
                                      sort_masks=False)
            st = B.ConvStack(tabs, self.channels, self.dt, self.weights, s.cuda_stream, train=False)
            st.bind(x_dev)
            st.step()
            ev = torch.cuda.Event()
            ev.record(s)
        b.keep = [c_dev, x_dev, tabs, st]
        b.event = ev

    def _poll(self) -> None:
        for si, b in self.busy.items():
            if b is None or not b.launched.is_set():
                continue
            if b.error is not None:
                raise RuntimeError(f"launching a batch on stream {si} failed") from b.error
            if not b.event.query():
                continue
            t = self.now()
            st, tabs = b.keep[3], b.keep[2]
            out = st.layers[-1].out
            oc = tabs[-1].out_coords
            if self.keep_outputs:
                # output rows of frame k are contiguous: keys sort frame-major
                counts = torch.bincount(oc[:, 0].long(), minlength=len(b.frames)).tolist()
                offs = np.cumsum([0] + counts)
                for k, f in enumerate(b.frames):
                    f.out = out[offs[k]:offs[k + 1]]
                    f.out_coords = oc[offs[k]:offs[k + 1]]
            for f in b.frames:
                f.done = t
                self.completed.append(f)
            self.reserved -= b.reserved
            self.inflight -= len(b.frames)
            self.busy[si] = None

    def pump(self) -> bool:
        """One scheduler pass; True if it did anything."""
        before = (len(self.completed), len(self.admitted), len(self.waiting))
        self._poll()
        self._admit()
        si = self._idle_stream()
        launched = False
        if self.admitted and si is not None:
            self._launch(si)
            launched = True
        return launched or before != (len(self.completed), len(self.admitted), len(self.waiting))

    def drain(self) -> None:
        while self.waiting or self.admitted or any(b is not None for b in self.busy.values()):
            if not self.pump():
                time.sleep(5e-5)

    # ------------------------------------------------------------ trace replay
    def run_trace(self, frames: Sequence[Frame], arrivals: Sequence[float]) -> dict:
        """Replays ``frames`` arriving at ``arrivals`` (seconds from start);
        returns latency / throughput statistics."""
        order = np.argsort(arrivals, kind="stable")
        self.t0 = time.perf_counter()
        i = 0
        n = len(frames)
        while i < n or self.waiting or self.admitted or any(b is not None for b in self.busy.values()):
            t = self.now()
            while i < n and arrivals[order[i]] <= t:
                f = frames[order[i]]
                f.arrival = float(arrivals[order[i]])
                self.submit(f)
                i += 1
            if not self.pump():
                nxt = arrivals[order[i]] - self.now() if i < n else 0.0
                time.sleep(min(max(nxt, 0.0), 2e-4) if nxt > 0 else 5e-5)
        lat = np.array([f.done - f.arrival for f in frames])
        span = max(f.done for f in frames) - min(f.arrival for f in frames)
        return {"frames": n, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
                "max_ms": float(lat.max() * 1e3), "frames_per_s": n / span if span > 0 else float("inf"),
                "mean_batch": float(np.mean([f.batch_size for f in frames]))}
