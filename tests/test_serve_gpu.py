"""The continuous-batching server on the GPU.

* every frame's output is the same alone (batch of one) and inside a mixed
  batch: bit-identical rows and output coordinates (stated tolerance: 0);
* admission never reserves more than the HBM budget or the slot count, and a
  frame larger than the budget is refused;
* an FP32 server matches the FP64 oracle chained over the host pair tables
  (relative error <= 1e-5);
* a Poisson trace completes every frame and reports latency percentiles.
"""
import numpy as np
import pytest
import torch

from oracle import spconv_ref as ref

from cumm_amd import batch as B
from cumm_amd import pointcloud as pc
from cumm_amd.serve import Frame, SparseConvServer

pytestmark = pytest.mark.gpu
GRID = (16, 128, 128)
KINDS = ("subm", "strided", "subm", "strided")
CH = (16, 32, 64, 64, 128)


def _weights(seed=0):
    rng = np.random.default_rng(seed)
    return [rng.normal(0, 1 / np.sqrt(27 * CH[i]), (CH[i + 1], 27, CH[i])).astype(np.float32) for i in range(4)]


def _frames(sizes, seed=0):
    out = []
    rng = np.random.default_rng(seed + 100)
    for i, n in enumerate(sizes):
        c = pc.plane_scene(n, grid=GRID, seed=seed + i)
        out.append(Frame(i, c, rng.standard_normal((n, CH[0])).astype(np.float32)))
    return out


def _server(**kw):
    return SparseConvServer(KINDS, CH, _weights(), grid=GRID, **kw)


def test_alone_equals_mix():
    print("[serve] alone vs mix", flush=True)
    sizes = [1500, 400, 3000, 900, 2200]
    alone = []
    s1 = _server(dt="bf16", max_batch=1)
    for f in _frames(sizes):
        s1.submit(f)
        s1.drain()
        alone.append(f)
    s2 = _server(dt="bf16", max_batch=8, n_streams=1)
    mixed = _frames(sizes)
    for f in mixed:
        s2.submit(f)
    s2.drain()
    torch.cuda.synchronize()
    assert max(f.batch_size for f in mixed) == len(sizes)
    for a, m in zip(alone, mixed):
        assert a.batch_size == 1
        assert torch.equal(a.out_coords[:, 1:], m.out_coords[:, 1:])
        assert torch.equal(a.out, m.out), f"frame {a.fid}"


def test_admission_respects_budget():
    print("[serve] admission", flush=True)
    frames = _frames([2000] * 6, seed=3)
    srv = _server(dt="bf16", max_batch=8, hbm_budget=0, max_inflight=3)
    need = srv.frame_bytes(2000)
    srv.budget = 2 * need + need // 2  # room for two frames
    with pytest.raises(ValueError):  # larger than the whole budget
        srv.submit(Frame(99, np.zeros((10 ** 6, 4), np.int32), np.zeros((1, CH[0]), np.float32)))
    for f in frames:
        srv.submit(f)
    peak = 0
    while srv.waiting or srv.admitted or any(b is not None for b in srv.busy.values()):
        srv.pump()
        assert srv.reserved <= srv.budget
        assert srv.inflight <= 2
        peak = max(peak, srv.inflight)
    assert peak == 2
    assert len(srv.completed) == 6


def test_fp32_server_matches_oracle():
    print("[serve] fp32 vs oracle", flush=True)
    f = _frames([1200], seed=7)[0]
    srv = _server(dt="f32", max_batch=1)
    srv.submit(f)
    srv.drain()
    torch.cuda.synchronize()
    tabs = B.build_stack_tables(KINDS, f.coords, GRID)
    x = f.feats.astype(np.float64)
    for t, w in zip(tabs, _weights()):
        x, _ = ref.conv_forward(x, w, t.pair_fwd, t.mask_fwd, t.argsort_fwd, 16)
    assert np.array_equal(f.out_coords.cpu().numpy()[:, 1:], tabs[-1].out_coords[:, 1:])
    assert ref.relative_error(f.out.double().cpu().numpy(), x)[1] <= 1e-5


def test_poisson_trace_completes():
    print("[serve] trace", flush=True)
    frames = _frames([800, 2500, 1200, 3000, 600, 1800, 2200, 900], seed=11)
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1 / 400.0, len(frames)))
    srv = _server(dt="bf16", keep_outputs=False)
    st = srv.run_trace(frames, arrivals)
    assert st["frames"] == len(frames) and all(f.done >= f.arrival for f in frames)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    print(st)


def test_two_streams_overlap_and_match_alone():
    """Two launch threads / streams: every batch of a trace arriving faster
    than one stream drains runs on either stream, and each frame's output
    still equals its batch-of-one result bit for bit."""
    print("[serve] two streams", flush=True)
    sizes = [1500, 400, 3000, 900, 2200, 1700, 2600, 800]
    alone = []
    s1 = _server(dt="bf16", max_batch=1)
    for f in _frames(sizes, seed=21):
        s1.submit(f)
        s1.drain()
        alone.append(f)
    s1.close()
    s2 = _server(dt="bf16", max_batch=3, n_streams=2)
    used = set()
    launch = s2._launch

    def spy(si):
        used.add(si)
        launch(si)

    s2._launch = spy
    frames = _frames(sizes, seed=21)
    arrivals = np.arange(len(frames)) * 2e-4
    st = s2.run_trace(frames, arrivals)
    s2.close()
    torch.cuda.synchronize()
    assert st["frames"] == len(frames)
    assert used == {0, 1}
    for a, m in zip(alone, sorted(frames, key=lambda f: f.fid)):
        assert torch.equal(a.out_coords[:, 1:], m.out_coords[:, 1:])
        assert torch.equal(a.out, m.out), f"frame {a.fid}"
