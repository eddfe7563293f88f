"""The batched and serving workloads at their stated sizes, in their 16-bit
dtypes, against the FP64 oracle (the reference's mask-walk semantics,
cumm/conv/main_real_spconv.py:309-403, chained over the host pair tables).

* Batched: 32 point clouds with voxel counts log-uniform in [5k, 150k] on the
  41 x 1600 x 1408 grid, strided 3x3x3 stride-2 32 -> 64 (outputs
  unique(floor(c / 2))) then submanifold 64 -> 128, FP16 with FP32
  accumulation, one grouped launch per layer and op, the step captured into a
  HIP graph and replayed.  Every result of a replay (both forwards, the second
  layer's dgrad, both wgrads) equals the eager step bit for bit, and each is
  compared with the FP64 oracle evaluated on that op's actual FP16 inputs:
  relative error <= 1e-3 (forward, dgrad) and <= 2e-3 (wgrad, FP16 output of
  an FP32 split-K tree over ~10^6-row reductions).
* Serving: frames of 20k - 200k voxels through the 4-layer BF16 stack (subm
  16 -> 32, strided 32 -> 64, subm 64 -> 64, strided 64 -> 128) served alone
  and in one mixed batch: bit-identical rows and coordinates, and each frame's
  output against the FP64 chain with BF16 rounding after every layer,
  relative error <= 8e-3.
"""
import numpy as np
import pytest
import torch

from oracle import spconv_ref as ref

from cumm_amd import batch as B
from cumm_amd import pointcloud as pc
from cumm_amd.serve import Frame, SparseConvServer

pytestmark = pytest.mark.gpu


def _q(x, dt):
    t = torch.from_numpy(np.asarray(x, dtype=np.float32)).to(B.TORCH_DT[dt])
    return t.to(torch.float64).numpy()


def _np64(t):
    return t.double().cpu().numpy()


def _rel(a, b):
    return ref.relative_error(a, b)[1]


@pytest.fixture(scope="module")
def c4():
    sizes = B.cloud_sizes(32, 5000, 150000, 0)
    clouds = B.make_clouds(sizes, pc.GRID_KITTI, seed=0)
    coords = np.concatenate(clouds)
    tabs = B.build_stack_tables(("strided", "subm"), coords, pc.GRID_KITTI)
    chans = (32, 64, 128)
    rng = np.random.default_rng(1)
    Ws = [rng.normal(0.0, 1.0 / np.sqrt(27 * chans[i]), size=(chans[i + 1], 27, chans[i])).astype(np.float32)
          for i in range(2)]
    X = rng.standard_normal((coords.shape[0], chans[0])).astype(np.float32)
    dO = rng.standard_normal((tabs[-1].n_out, chans[-1])).astype(np.float32)
    return sizes, coords, tabs, chans, Ws, X, dO


def test_c4_grouped_graph_f16_vs_oracle(c4):
    sizes, coords, tabs, chans, Ws, X, dO = c4
    print(f"[c4] {len(sizes)} clouds, {coords.shape[0]} voxels, rows {[t.n_out for t in tabs]}, "
          f"pairs {[t.num_pairs for t in tabs]}", flush=True)
    assert len(sizes) == 32 and min(sizes) >= 5000 and max(sizes) <= 150000
    st = B.ConvStack(tabs, chans, "f16", Ws, torch.cuda.current_stream().cuda_stream, train=True)
    st.bind(torch.from_numpy(X).cuda().half(), torch.from_numpy(dO).cuda().half())
    st.step()
    torch.cuda.synchronize()
    L0, L1 = st.layers

    def results():
        return [L0.out, L1.out, L1.din, L0.dw, L1.dw]

    eager = [t.clone() for t in results()]
    st.capture()
    for r in range(2):
        for t in results():
            t.zero_()
        torch.cuda.synchronize()
        st.replay()
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(results(), eager)):
            assert torch.equal(a, b), f"replay {r}: result {i} differs from the eager step"
    t0, t1 = tabs
    x = _np64(L0.x)
    w0 = _q(Ws[0], "f16")
    w1 = _q(Ws[1], "f16")
    y0 = _np64(L0.out)
    y1 = _np64(L1.out)
    d1 = _np64(L1.dout)
    din1 = _np64(L1.din)
    exp0, tm0 = ref.conv_forward(x, w0, t0.pair_fwd, t0.mask_fwd, t0.argsort_fwd, 16)
    e = _rel(y0, _q(exp0, "f16"))
    print(f"[c4] layer 0 forward rel {e:.2e}", flush=True)
    assert e <= 1e-3
    exp1, tm1 = ref.conv_forward(y0, w1, t1.pair_fwd, t1.mask_fwd, t1.argsort_fwd, 16)
    e = _rel(y1, _q(exp1, "f16"))
    print(f"[c4] layer 1 forward rel {e:.2e}", flush=True)
    assert e <= 1e-3
    assert np.array_equal(L1.mask_out.cpu().numpy().view(np.uint32)[:tm1.shape[0]], tm1)
    expd = ref.conv_dgrad(d1, w1, t1.pair_bwd, t1.mask_fwd, t1.argsort_fwd, 16, reverse_mask=True)
    e = _rel(din1, _q(expd, "f16"))
    print(f"[c4] layer 1 dgrad rel {e:.2e}", flush=True)
    assert e <= 1e-3
    expw1 = ref.conv_wgrad(y0, d1, t1.pair_fwd, tm1, t1.argsort_fwd, 16)
    e = _rel(_np64(L1.dw), _q(expw1, "f16"))
    print(f"[c4] layer 1 wgrad rel {e:.2e}", flush=True)
    assert e <= 2e-3
    expw0 = ref.conv_wgrad(x, din1, t0.pair_fwd, tm0, t0.argsort_fwd, 16)
    e = _rel(_np64(L0.dw), _q(expw0, "f16"))
    print(f"[c4] layer 0 wgrad rel {e:.2e}", flush=True)
    assert e <= 2e-3


KINDS = ("subm", "strided", "subm", "strided")
CH = (16, 32, 64, 64, 128)


def test_c5_serving_bf16_frames_vs_oracle():
    sizes = [20000, 64000, 131000, 200000]
    rng = np.random.default_rng(2)
    Ws = [rng.normal(0.0, 1.0 / np.sqrt(27 * CH[i]), size=(CH[i + 1], 27, CH[i])).astype(np.float32)
          for i in range(4)]
    coords = [pc.plane_scene(n, pc.GRID_KITTI, seed=50 + i) for i, n in enumerate(sizes)]
    feats = [rng.standard_normal((n, CH[0])).astype(np.float32) for n in sizes]
    print(f"[c5] frames {sizes}", flush=True)

    def frames():
        return [Frame(i, c, x) for i, (c, x) in enumerate(zip(coords, feats))]

    alone = []
    s1 = SparseConvServer(KINDS, CH, Ws, dt="bf16", max_batch=1, hbm_budget=32 << 30)
    for f in frames():
        s1.submit(f)
        s1.drain()
        alone.append(f)
    s2 = SparseConvServer(KINDS, CH, Ws, dt="bf16", max_batch=8, hbm_budget=32 << 30)
    mixed = frames()
    for f in mixed:
        s2.submit(f)
    s2.drain()
    torch.cuda.synchronize()
    assert max(f.batch_size for f in mixed) == len(sizes)
    wq = [_q(w, "bf16") for w in Ws]
    for a, m, c, x in zip(alone, mixed, coords, feats):
        assert a.batch_size == 1
        assert torch.equal(a.out_coords[:, 1:], m.out_coords[:, 1:])
        assert torch.equal(a.out, m.out), f"frame {a.fid}: alone and mixed differ"
        tabs = B.build_stack_tables(KINDS, c, pc.GRID_KITTI)
        y = _q(x, "bf16")
        for t, w in zip(tabs, wq):
            y, _ = ref.conv_forward(y, w, t.pair_fwd, t.mask_fwd, t.argsort_fwd, 16)
            y = _q(y, "bf16")
        assert np.array_equal(a.out_coords.cpu().numpy()[:, 1:], tabs[-1].out_coords[:, 1:])
        e = _rel(_np64(a.out), y)
        print(f"[c5] frame {a.fid} ({c.shape[0]} voxels -> {y.shape[0]} rows) rel {e:.2e}", flush=True)
        assert e <= 8e-3
