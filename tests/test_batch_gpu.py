"""Batched stacks, the autotuner and HIP-graph capture on the GPU.

* grouped (all clouds in one launch per layer and op) vs per-cloud launches:
  forward and dgrad rows bit-identical, wgrad equal to the per-cloud sum within
  FP32 rounding (1e-5 relative);
* a captured step replayed several times, with eager work and synchronisation
  between replays, equals the eager step bit for bit every time;
* the tuner picks a catalogue kernel that gives the oracle's result and leaves
  the layer's accumulating output as it was.
"""
import numpy as np
import pytest
import torch

from oracle import spconv_ref as ref

from cumm_amd import batch as B
from cumm_amd import pointcloud as pc
from cumm_amd.tune import ConvTuner

pytestmark = pytest.mark.gpu
GRID = (12, 96, 96)
CH = (16, 32, 64)


@pytest.fixture(scope="module")
def clouds():
    sizes = [900, 2500, 400]
    cl = [pc.plane_scene(s, grid=GRID, seed=40 + b, batch_index=b) for b, s in enumerate(sizes)]
    coords = np.concatenate(cl)
    tabs = B.build_stack_tables(("strided", "subm"), coords, GRID)
    per = [B.build_stack_tables(("strided", "subm"), c, GRID) for c in cl]
    rng = np.random.default_rng(9)
    Ws = [rng.normal(0, 1 / np.sqrt(27 * CH[i]), (CH[i + 1], 27, CH[i])).astype(np.float32) for i in range(2)]
    X = rng.standard_normal((coords.shape[0], CH[0])).astype(np.float32)
    dO = rng.standard_normal((tabs[-1].n_out, CH[-1])).astype(np.float32)
    return cl, tabs, per, Ws, X, dO


def _stack(tabs, Ws, X, dO, dt="f16"):
    st = B.ConvStack(tabs, CH, dt, Ws, torch.cuda.current_stream().cuda_stream)
    st.bind(torch.from_numpy(X).cuda().to(B.TORCH_DT[dt]), torch.from_numpy(dO).cuda().to(B.TORCH_DT[dt]))
    return st


def test_grouped_equals_per_cloud(clouds):
    print("[batch] grouped vs per-cloud", flush=True)
    cl, tabs, per, Ws, X, dO = clouds
    g = _stack(tabs, Ws, X, dO)
    g.step()
    io = np.cumsum([0] + [c.shape[0] for c in cl])
    oo = np.cumsum([0] + [t[-1].n_out for t in per])
    mo = np.cumsum([0] + [t[0].n_out for t in per])
    dw_sum = [0, 0]
    for b in range(len(cl)):
        s = _stack(per[b], Ws, X[io[b]:io[b + 1]], dO[oo[b]:oo[b + 1]])
        s.step()
        torch.cuda.synchronize()
        assert torch.equal(g.layers[-1].out[oo[b]:oo[b + 1]], s.layers[-1].out)
        assert torch.equal(g.layers[0].out[mo[b]:mo[b + 1]], s.layers[0].out)
        assert torch.equal(g.layers[1].din[mo[b]:mo[b + 1]], s.layers[1].din)
        for i in range(2):
            dw_sum[i] = dw_sum[i] + s.layers[i].dw.double()
    for i in range(2):
        rel = float((g.layers[i].dw.double() - dw_sum[i]).norm() / dw_sum[i].norm())
        assert rel < 1e-3, rel  # f16 outputs of different partial sums


def test_grouped_matches_oracle(clouds):
    print("[batch] grouped vs oracle", flush=True)
    cl, tabs, per, Ws, X, dO = clouds
    g = _stack(tabs, Ws, X, dO, "f32")
    g.step()
    torch.cuda.synchronize()
    t0, t1 = tabs
    y0, tm0 = ref.conv_forward(X, Ws[0], t0.pair_fwd, t0.mask_fwd, t0.argsort_fwd, 16)
    y1, tm1 = ref.conv_forward(y0, Ws[1], t1.pair_fwd, t1.mask_fwd, t1.argsort_fwd, 16)
    assert ref.relative_error(g.layers[-1].out.double().cpu().numpy(), y1)[1] < 5e-6
    d1 = ref.conv_dgrad(dO, Ws[1], t1.pair_bwd, t1.mask_fwd, t1.argsort_fwd, 16, reverse_mask=True)
    assert ref.relative_error(g.layers[1].din.double().cpu().numpy(), d1)[1] < 5e-6
    w1 = ref.conv_wgrad(y0, dO, t1.pair_fwd, tm1, t1.argsort_fwd, 16)
    assert ref.relative_error(g.layers[1].dw.double().cpu().numpy(), w1)[1] < 5e-6
    w0 = ref.conv_wgrad(X, d1, t0.pair_fwd, tm0, t0.argsort_fwd, 16)
    assert ref.relative_error(g.layers[0].dw.double().cpu().numpy(), w0)[1] < 5e-6


def test_graph_replay_equals_eager(clouds):
    print("[batch] graph replay", flush=True)
    cl, tabs, per, Ws, X, dO = clouds
    g = _stack(tabs, Ws, X, dO)
    g.step()
    torch.cuda.synchronize()
    ref_out = [t.clone() for t in (g.layers[0].out, g.layers[1].out, g.layers[1].din, g.layers[0].dw,
                                   g.layers[1].dw)]
    g.capture()
    for r in range(3):
        for t in (g.layers[0].out, g.layers[1].out, g.layers[1].din, g.layers[0].dw, g.layers[1].dw):
            t.zero_()
        junk = torch.randn(1 << 20, device="cuda").sum()  # unrelated eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = (g.layers[0].out, g.layers[1].out, g.layers[1].din, g.layers[0].dw, g.layers[1].dw)
        for a, b in zip(got, ref_out):
            assert torch.equal(a, b), f"replay {r}"
        del junk


def test_tuner_restores_and_matches(clouds):
    print("[batch] tuner", flush=True)
    cl, tabs, per, Ws, X, dO = clouds
    g = _stack(tabs, Ws, X, dO, "f32")
    lay = g.layers[1]
    p = lay.params["fwd"]
    p.beta = 1.0  # accumulate into the output: tuning must leave it untouched
    init = torch.randn_like(lay.out)
    lay.out.copy_(init)
    tuner = ConvTuner(warmup=1, iters=2)
    res = tuner.tune(p)
    torch.cuda.synchronize()
    assert torch.equal(lay.out, init)
    assert len(res.tried) >= 1
    lay.run("fwd")
    torch.cuda.synchronize()
    t1 = tabs[1]
    exp, _ = ref.conv_forward(lay.x.double().cpu().numpy(), Ws[1], t1.pair_fwd, t1.mask_fwd, t1.argsort_fwd, 16,
                              beta=1.0, out_init=init.double().cpu().numpy())
    assert ref.relative_error(lay.out.double().cpu().numpy(), exp)[1] < 5e-6
    wl = g.layers[0]
    wl.run("fwd")
    r2 = tuner.tune(wl.params["wgrad"])
    assert r2.split_k in (16, 32, 64, 128)
