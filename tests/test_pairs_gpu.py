"""Device pair / mask / argsort generation (csrc/cs_pairs.hip) against the host
builder (cumm_amd.pointcloud, itself pinned by brute force in
test_oracle.py) and the golden tables: everything int32 / uint32 bit-exact."""
from pathlib import Path

import numpy as np
import pytest
import torch

from cumm_amd import batch as B
from cumm_amd import pairs
from cumm_amd import pointcloud as pc

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def _np(t):
    return t.cpu().numpy()


def _coords(sizes, grid, seed):
    return np.concatenate([pc.plane_scene(s, grid=grid, seed=seed + b, batch_index=b) for b, s in enumerate(sizes)])


@pytest.mark.parametrize("sizes,grid", [([1], (4, 8, 8)), ([300], (6, 20, 20)), ([5000, 1200, 800], (16, 96, 96)),
                                        ([20000], pc.GRID_KITTI), ([150000], pc.GRID_KITTI)])
def test_subm_tables_match_host(sizes, grid):
    print(f"[pairs] subm {sizes} {grid}", flush=True)
    c = _coords(sizes, grid, 3)
    host = pc.build_subm(c, grid)
    dev = pairs.subm_tables(torch.from_numpy(c).cuda(), grid)
    torch.cuda.synchronize()
    assert np.array_equal(_np(dev.pair_fwd), host.pair_fwd)
    assert np.array_equal(_np(dev.pair_bwd), host.pair_bwd)
    assert np.array_equal(_np(dev.mask_fwd).view(np.uint32), host.mask_sorted)
    assert np.array_equal(_np(dev.argsort_fwd), host.argsort)


@pytest.mark.parametrize("rule", ["down", "reach"])
@pytest.mark.parametrize("sizes,grid", [([1], (4, 8, 8)), ([300], (6, 20, 20)), ([5000, 1200, 800], (16, 96, 96)),
                                        ([20000], pc.GRID_KITTI)])
def test_strided_tables_match_host(sizes, grid, rule):
    print(f"[pairs] strided {rule} {sizes} {grid}", flush=True)
    c = _coords(sizes, grid, 5)
    host = pc.build_strided(c, grid, rule=rule)
    dev = pairs.strided_tables(torch.from_numpy(c).cuda(), grid, rule=rule)
    torch.cuda.synchronize()
    assert dev.out_shape == tuple(host.out_shape)
    assert np.array_equal(_np(dev.out_coords), host.out_coords)
    assert np.array_equal(_np(dev.pair_fwd), host.pair_fwd)
    assert np.array_equal(_np(dev.pair_bwd), host.pair_bwd)
    assert np.array_equal(_np(dev.mask_fwd).view(np.uint32), host.mask_fwd_sorted)
    assert np.array_equal(_np(dev.argsort_fwd), host.argsort_fwd)
    assert np.array_equal(_np(dev.mask_bwd).view(np.uint32), host.mask_bwd_sorted)
    assert np.array_equal(_np(dev.argsort_bwd), host.argsort_bwd)


def test_golden_tables_on_device():
    g = np.load(GOLD / "subm_tiny.npz")
    dev = pairs.subm_tables(torch.from_numpy(g["coords"]).cuda(), (6, 14, 14))
    assert np.array_equal(_np(dev.pair_fwd), g["pair_fwd"])
    assert np.array_equal(_np(dev.argsort_fwd), g["argsort"])
    g = np.load(GOLD / "strided_tiny.npz")
    dev = pairs.strided_tables(torch.from_numpy(g["in_coords"]).cuda(), (6, 14, 14), rule="reach")
    assert np.array_equal(_np(dev.out_coords), g["out_coords"])
    assert np.array_equal(_np(dev.pair_fwd), g["pair_fwd"])
    assert np.array_equal(_np(dev.pair_bwd), g["pair_bwd"])


def test_device_tables_drive_a_stack():
    """A stack over device-built tables gives the same bits as over host tables."""
    grid = (16, 96, 96)
    c = _coords([3000, 900], grid, 11)
    ch = (16, 32, 64)
    rng = np.random.default_rng(2)
    Ws = [rng.normal(0, 0.1, (ch[i + 1], 27, ch[i])).astype(np.float32) for i in range(2)]
    host = B.build_stack_tables(("strided", "subm"), c, grid)
    dev = pairs.stack_tables(("strided", "subm"), torch.from_numpy(c).cuda(), grid)
    X = rng.standard_normal((c.shape[0], ch[0])).astype(np.float32)
    dO = rng.standard_normal((host[-1].n_out, ch[-1])).astype(np.float32)
    outs = []
    for tabs in (host, dev):
        st = B.ConvStack(tabs, ch, "f16", Ws, torch.cuda.current_stream().cuda_stream)
        st.bind(torch.from_numpy(X).cuda().half(), torch.from_numpy(dO).cuda().half())
        st.step()
        torch.cuda.synchronize()
        outs.append([st.layers[-1].out.clone(), st.layers[1].din.clone(), st.layers[0].dw.clone()])
    for a, b in zip(*outs):
        assert torch.equal(a, b)


@pytest.mark.parametrize("kind", ["subm", "strided"])
def test_forward_only_tables(kind):
    """backward=False (inference) gives the same forward tables and no backward ones."""
    print(f"[pairs] forward-only {kind}", flush=True)
    grid = (16, 96, 96)
    c = torch.from_numpy(_coords([4000, 700], grid, 21)).cuda()
    build = pairs.subm_tables if kind == "subm" else pairs.strided_tables
    full, fwd = build(c, grid), build(c, grid, backward=False)
    torch.cuda.synchronize()
    for name in ("out_coords", "pair_fwd", "mask_fwd", "argsort_fwd"):
        assert torch.equal(getattr(full, name), getattr(fwd, name)), name
    assert fwd.pair_bwd is None and fwd.mask_bwd is None and fwd.argsort_bwd is None
    with pytest.raises(ValueError, match="backward"):
        B.DeviceLayer(fwd, 16, 16, "f16", np.zeros((16, 27, 16), np.float32), 0)


@pytest.mark.parametrize("ksize,dilation", [((5, 5, 5), (1, 1, 1)), ((3, 5, 3), (2, 1, 3)), ((1, 1, 1), (1, 1, 1)),
                                            ((7, 3, 1), (1, 2, 1))])
def test_forward_only_subm_kernel_shapes(ksize, dilation):
    """Forward-only submanifold tables look up half the offsets and fill the
    mirrored ones by symmetry; the result equals the full build's forward table
    and the host builder's for anisotropic and dilated kernels."""
    print(f"[pairs] forward-only subm {ksize} dil {dilation}", flush=True)
    grid = (16, 96, 96)
    c = _coords([3000, 900], grid, 33)
    dc = torch.from_numpy(c).cuda()
    full = pairs.subm_tables(dc, grid, ksize=ksize, dilation=dilation)
    fwd = pairs.subm_tables(dc, grid, backward=False, ksize=ksize, dilation=dilation)
    uns = pairs.subm_tables(dc, grid, backward=False, ksize=ksize, dilation=dilation, sort_masks=False)
    torch.cuda.synchronize()
    for name in ("pair_fwd", "mask_fwd", "argsort_fwd"):
        assert torch.equal(getattr(full, name), getattr(fwd, name)), name
    assert torch.equal(full.pair_fwd, uns.pair_fwd)
    host = pc.build_subm(c, grid, ksize=ksize, dilation=dilation)
    np.testing.assert_array_equal(fwd.pair_fwd.cpu().numpy(), host.pair_fwd)


@pytest.mark.parametrize("kind", ["subm", "strided"])
def test_tables_on_a_side_stream(kind):
    """Tables built on a stream other than the current one: scratch and outputs
    belong to that stream, so allocations made on the current stream right
    after the call (while the kernels may still run) cannot take their memory."""
    print(f"[pairs] side stream {kind}", flush=True)
    grid = (16, 96, 96)
    c = _coords([6000, 3000], grid, 9)
    dc = torch.from_numpy(c).cuda()
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    if kind == "subm":
        dev = pairs.subm_tables(dc, grid, stream=side.cuda_stream)
    else:
        dev = pairs.strided_tables(dc, grid, rule="down", stream=side.cuda_stream)
    # churn the current stream's pool while the side stream may still be busy
    junk = [torch.full((1 << 16,), -7, dtype=torch.int32, device="cuda") for _ in range(32)]
    side.synchronize()
    torch.cuda.synchronize()
    del junk
    if kind == "subm":
        host = pc.build_subm(c, grid)
        assert np.array_equal(_np(dev.pair_fwd), host.pair_fwd)
        assert np.array_equal(_np(dev.pair_bwd), host.pair_bwd)
        assert np.array_equal(_np(dev.mask_fwd).view(np.uint32), host.mask_sorted)
        assert np.array_equal(_np(dev.argsort_fwd), host.argsort)
    else:
        host = pc.build_strided(c, grid, rule="down")
        assert np.array_equal(_np(dev.out_coords), host.out_coords)
        assert np.array_equal(_np(dev.pair_fwd), host.pair_fwd)
        assert np.array_equal(_np(dev.pair_bwd), host.pair_bwd)
        assert np.array_equal(_np(dev.mask_fwd).view(np.uint32), host.mask_fwd_sorted)
        assert np.array_equal(_np(dev.argsort_bwd), host.argsort_bwd)


@pytest.mark.parametrize("ksize,dilation", [((5, 5, 5), 1), ((7, 7, 7), 1), ((3, 3, 3), 2), ((5, 3, 3), (1, 2, 1))])
def test_subm_geometry_tables_match_host(ksize, dilation):
    """Other kernel sizes and dilations; kv > 32 gives [N, W] masks sorted by
    the words as one number (last word most significant)."""
    print(f"[pairs] subm ksize={ksize} dilation={dilation}", flush=True)
    grid = (16, 96, 96)
    c = _coords([3000, 800], grid, 13)
    host = pc.build_subm(c, grid, ksize, dilation=np.asarray(dilation))
    dev = pairs.subm_tables(torch.from_numpy(c).cuda(), grid, ksize=ksize, dilation=dilation)
    torch.cuda.synchronize()
    assert dev.pair_fwd.shape[0] == host.kv
    assert np.array_equal(_np(dev.pair_fwd), host.pair_fwd)
    assert np.array_equal(_np(dev.pair_bwd), host.pair_bwd)
    assert np.array_equal(_np(dev.mask_fwd).view(np.uint32), host.mask_sorted)
    assert np.array_equal(_np(dev.argsort_fwd), host.argsort)


@pytest.mark.parametrize("rule", ["down", "reach"])
@pytest.mark.parametrize("ksize,stride,padding", [((2, 2, 2), 2, 0), ((3, 3, 3), 1, 1), ((5, 5, 5), 2, 2),
                                                  ((3, 3, 3), (2, 2, 1), (1, 1, 0))])
def test_strided_geometry_tables_match_host(ksize, stride, padding, rule):
    print(f"[pairs] strided ksize={ksize} stride={stride} padding={padding} {rule}", flush=True)
    grid = (16, 96, 96)
    c = _coords([3000, 800], grid, 17)
    st, pd = np.broadcast_to(np.asarray(stride), 3), np.broadcast_to(np.asarray(padding), 3)
    host = pc.build_strided(c, grid, ksize, st, pd, rule=rule)
    dev = pairs.strided_tables(torch.from_numpy(c).cuda(), grid, rule=rule, ksize=ksize, stride=stride,
                               padding=padding)
    torch.cuda.synchronize()
    assert dev.out_shape == tuple(host.out_shape)
    assert np.array_equal(_np(dev.out_coords), host.out_coords)
    assert np.array_equal(_np(dev.pair_fwd), host.pair_fwd)
    assert np.array_equal(_np(dev.pair_bwd), host.pair_bwd)
    assert np.array_equal(_np(dev.mask_fwd).view(np.uint32), host.mask_fwd_sorted)
    assert np.array_equal(_np(dev.argsort_fwd), host.argsort_fwd)
    assert np.array_equal(_np(dev.mask_bwd).view(np.uint32), host.mask_bwd_sorted)
    assert np.array_equal(_np(dev.argsort_bwd), host.argsort_bwd)


def test_strided_tables_for_a_given_output_list():
    """cs_pairs_strided_geom over an output list the caller chose (a shuffled
    subset of the outputs plus one row outside the output grid): listed outputs
    get the host builder's pairs, unlisted outputs leave their inputs' pair_bwd
    entries at -1, and the out-of-grid row gets no pairs."""
    import ctypes as C
    from cumm_amd import _abi
    print("[pairs] strided over a given output list", flush=True)
    grid = (16, 96, 96)
    c = _coords([2500, 700], grid, 41)
    oc, oshape = pc.strided_out_coords(c, grid, (3, 3, 3), np.array([2, 2, 2]), np.array([1, 1, 1]), "reach")
    rng = np.random.default_rng(3)
    keep = np.sort(rng.choice(len(oc), len(oc) // 2, replace=False))
    sub = oc[keep][rng.permutation(len(keep))]
    sub = np.concatenate([sub, np.array([[0, oshape[0], 3, 3]], np.int32)])  # z one past the grid
    hf, hb = pc.strided_pairs(c, grid, sub[:-1], (3, 3, 3), np.array([2, 2, 2]), np.array([1, 1, 1]))
    L = _abi.lib()
    g = pairs.LayerSpec("strided").cs()
    n_in, n_out, kv = len(c), len(sub), 27
    dc, do = torch.from_numpy(c).cuda(), torch.from_numpy(np.ascontiguousarray(sub)).cuda()
    pf = torch.empty((kv, n_out), dtype=torch.int32, device="cuda")
    pb = torch.empty((kv, n_in), dtype=torch.int32, device="cuda")
    mf = torch.empty((n_out,), dtype=torch.int32, device="cuda")
    af = torch.empty((n_out,), dtype=torch.int32, device="cuda")
    mb = torch.empty((n_in,), dtype=torch.int32, device="cuda")
    ab = torch.empty((n_in,), dtype=torch.int32, device="cuda")
    nb = C.c_int64(0)
    assert L.cs_pairs_strided_geom_workspace(n_in, n_out, C.byref(g), C.byref(nb)) == 0
    ws = torch.empty((max(nb.value, 1),), dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    assert L.cs_pairs_strided_geom(dc.data_ptr(), n_in, pairs._shape_ptr(grid), C.byref(g), do.data_ptr(), n_out,
                                   pf.data_ptr(), pb.data_ptr(), mf.data_ptr(), af.data_ptr(), mb.data_ptr(),
                                   ab.data_ptr(), ws.data_ptr(), ws.numel(), s) == 0
    torch.cuda.synchronize()
    got_f, got_b = _np(pf), _np(pb)
    np.testing.assert_array_equal(got_f[:, :-1], hf)
    assert (got_f[:, -1] == -1).all()
    np.testing.assert_array_equal(got_b, hb)
    assert (got_b >= 0).any() and (got_b == -1).any()


def test_device_tables_drive_a_dynamic_stack():
    """5x5x5 submanifold layers (dynamic masks) and a 2x2x2 stride-2 layer,
    device-resident: the stack gives the same bits over device and host tables."""
    grid = (16, 96, 96)
    kinds = (("subm", 5), ("strided", 2, 2, 0), ("subm", 5))
    c = _coords([2500, 700], grid, 19)
    ch = (16, 32, 32, 64)
    kvs = (125, 8, 125)
    rng = np.random.default_rng(4)
    Ws = [rng.normal(0, 0.05, (ch[i + 1], kvs[i], ch[i])).astype(np.float32) for i in range(3)]
    host = B.build_stack_tables(kinds, c, grid)
    dev = pairs.stack_tables(kinds, torch.from_numpy(c).cuda(), grid)
    assert [t.pair_fwd.shape[0] for t in host] == list(kvs)
    X = rng.standard_normal((c.shape[0], ch[0])).astype(np.float32)
    dO = rng.standard_normal((host[-1].n_out, ch[-1])).astype(np.float32)
    outs = []
    for tabs in (host, dev):
        st = B.ConvStack(tabs, ch, "f16", Ws, torch.cuda.current_stream().cuda_stream)
        st.bind(torch.from_numpy(X).cuda().half(), torch.from_numpy(dO).cuda().half())
        st.step()
        torch.cuda.synchronize()
        outs.append([st.layers[-1].out.clone(), st.layers[2].din.clone(), st.layers[1].din.clone(),
                     st.layers[0].dw.clone(), st.layers[2].dw.clone()])
    for a, b in zip(*outs):
        assert torch.equal(a, b)


@pytest.mark.parametrize("kind", ["subm", "strided"])
def test_unsorted_row_masks(kind):
    """sort_masks=False: the same pair table, masks in row order (bit f of row
    i <=> pair[f][i] != -1, as the host builder's masks before its sort) and
    the identity as argsort."""
    print(f"[pairs] unsorted row masks {kind}", flush=True)
    grid = (16, 96, 96)
    c = torch.from_numpy(_coords([4000, 1500], grid, 21)).cuda()
    if kind == "subm":
        srt = pairs.subm_tables(c, grid, backward=False)
        uns = pairs.subm_tables(c, grid, backward=False, sort_masks=False)
    else:
        srt = pairs.strided_tables(c, grid, backward=False)
        uns = pairs.strided_tables(c, grid, backward=False, sort_masks=False)
    torch.cuda.synchronize()
    assert torch.equal(srt.pair_fwd, uns.pair_fwd)
    pf = _np(uns.pair_fwd)
    exp = np.zeros(pf.shape[1], np.uint32)
    for f in range(pf.shape[0]):
        exp |= (pf[f] >= 0).astype(np.uint32) << np.uint32(f)
    assert np.array_equal(_np(uns.mask_fwd).view(np.uint32), exp)
    assert np.array_equal(_np(uns.argsort_fwd), np.arange(pf.shape[1], dtype=np.int32))
    # the sorted masks are the row masks permuted by the argsort
    assert np.array_equal(_np(srt.mask_fwd).view(np.uint32), exp[_np(srt.argsort_fwd)])


def test_unsorted_masks_drive_a_stack_bit_identically():
    """A forward-only bf16 stack (subm, strided, subm, strided) over unsorted
    masks gives the same bits as over mask-sorted tables: a row's sum runs over
    its own offsets in ascending order whatever tile it lands in."""
    print("[pairs] unsorted stack", flush=True)
    grid = (16, 96, 96)
    c = torch.from_numpy(_coords([6000, 2500], grid, 23)).cuda()
    kinds = ("subm", "strided", "subm", "strided")
    chans = (16, 32, 64, 64, 128)
    rng = np.random.default_rng(3)
    Ws = [torch.from_numpy(rng.normal(0, 1 / np.sqrt(27 * chans[i]), size=(chans[i + 1], 27, chans[i]))
                           .astype(np.float32)).cuda().to(torch.bfloat16) for i in range(4)]
    x = torch.from_numpy(rng.standard_normal((c.shape[0], 16)).astype(np.float32)).cuda().to(torch.bfloat16)
    outs = []
    for srt in (True, False):
        tabs = pairs.stack_tables(kinds, c, grid, backward=False, sort_masks=srt)
        st = B.ConvStack(tabs, chans, "bf16", Ws, torch.cuda.current_stream().cuda_stream, train=False)
        st.bind(x)
        st.step()
        torch.cuda.synchronize()
        outs.append(st.layers[-1].out.clone())
    assert torch.equal(outs[0], outs[1])
