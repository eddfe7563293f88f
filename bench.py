"""Benchmark: submanifold 3x3x3 sparse conv, 120k voxels, 64 -> 64 channels.

One step = forward + dgrad + wgrad through ConvMainUnitTest.implicit_gemm2
(the drop-in boundary, C ABI, hand-written gfx950 kernels) on one synthetic
scene: 120 000 unique voxels on 8 random planes (+N(0, 0.5) voxel jitter) in a
41 x 1600 x 1408 grid, features ~ N(0, 1), weights ~ N(0, 1/sqrt(27*64)).
Inputs, pair tables, masks and workspaces are resident in HBM before timing.

Reported ``value``: effective TFLOP/s of the FP16 step (FP16 in, FP32
accumulate), FLOPs = 3 * 2 * P * C * K with P = count(pair_fwd != -1).
``per_dtype`` gives FP32 / FP16 / BF16 TFLOP/s, fraction of the MI355X dense
MFMA peak and the speedup over the plain FP32 path on the same GPU.
``roofline`` covers the step's dominant kernel, the row-gather implicit GEMM
(launched once by forward and once by dgrad), timed with HIP events on its
stream around the forward launch; ``traffic`` is its measured HBM bytes per
launch, read from the counter summary in profiles/ when that summary was taken
on the same workload.  ``cpu_baseline`` times the reference's CPU path (numpy
FP32 gather-GEMM-scatter over compact pairs) on a bounded sample on the host,
and ``accuracy_vs_fp64`` records the FP64-oracle error of every GPU result.

Profiling helpers: ``--ops fwd`` restricts the step to the listed ops and
``--copy-bytes N`` first runs a device-to-device copy of N bytes three times,
the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "cumm-rocm_amd"))

PEAK = {"f32": 157.3e12, "f16": 2.5e15, "bf16": 2.5e15, "e4m3": 5.0e15}
HBM_PEAK = 8.0e12
DT_BYTES = {"f32": 4, "f16": 2, "bf16": 2, "e4m3": 1}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", choices=("c2", "c3", "c4", "c5", "f3", "gemm"), default="c2",
                    help="c2: 120k voxels 64->64 (default); c3: precision form, 200k voxels 128->128; "
                         "c4: 32 clouds, strided 32->64 + subm 64->128, grouped launches under a HIP graph; "
                         "c5: serving, Poisson 20-200 frames/s, 4-layer BF16 stack; "
                         "f3: int8-inference forward and the gather / scatter-add glue on the c2 scene")
    ap.add_argument("--rates", default="20,50,100,200,400", help="c5 arrival rates (frames/s)")
    ap.add_argument("--seconds", type=float, default=3.0, help="c5 traffic per rate")
    ap.add_argument("--forms", default="", help="c5: comma-separated serving forms to run (default: all)")
    ap.add_argument("--voxels", type=int, default=None)
    ap.add_argument("--channels", type=int, default=None)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--dtypes", default=None)
    ap.add_argument("--ops", default="fwd,dgrad,wgrad")
    ap.add_argument("--copy-bytes", type=int, default=0)
    ap.add_argument("--no-extra", action="store_true",
                    help="c2 only: skip the precision, batched and serving forms in the same line")
    args = ap.parse_args()
    dflt = {"c2": (120000, 64, "f32,f16,bf16"), "c3": (200000, 128, "f32,bf16,f16,e4m3"),
            "c4": (0, 0, "f16"), "c5": (0, 0, "bf16"), "f3": (120000, 64, "i8"), "gemm": (0, 0, "f16")}[args.config]
    args.voxels = args.voxels or dflt[0]
    args.channels = args.channels or dflt[1]
    args.dtypes = args.dtypes or dflt[2]
    return args


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def out_bytes(dt: str, op: str) -> int:
    """Output element size: FP8 forward / dgrad write f16, FP8 wgrad FP32."""
    if dt == "e4m3":
        return 4 if op == "wgrad" else 2
    return DT_BYTES[dt]


def op_bytes(op: str, n: int, kv: int, C: int, K: int, s: int, s_out: int) -> int:
    """Compulsory HBM bytes of one launch (each tensor touched once)."""
    tiles = (n + 15) // 16
    if op == "fwd":
        return n * C * s + n * K * s_out + kv * C * K * s + kv * n * 4 + 2 * n * 4 + tiles * 4
    if op == "dgrad":
        return n * K * s + n * C * s_out + kv * C * K * s + kv * n * 4 + 2 * n * 4
    return n * K * s + n * C * s + kv * n * 4 + n * 4 + tiles * 4 + kv * C * K * s_out


class Layer:
    """Device-resident buffers and prebuilt ConvParams for one dtype."""

    def __init__(self, prob, X, W, dO, dt: str, stream: int, ops=("fwd", "dgrad", "wgrad")):
        import torch
        from cumm_amd import tensorview as tv
        from cumm_amd.conv.main import ConvMainUnitTest, select_conv_algo
        from cumm_amd.tensorview.gemm import ConvOpType, ConvParams

        TD = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16, "e4m3": torch.float8_e4m3fn}
        TV = {"f32": tv.float32, "f16": tv.float16, "bf16": tv.bfloat16, "e4m3": tv.float_e4m3}
        n, C, K, kv = prob.n, X.shape[1], W.shape[0], prob.kv
        dev = "cuda"
        self.scale_f = self.scale_d = None
        self.alpha_w = 1.0
        if dt == "e4m3":
            # FP8-E4M3 operands with absmax scales (448 = largest finite e4m3):
            # forward weights per output channel k, dgrad weights per output
            # channel c, features and gradients per tensor; FP32 accumulation,
            # f16 outputs for forward / dgrad, FP32 dW.
            s_x = float(np.abs(X).max()) / 448.0
            s_do = float(np.abs(dO).max()) / 448.0
            s_wk = np.abs(W).reshape(K, -1).max(axis=1) / 448.0
            s_wc = np.abs(W).max(axis=(0, 1)) / 448.0
            q = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev).to(TD["e4m3"])  # noqa: E731
            self.x = q(X / s_x)
            self.w = q(W / s_wk[:, None, None])
            self.w_d = q(W / s_wc[None, None, :])
            self.dout = q(dO / s_do)
            self.scale_f = torch.from_numpy((s_x * s_wk).astype(np.float32)).to(dev)
            self.scale_d = torch.from_numpy((s_do * s_wc).astype(np.float32)).to(dev)
            self.alpha_w = s_do * s_x
            o_fd, o_w = "f16", "f32"
        else:
            self.x = torch.from_numpy(X).to(dev, TD[dt])
            self.w = torch.from_numpy(W).to(dev, TD[dt])
            self.w_d = self.w
            self.dout = torch.from_numpy(dO).to(dev, TD[dt])
            o_fd = o_w = dt
        self.out = torch.empty((n, K), dtype=TD[o_fd], device=dev)
        self.din = torch.zeros((n, C), dtype=TD[o_fd], device=dev)
        self.dw = torch.empty((K, kv, C), dtype=TD[o_w], device=dev)
        pf = torch.from_numpy(prob.pair_fwd).to(dev)
        pb = torch.from_numpy(prob.pair_bwd).to(dev)
        ms = torch.from_numpy(prob.mask_sorted.view(np.int32)).to(dev)
        am = torch.from_numpy(prob.argsort).to(dev)
        fd = select_conv_algo(0, TV[dt], TV[dt], TV[o_fd], C, K)
        dd = select_conv_algo(1, TV[o_fd], TV[dt], TV[dt], C, K)
        wd = select_conv_algo(2, TV[dt], TV[o_w], TV[dt], C, K)
        tile = fd.tile_shape[0]
        self.mask_out = torch.zeros(((n + tile - 1) // tile,), dtype=torch.int32, device=dev)

        def params(op, desp):
            p = ConvParams(3, ConvOpType(op))
            p.conv_algo_desp = desp
            p.indices = tv.from_torch(pb if op == 1 else pf)
            p.mask = tv.from_torch(ms if op != 2 else self.mask_out)
            p.mask_argsort = tv.from_torch(am)
            p.stream = stream
            if op == 0:
                p.input, p.weight, p.output = tv.from_torch(self.x), tv.from_torch(self.w), tv.from_torch(self.out)
                p.mask_output = tv.from_torch(self.mask_out)
                if self.scale_f is not None:
                    p.scale = tv.from_torch(self.scale_f)
            elif op == 1:
                p.input, p.weight, p.output = tv.from_torch(self.din), tv.from_torch(self.w_d), tv.from_torch(self.dout)
                p.reverse_mask = True
                if self.scale_d is not None:
                    p.scale = tv.from_torch(self.scale_d)
            else:
                p.input, p.weight, p.output = tv.from_torch(self.x), tv.from_torch(self.dw), tv.from_torch(self.dout)
                p.mask_width = tile
                p.alpha = self.alpha_w
            ws = ConvMainUnitTest.query_workspace_size(p)
            if ws > 0:
                p.workspace = tv.from_torch(torch.empty((ws,), dtype=torch.uint8, device=dev))
            return p

        self.params = {"fwd": params(0, fd), "dgrad": params(1, dd), "wgrad": params(2, wd)}
        self.names = {k: repr(v.conv_algo_desp) for k, v in self.params.items()}
        self.run = ConvMainUnitTest.implicit_gemm2
        self.ops = [k for k in ("fwd", "dgrad", "wgrad") if k in ops]
        if "wgrad" in self.ops and "fwd" not in self.ops:
            self.run(self.params["fwd"])  # wgrad reads the forward's per-tile masks

    def step(self):
        for k in self.ops:
            self.run(self.params[k])


def time_dtype(layer: Layer, steps: int, warmup: int, stream):
    import torch

    for _ in range(warmup):
        layer.step()
    torch.cuda.synchronize()
    # Per-op device time with HIP events on the launch stream: `steps`
    # back-to-back launches of one op between two events, so the queue stays
    # ahead of the GPU and the average is the kernels' own duration.
    mean_op = {}
    for k in layer.ops:
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(steps):
            layer.run(layer.params[k])
        b.record(stream)
        torch.cuda.synchronize()
        mean_op[k] = a.elapsed_time(b) / steps
    per_op = dict(mean_op)
    # whole-step wall clock, synchronize on both sides, no events inside
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        layer.step()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall, per_op, mean_op


def copy_calibration(nbytes: int):
    """A device-to-device copy of nbytes, three times: the known-size kernel
    whose FETCH_SIZE / WRITE_SIZE readings scale the counters of the others."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy calibration: {nbytes} bytes x 3")


def kernel_name(op: str, dt: str, C: int, K: int) -> str:
    """The main kernel of a subm 3x3x3 op at these channels (the launcher's
    choice for the default catalogue kernel): wgrad's partial products; for
    forward / dgrad the weight-stationary row gather while the weights of all
    offsets fit two column slices (four for FP32), the streamed-weights kernel
    for wider 16-bit / fp8 layers, the block row gather for wide FP32 ones."""
    if op == "wgrad":
        return "wgrad_kernel"
    nout = K if op == "fwd" else C
    if nout <= 64:
        return "rowgather_ws_kernel"
    return "rowgather_kernel" if dt == "f32" else "rowgather_stream_kernel"


def load_traffic(kernel: str, workload: str):
    """Measured HBM bytes per launch of `kernel` on `workload` from the counter
    summary (profiles/traffic.json), or None when none was taken for them."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        rows = json.loads(f.read_text())
    except ValueError:
        return None
    for r in rows:
        if r.get("kernel") == kernel and r.get("workload") == workload:
            return r
    return None


def host_info() -> dict:
    """The host the CPU baseline runs on: logical CPUs, those this process may
    use, the CPU model and the BLAS thread count."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        from threadpoolctl import threadpool_info
        blas = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        blas = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu_count": os.cpu_count(), "cpus_usable": usable, "cpu_model": model, "blas_threads": int(blas)}


def accuracy_vs_fp64(prob, X, W, dO, gpu_outputs):
    """Max-abs and relative error of every GPU result against the FP64 oracle
    on the unquantised FP32 inputs, and mask_output bit-exactness."""
    sys.path.insert(0, str(ROOT))
    from oracle import spconv_ref as ref

    n = prob.n
    acc = {}
    y64, _ = ref.conv_forward(X, W, prob.pair_fwd, prob.mask_sorted, prob.argsort, 16)
    d64 = ref.conv_dgrad(dO, W, prob.pair_bwd, prob.mask_sorted, prob.argsort, 16, reverse_mask=True)
    tm = ref.tile_masks(prob.mask_sorted, n, 16).astype(np.uint32)
    w64 = ref.conv_wgrad(X, dO, prob.pair_fwd, tm, prob.argsort, 16)
    for dt_name, (y, d, w, mo) in gpu_outputs.items():
        e = {}
        for name, got, exp in (("fwd", y, y64), ("dgrad", d, d64), ("wgrad", w, w64)):
            mx, rel = ref.relative_error(got, exp)
            e[name] = {"max_abs": mx, "rel": rel}
        e["mask_output_bit_exact"] = bool(np.array_equal(mo, tm))
        acc[dt_name] = e
    return acc


def cpu_baseline(prob, X, W, dO, budget_s: float):
    """The reference's CPU path (numpy FP32 gather-GEMM-scatter, compact pairs)
    timed on the host."""
    sys.path.insert(0, str(ROOT))
    from oracle import spconv_ref as ref

    host = host_info()
    threads = host["blas_threads"]
    pairs, nums = ref.compact_pairs(prob.pair_fwd)
    n, C, K = prob.n, X.shape[1], W.shape[0]
    flops_step = 3 * 2 * prob.num_pairs * C * K
    # bounded sample: whole fwd+dgrad+wgrad steps until the budget is used
    steps = 0
    t0 = time.perf_counter()
    while True:
        ref.cpu_fwd(X, W, pairs, nums, n)
        ref.cpu_dgrad(dO, W, pairs, nums, n)
        ref.cpu_wgrad(X, dO, pairs, nums, K, C)
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = (time.perf_counter() - t0) / steps
    return {
        "value": flops_step / dt / 1e12, "unit": "TFLOP/s", "cores": int(threads), "kind": "port",
        "sample": f"{steps} full fwd+dgrad+wgrad steps of the same {n}-voxel {C}->{K} layer "
                  f"(numpy FP32, compact pairs), {dt * 1e3:.0f} ms/step",
        "ms_per_step": dt * 1e3,
        "host": host,
    }


def _clouds_worker(job):
    """Host-side scene and pair tables of one cloud (runs in a worker process)."""
    from cumm_amd import batch as B

    from cumm_amd import pointcloud as pc

    size, seed, b = job
    c = pc.plane_scene(size, seed=seed, batch_index=b)
    return c, B.build_stack_tables(("strided", "subm"), c, pc.GRID_KITTI)


def run_batched(args):
    """Batched form: 32 clouds (log-uniform 5k..150k voxels) through a strided
    3x3x3 s2 32->64 layer (outputs = unique(floor(c / 2))) and a subm 64->128
    layer, forward + backward (dgrad of the second layer, wgrad of both).
    Compared: the plain FP32 per-call path (one launch per cloud, layer and op),
    FP16 per cloud, FP16 grouped (one launch per layer and op over all clouds)
    and FP16 grouped with tuned kernels replayed from one HIP graph."""
    from concurrent.futures import ProcessPoolExecutor

    from cumm_amd import batch as B
    from cumm_amd import pointcloud as pc

    t0 = time.perf_counter()
    sizes = B.cloud_sizes(32, 5000, 150000, args.seed)
    jobs = [(s, args.seed + 7919 * b, b) for b, s in enumerate(sizes)]
    with ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        res = list(ex.map(_clouds_worker, jobs))
    clouds = [r[0] for r in res]
    per_cloud = [r[1] for r in res]
    coords = np.concatenate(clouds)
    tabs = B.build_stack_tables(("strided", "subm"), coords, pc.GRID_KITTI)
    log(f"c4: {len(sizes)} clouds, {coords.shape[0]} voxels; layer outputs "
        f"{[t.n_out for t in tabs]}, pairs {[t.num_pairs for t in tabs]}; host tables "
        f"{time.perf_counter() - t0:.1f}s")
    import torch
    from cumm_amd.tune import ConvTuner

    chans = (32, 64, 128)
    rng = np.random.default_rng(args.seed + 1)
    Ws = [rng.normal(0.0, 1.0 / np.sqrt(27 * chans[i]), size=(chans[i + 1], 27, chans[i])).astype(np.float32)
          for i in range(2)]
    X = rng.standard_normal((coords.shape[0], chans[0])).astype(np.float32)
    dO = rng.standard_normal((tabs[-1].n_out, chans[-1])).astype(np.float32)
    in_off = np.cumsum([0] + [c.shape[0] for c in clouds])
    out_off = np.cumsum([0] + [t[-1].n_out for t in per_cloud])
    stream = torch.cuda.current_stream()
    sid = stream.cuda_stream

    def stack(tables, dt, x, d):
        st = B.ConvStack(tables, chans, dt, Ws, sid, train=True)
        st.bind(torch.from_numpy(x).cuda().to(B.TORCH_DT[dt]), torch.from_numpy(d).cuda().to(B.TORCH_DT[dt]))
        return st

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    flops = None
    out = {}
    for dt in ("f32", "f16"):
        sts = [stack(per_cloud[b], dt, X[in_off[b]:in_off[b + 1]], dO[out_off[b]:out_off[b + 1]])
               for b in range(len(clouds))]
        flops = sum(st.flops() for st in sts)
        t = timed(lambda: [st.step() for st in sts], args.steps, args.warmup)
        out[f"{dt}_per_cloud"] = t
        log(f"c4 {dt} per-cloud eager: {t * 1e3:.3f} ms/step")
        if dt == "f16":
            per_cloud_y = [st.layers[-1].out.clone() for st in sts]
            per_cloud_dw = [st.layers[0].dw.double() for st in sts]
        del sts
    g = stack(tabs, "f16", X, dO)
    assert g.flops() == flops
    out["f16_grouped"] = timed(g.step, args.steps, args.warmup)
    log(f"c4 f16 grouped eager: {out['f16_grouped'] * 1e3:.3f} ms/step")
    y_eager = g.layers[-1].out.clone()
    dw_eager = [l.dw.clone() for l in g.layers]
    din_eager = g.layers[1].din.clone()
    tuner = ConvTuner()
    choices = g.tune(tuner)
    for c in choices:
        log(f"c4 tuned {c[0]} {c[1]}->{c[2]}: {c[3]} split {c[4]} {c[5]:.1f}us")
    g.step()
    torch.cuda.synchronize()
    y_tuned = g.layers[-1].out.clone()
    g.capture()
    out["f16_grouped_graph"] = timed(g.replay, args.steps, args.warmup)
    log(f"c4 f16 grouped tuned graph: {out['f16_grouped_graph'] * 1e3:.3f} ms/step")
    torch.cuda.synchronize()
    checks = {
        "graph_equals_eager_tuned_fwd": bool(torch.equal(g.layers[-1].out, y_tuned)),
        "tuned_equals_default_fwd": bool(torch.equal(y_tuned, y_eager)),
        "grouped_equals_per_cloud_fwd": bool(all(
            torch.equal(y_eager[out_off[b]:out_off[b + 1]], per_cloud_y[b]) for b in range(len(clouds)))),
        "dgrad_graph_equals_eager": bool(torch.equal(g.layers[1].din, din_eager)),
    }
    # wgrad of the first layer: grouped vs the sum of the per-cloud results
    dw_sum = torch.stack(per_cloud_dw).sum(0)
    rel = float((dw_eager[0].double() - dw_sum).norm() / dw_sum.norm())
    checks["wgrad_grouped_vs_sum_per_cloud_rel"] = rel
    tf = {k: flops / v / 1e12 for k, v in out.items()}
    line = {
        "metric": "sparse-conv TFLOP/s + % MFMA peak per dtype vs plain FP32; serving p50/p99 ms",
        "value": tf["f16_grouped_graph"], "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": out["f16_grouped_graph"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f16", "data": "synthetic",
        "config": {"workload": f"32 clouds log-uniform[5k,150k] ({coords.shape[0]} voxels), strided 3x3x3 s2 "
                               f"32->64 (unique(floor(c/2))) + subm 64->128, fwd+dgrad+wgrad",
                   "clouds": len(sizes), "voxels": int(coords.shape[0]),
                   "layer_out_rows": [t.n_out for t in tabs], "pairs": [t.num_pairs for t in tabs],
                   "flops_per_step": flops},
        "forms_ms": {k: v * 1e3 for k, v in out.items()},
        "forms_tflops": tf,
        "speedup_vs_f32_per_call": {k: out["f32_per_cloud"] / v for k, v in out.items()},
        "frac_mfma_peak_f16": tf["f16_grouped_graph"] * 1e12 / PEAK["f16"],
        "tuned": [{"op": c[0], "C": c[1], "K": c[2], "kernel": c[3], "split_k": c[4], "us": c[5]} for c in choices],
        "checks": checks,
    }
    return line


def _frame_worker(job):
    from cumm_amd import pointcloud as pc

    size, seed = job
    return pc.plane_scene(size, seed=seed)


def run_serving(args):
    """Serving form: Poisson arrivals at each rate, frames of 20k-200k voxels
    (a pool of 16 scenes generated as for the single-layer workload), 4-layer
    stack subm 16->32, strided 32->64, subm 64->64, strided 64->128.  The BF16
    server batches continuously on two streams (a launch thread each) under an
    HBM budget, and on one stream for comparison; the plain FP32 per-call path
    serves the same trace one frame at a time."""
    from concurrent.futures import ProcessPoolExecutor

    from cumm_amd import batch as B

    t0 = time.perf_counter()
    sizes = B.cloud_sizes(16, 20000, 200000, args.seed + 5)
    with ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        pool = list(ex.map(_frame_worker, [(s, args.seed + 31 * i) for i, s in enumerate(sizes)]))
    log(f"c5: frame pool {sizes} generated in {time.perf_counter() - t0:.1f}s")
    import torch
    from cumm_amd.serve import Frame, SparseConvServer

    kinds = ("subm", "strided", "subm", "strided")
    chans = (16, 32, 64, 64, 128)
    rng = np.random.default_rng(args.seed + 2)
    Ws = [rng.normal(0.0, 1.0 / np.sqrt(27 * chans[i]), size=(chans[i + 1], 27, chans[i])).astype(np.float32)
          for i in range(4)]
    feats = [rng.standard_normal((c.shape[0], chans[0])).astype(np.float32) for c in pool]

    def trace(rate, seed):
        r = np.random.default_rng(seed)
        n = max(20, int(rate * args.seconds))
        idx = r.integers(0, len(pool), n)
        arr = np.cumsum(r.exponential(1.0 / rate, n))
        return [Frame(i, pool[j], feats[j]) for i, j in enumerate(idx)], arr

    forms = {"bf16_continuous": dict(dt="bf16", max_batch=16, n_streams=2),
             "bf16_continuous_1stream": dict(dt="bf16", max_batch=16, n_streams=1),
             "f32_per_call": dict(dt="f32", max_batch=1, n_streams=1)}
    if getattr(args, "forms", ""):
        forms = {k: forms[k] for k in args.forms.split(",")}
    res = {}
    for name, kw in forms.items():
        srv = SparseConvServer(kinds, chans, Ws, keep_outputs=False, hbm_budget=16 << 30, **kw)
        # warm-up: first use of every kernel, and each stream's allocator pool
        # and pinned staging grown by the whole frame pool arriving at once
        for _ in range(2):
            srv.run_trace([Frame(i, c, x) for i, (c, x) in enumerate(zip(pool, feats))], np.zeros(len(pool)))
            srv.run_trace([Frame(i, c, x) for i, (c, x) in enumerate(zip(pool, feats))],
                          np.arange(len(pool)) * 1e-3)
        torch.cuda.synchronize()
        for rate in [float(v) for v in args.rates.split(",")]:
            fr, arr = trace(rate, int(rate) + args.seed)
            st = srv.run_trace(fr, arr)
            res.setdefault(name, {})[str(int(rate))] = st
            log(f"c5 {name} {rate:.0f} fps: p50 {st['p50_ms']:.2f} ms  p99 {st['p99_ms']:.2f} ms  "
                f"{st['frames_per_s']:.1f} frames/s  mean batch {st['mean_batch']:.2f}")
        srv.close()
        del srv
        torch.cuda.synchronize()
    top = str(int(float(args.rates.split(",")[-1])))
    main = res[next(iter(res))][top]
    line = {
        "metric": "sparse-conv TFLOP/s + % MFMA peak per dtype vs plain FP32; serving p50/p99 ms",
        "value": main["p99_ms"], "unit": "ms (p99 latency)", "n_gpus": 1, "steps": main["frames"], "warmup": 4 * 16,
        "ms_per_step": main["mean_ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "serving: Poisson arrivals, frames of 20k-200k voxels, 4-layer stack subm 16->32, "
                               "strided 32->64, subm 64->64, strided 64->128 (unique(floor(c/2)) outputs)",
                   "rates_fps": [float(v) for v in args.rates.split(",")], "seconds_per_rate": args.seconds,
                   "frame_pool_voxels": sizes, "hbm_budget_bytes": 16 << 30},
        "serving": res,
    }
    return line


def run_f3(args):
    """int8-inference forward (int8 features / weights, int32 accumulation,
    float scale + bias epilogue, int8 out) and the gather-GEMM-scatter glue of
    spconv's native algorithm (per kernel offset: gather the offset's input
    rows, scatter-add its FP32 output rows), on the c2 scene."""
    import torch
    from cumm_amd import pointcloud as pc
    from cumm_amd import tensorview as tv
    from cumm_amd.conv.main import ConvMainUnitTest, select_conv_algo
    from cumm_amd.gemm.gather import GatherAll, ScatterAll
    from cumm_amd.tensorview.gemm import ConvOpType, ConvParams

    coords = pc.plane_scene(args.voxels, seed=args.seed)
    prob = pc.build_subm(coords, pc.GRID_KITTI)
    n, C, K, kv = prob.n, args.channels, args.channels, prob.kv
    P = prob.num_pairs
    rng = np.random.default_rng(args.seed + 1)
    dev = "cuda"
    stream = torch.cuda.current_stream()
    sid = stream.cuda_stream

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(args.steps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / args.steps * 1e-3

    # ---- int8 inference forward
    xq = torch.from_numpy(rng.integers(-127, 128, (n, C)).astype(np.int8)).to(dev)
    wq = torch.from_numpy(rng.integers(-127, 128, (K, kv, C)).astype(np.int8)).to(dev)
    out = torch.empty((n, K), dtype=torch.int8, device=dev)
    p = ConvParams(3, ConvOpType.Forward)
    p.conv_algo_desp = select_conv_algo(0, tv.int8, tv.int8, tv.int8, C, K)
    p.input, p.weight, p.output = tv.from_torch(xq), tv.from_torch(wq), tv.from_torch(out)
    p.indices = tv.from_torch(torch.from_numpy(prob.pair_fwd).to(dev))
    p.mask = tv.from_torch(torch.from_numpy(prob.mask_sorted.view(np.int32)).to(dev))
    p.mask_argsort = tv.from_torch(torch.from_numpy(prob.argsort).to(dev))
    p.mask_output = tv.from_torch(torch.zeros(((n + 15) // 16,), dtype=torch.int32, device=dev))
    p.scale = tv.from_torch(torch.full((K,), 1e-4, device=dev))
    p.bias = tv.from_torch(torch.zeros((K,), device=dev))
    p.stream = sid
    t_i8 = timed(lambda: ConvMainUnitTest.implicit_gemm2(p))
    b_i8 = op_bytes("fwd", n, kv, C, K, 1, 1)
    ops_i8 = 2 * P * C * K
    log(f"int8 fwd: {t_i8 * 1e6:.1f} us  {ops_i8 / t_i8 / 1e12:.1f} TOP/s  {b_i8 / t_i8 / 1e9:.0f} GB/s")

    # ---- gather / scatter-add glue, per kernel offset (compact pairs)
    ins, outs = [], []
    for f in range(kv):
        o = np.nonzero(prob.pair_fwd[f] >= 0)[0].astype(np.int32)
        ins.append(torch.from_numpy(prob.pair_fwd[f][o].astype(np.int32)).to(dev))
        outs.append(torch.from_numpy(o).to(dev))
    feat = torch.from_numpy(rng.standard_normal((n, C)).astype(np.float16)).to(dev)
    buf = torch.empty((int(max(t.numel() for t in ins)), C), dtype=torch.float16, device=dev)
    part = torch.from_numpy(rng.standard_normal((buf.shape[0], K)).astype(np.float32)).to(dev)
    acc = torch.zeros((n, K), dtype=torch.float32, device=dev)
    gp, sp = (8, 2048, 16, 256), (8, 2048, 16, 256)
    ga, sa = GatherAll(), ScatterAll()
    tvs = [(tv.from_torch(buf[:i.numel()]), tv.from_torch(i), tv.from_torch(part[:i.numel()]), tv.from_torch(o))
           for i, o in zip(ins, outs)]
    tfeat, tacc = tv.from_torch(feat), tv.from_torch(acc)

    def gather_all():
        for b, i, _, _ in tvs:
            ga.gather(b, tfeat, i, *gp, sid)

    def scatter_all():
        for _, _, pr, o in tvs:
            sa.scatter(tacc, pr, o, *sp, sid)

    t_g = timed(gather_all)
    t_s = timed(scatter_all)
    b_g = P * C * 2 * 2 + P * 4               # read + write each gathered row, indices
    b_s = P * K * 4 * 3 + P * 4               # read partial, read + write target, indices
    log(f"gather: {t_g * 1e6:.1f} us {b_g / t_g / 1e9:.0f} GB/s; scatter-add: {t_s * 1e6:.1f} us "
        f"{b_s / t_s / 1e9:.0f} GB/s ({kv} launches each)")
    workload = f"subm3x3x3 forward, {n} voxels on 8 planes in 41x1600x1408, {C}->{K} ch"
    line = {
        "metric": "int8-inference sparse-conv TOP/s; gather / scatter-add GB/s",
        "value": ops_i8 / t_i8 / 1e12, "unit": "TOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_i8 * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "i8", "data": "synthetic",
        "config": {"workload": workload, "voxels": n, "pairs": P, "channels_in": C, "channels_out": K,
                   "desp": repr(p.conv_algo_desp)},
        "roofline": {"kernel": "rowgather_ws_kernel<int8_t> (fwd)", "bound": "hbm", "achieved": b_i8 / t_i8 / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": b_i8 / t_i8 / HBM_PEAK, "traffic": None,
                     "algorithmic_bytes": b_i8, "avg_us": t_i8 * 1e6,
                     "mfma_frac": ops_i8 / t_i8 / 5.0e15},
        "glue": {"gather_us": t_g * 1e6, "gather_GB/s": b_g / t_g / 1e9, "gather_frac_hbm": b_g / t_g / HBM_PEAK,
                 "scatter_us": t_s * 1e6, "scatter_GB/s": b_s / t_s / 1e9,
                 "scatter_frac_hbm": b_s / t_s / HBM_PEAK, "launches": kv, "rows": P,
                 "gather_dtype": "f16", "scatter_dtype": "f32"},
    }
    print(json.dumps(line))


def run_gemm(args, rows: int = 64000, n: int = 128, k: int = 128) -> dict:
    """matmul2 in its spconv roles at rows = 64000, n = k = 128: ShuffleAC
    (C[c_inds] = A[a_inds] @ B, gathered rows in, scattered rows out; B
    row-major [k, n], read k-major in place, and n-major [n, k]) and ShuffleAB
    (C = A[a_inds]^T @ B[b_inds], the weight-gradient form, default split) for
    f32, f16, bf16 and int8 (int8 and int32 outputs).  Indices are random
    permutations.  Per call: mean over `steps` back-to-back calls between HIP
    events; algorithmic HBM bytes = gathered rows + written rows + B +
    indices (ShuffleAB: both gathered operands + indices + the output)."""
    import torch
    from cumm_amd import tensorview as tv
    from cumm_amd.gemm.main import GemmMainUnitTest
    from cumm_amd.tensorview.gemm import GemmParams

    rng = np.random.default_rng(args.seed)
    stream = torch.cuda.current_stream()
    tdt = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16, "s8": torch.int8,
           "s32": torch.int32}
    vdt = {"f32": tv.float32, "f16": tv.float16, "bf16": tv.bfloat16, "s8": tv.int8, "s32": tv.int32}
    size = {"f32": 4, "f16": 2, "bf16": 2, "s8": 1, "s32": 4}
    peak = {"f32": PEAK["f32"], "f16": PEAK["f16"], "bf16": PEAK["bf16"], "s8": 5.0e15}
    descs = GemmMainUnitTest.get_all_algo_desp()

    def desp(din, dout, shuffle, ta, tb):
        for d in descs:
            if (d.dtype_a == vdt[din] and d.dtype_c == vdt[dout] and int(d.shuffle_type) == shuffle
                    and d.trans_a == ta and d.trans_b == tb and not d.trans_c and d.dacc != tv.float16):
                return d
        raise LookupError((din, dout, shuffle, ta, tb))

    def operand(shape, dt):
        if dt == "s8":
            return torch.from_numpy(rng.integers(-128, 128, shape).astype(np.int8)).cuda()
        return torch.from_numpy(rng.uniform(-1, 1, shape).astype(np.float32)).cuda().to(tdt[dt])

    perm = lambda: torch.from_numpy(rng.permutation(rows).astype(np.int32)).cuda()
    a_inds, c_inds, b_inds = perm(), perm(), perm()

    def timed(p):
        # a caller that runs the same shape again keeps its workspace (split-K
        # partials, transposes): queried once and allocated before timing, as
        # the conv layers do; without one the library allocates per call
        ws = GemmMainUnitTest.query_workspace_size(p)
        if ws > 0:
            p.workspace = tv.from_torch(torch.empty((ws,), dtype=torch.uint8, device="cuda"))
        for _ in range(args.warmup):
            GemmMainUnitTest.matmul2(p)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            GemmMainUnitTest.matmul2(p)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps * 1e-3

    forms = {}
    for din, dout in (("f32", "f32"), ("f16", "f16"), ("bf16", "bf16"), ("s8", "s8"), ("s8", "s32")):
        es, eo = size[din], size[dout]
        A = operand((rows, k), din)
        flops = 2.0 * rows * n * k
        for tb in (False, True):
            B = operand((n, k) if tb else (k, n), din)
            C = torch.zeros((rows, n), dtype=tdt[dout], device="cuda")
            p = GemmParams()
            p.algo_desp = desp(din, dout, 1, False, tb)
            p.a, p.b, p.c = tv.from_torch(A), tv.from_torch(B), tv.from_torch(C)
            p.a_inds, p.c_inds = tv.from_torch(a_inds), tv.from_torch(c_inds)
            p.stream = stream.cuda_stream
            t = timed(p)
            byt = rows * k * es + rows * n * eo + k * n * es + 2 * rows * 4
            forms[f"shuffle_ac_{din}_{dout}_{'nmajor_b' if tb else 'kmajor_b'}"] = {
                "us": t * 1e6, "GB/s": byt / t / 1e9, "frac_hbm": byt / t / HBM_PEAK,
                "TFLOP/s": flops / t / 1e12, "frac_mfma": flops / t / peak[din], "bytes": byt}
            log(f"gemm ShuffleAC {din}->{dout} {'b [n,k]' if tb else 'b [k,n]'}: {t * 1e6:.1f} us  "
                f"{byt / t / 1e9:.0f} GB/s  {flops / t / 1e12:.1f} TFLOP/s")
        Bg = operand((rows, n), din)
        C = torch.zeros((k, n), dtype=tdt[dout], device="cuda")
        p = GemmParams()
        p.algo_desp = desp(din, dout, 2, True, False)
        p.a, p.b, p.c = tv.from_torch(A), tv.from_torch(Bg), tv.from_torch(C)
        p.a_inds, p.b_inds = tv.from_torch(a_inds), tv.from_torch(b_inds)
        p.stream = stream.cuda_stream
        t = timed(p)
        byt = rows * k * es + rows * n * es + 2 * rows * 4 + k * n * eo
        forms[f"shuffle_ab_{din}_{dout}"] = {
            "us": t * 1e6, "GB/s": byt / t / 1e9, "frac_hbm": byt / t / HBM_PEAK, "TFLOP/s": flops / t / 1e12,
            "frac_mfma": flops / t / peak[din], "bytes": byt}
        log(f"gemm ShuffleAB {din}->{dout}: {t * 1e6:.1f} us  {byt / t / 1e9:.0f} GB/s  {flops / t / 1e12:.1f} TFLOP/s")
    main = forms["shuffle_ac_f16_f16_kmajor_b"]
    return {
        "metric": "matmul2 gather GEMM GB/s (ShuffleAC f16, rows 64000, n = k = 128)",
        "value": main["GB/s"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": main["us"] * 1e-3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16", "data": "synthetic",
        "config": {"workload": f"matmul2 ShuffleAC / ShuffleAB, rows {rows}, n {n}, k {k}, random permutation indices"},
        "roofline": {"bound": "hbm", "achieved": main["GB/s"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": main["frac_hbm"], "traffic": None},
        "forms": forms,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.config == "c4":
        print(json.dumps(run_batched(args)))
        return
    if args.config == "c5":
        print(json.dumps(run_serving(args)))
        return
    if args.config == "f3":
        return run_f3(args)
    if args.config == "gemm":
        print(json.dumps(run_gemm(args)))
        return
    line = run_single(args, with_cpu=not args.no_cpu_baseline)
    if args.config == "c2" and not args.no_extra:
        line["configs"] = run_extra(args)
    print(json.dumps(line))


def run_extra(args) -> dict:
    """The precision (200k voxels, 128 -> 128, f32 / bf16 / f16 / e4m3),
    batched (32 clouds under a HIP graph), matmul2 gather GEMM (rows 64000,
    n = k = 128) and serving (Poisson arrivals, 1 s per rate) forms,
    summarised for the same output line as the c2 headline."""
    import copy

    out = {}
    a3 = copy.copy(args)
    a3.config, a3.voxels, a3.channels, a3.dtypes = "c3", 200000, 128, "f32,bf16,f16,e4m3"
    a3.steps, a3.warmup = 20, 5
    t0 = time.perf_counter()
    try:
        l3 = run_single(a3, with_cpu=False)
        out["c3"] = {"workload": l3["config"]["workload"], "pairs": l3["config"]["pairs"],
                     "mac_efficiency": l3["config"]["mac_efficiency"],
                     "per_dtype": {k: {"tflops": v["tflops"], "speedup_vs_f32": v["speedup_vs_f32"],
                                       "frac_mfma_peak": v["frac_mfma_peak"], "op_us": v["op_us"]}
                                   for k, v in l3["per_dtype"].items()},
                     "accuracy_vs_fp64": l3["accuracy_vs_fp64"], "roofline_ops": l3["roofline_ops"]}
    except Exception as e:  # the headline must still print
        out["c3"] = {"error": repr(e)}
    log(f"extra c3: {time.perf_counter() - t0:.1f}s")
    a4 = copy.copy(args)
    a4.steps, a4.warmup = 10, 3
    t0 = time.perf_counter()
    try:
        l4 = run_batched(a4)
        out["c4"] = {k: l4[k] for k in ("value", "unit", "ms_per_step", "config", "forms_ms", "forms_tflops",
                                        "speedup_vs_f32_per_call", "frac_mfma_peak_f16", "checks")}
    except Exception as e:
        out["c4"] = {"error": repr(e)}
    log(f"extra c4: {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    try:
        ag = copy.copy(args)
        ag.steps, ag.warmup = 50, 10
        lg = run_gemm(ag)
        out["gemm"] = {k: lg[k] for k in ("value", "unit", "ms_per_step", "config", "roofline", "forms")}
    except Exception as e:
        out["gemm"] = {"error": repr(e)}
    log(f"extra gemm: {time.perf_counter() - t0:.1f}s")
    a5 = copy.copy(args)
    a5.seconds, a5.rates = 1.0, "20,50,100,200,400"
    t0 = time.perf_counter()
    try:
        l5 = run_serving(a5)
        out["c5"] = {"config": l5["config"], "serving": l5["serving"]}
    except Exception as e:
        out["c5"] = {"error": repr(e)}
    log(f"extra c5: {time.perf_counter() - t0:.1f}s")
    return out


def run_single(args, with_cpu: bool) -> dict:
    """One subm 3x3x3 layer (c2 / c3): per-dtype step time, per-op kernel
    time, roofline of every op, FP64 error of every result, and (c2) the CPU
    baseline."""
    import torch
    from cumm_amd import pointcloud as pc

    t0 = time.perf_counter()
    coords = pc.plane_scene(args.voxels, seed=args.seed)
    prob = pc.build_subm(coords, pc.GRID_KITTI)
    C = K = args.channels
    rng = np.random.default_rng(args.seed + 1)
    X = rng.standard_normal((prob.n, C)).astype(np.float32)
    W = (rng.normal(0.0, 1.0 / np.sqrt(prob.kv * C), size=(prob.kv, C, K))
         .transpose(2, 0, 1).astype(np.float32).copy())  # generated [27, C, K] -> KRSC [K, 27, C]
    dO = rng.standard_normal((prob.n, K)).astype(np.float32)
    log(f"scene: N={prob.n} P={prob.num_pairs} P/N={prob.num_pairs / prob.n:.3f} "
        f"gen {time.perf_counter() - t0:.1f}s")
    P = prob.num_pairs
    flops_op = 2 * P * C * K
    flops_step = 3 * flops_op

    if args.copy_bytes > 0:
        copy_calibration(args.copy_bytes)
    stream = torch.cuda.current_stream()
    sid = stream.cuda_stream
    ops = tuple(args.ops.split(","))
    results = {}
    gpu_outputs = {}
    for dt in args.dtypes.split(","):
        layer = Layer(prob, X, W, dO, dt, sid, ops)
        wall, per_op, mean_op = time_dtype(layer, args.steps, args.warmup, stream)
        results[dt] = (wall, per_op, layer.names, mean_op)
        torch.cuda.synchronize()
        gpu_outputs[dt] = (layer.out.double().cpu().numpy(), layer.din.double().cpu().numpy(),
                           layer.dw.double().cpu().numpy(), layer.mask_out.cpu().numpy().view(np.uint32))
        log(f"{dt}: step {wall * 1e3:.3f} ms  " + " ".join(f"{k}={v * 1e3:.1f}us" for k, v in per_op.items()))
        del layer

    main_dt = "f16" if "f16" in results else list(results)[0]
    wall, per_op, names, mean_op = results[main_dt]
    flops_step = len(ops) * flops_op
    f32_tf = flops_step / results["f32"][0] / 1e12 if "f32" in results else None
    per_dtype = {}
    for dt, (w, po, nm, _) in results.items():
        tf = flops_step / w / 1e12
        per_dtype[dt] = {
            "tflops": tf, "ms_per_step": w * 1e3, "frac_mfma_peak": tf * 1e12 / PEAK[dt],
            "speedup_vs_f32": (tf / f32_tf) if f32_tf else None,
            "op_us": {k: v * 1e3 for k, v in po.items()},
            "op_tflops": {k: flops_op / (v * 1e-3) / 1e12 for k, v in po.items()},
        }
    workload = (f"subm3x3x3 fwd+dgrad+wgrad, {prob.n} voxels on 8 planes in 41x1600x1408, "
                f"{C}->{K} ch")
    # Dominant kernel of the step: the row-gather implicit GEMM, launched by
    # forward and by dgrad (one kernel each); the forward launch is timed alone.
    s = DT_BYTES[main_dt]
    rop = "fwd" if "fwd" in mean_op else ops[0]
    byt = op_bytes(rop, prob.n, prob.kv, C, K, s, out_bytes(main_dt, rop))
    t_k = mean_op[rop] * 1e-3
    ach_bw = byt / t_k
    kname = kernel_name(rop, main_dt, C, K)
    tr = load_traffic(f"{kname}<{main_dt}> {rop}", workload)
    roofline = {"kernel": f"{kname}<{main_dt}> ({rop})", "desp": names[rop], "bound": "hbm",
                "achieved": ach_bw / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": ach_bw / HBM_PEAK,
                "traffic": tr["traffic_bytes"] if tr else None,
                "algorithmic_bytes": byt, "avg_us": mean_op[rop] * 1e3,
                "mfma_tflops": flops_op / t_k / 1e12, "mfma_frac": flops_op / t_k / PEAK[main_dt]}
    # every op of the reported dtype against its own roof (ops of several kernels
    # are timed as a whole: dgrad = weight transpose + row gather, wgrad = plan +
    # partial products + fixed-order reduce)
    roofline_ops = {}
    for k, v in mean_op.items():
        b = op_bytes(k, prob.n, prob.kv, C, K, s, out_bytes(main_dt, k))
        trk = load_traffic(f"{kernel_name(k, main_dt, C, K)}<{main_dt}> {k}", workload)
        roofline_ops[k] = {"avg_us": v * 1e3, "algorithmic_bytes": b, "GB/s": b / (v * 1e-3) / 1e9,
                           "frac_hbm": b / (v * 1e-3) / HBM_PEAK,
                           "frac_mfma": flops_op / (v * 1e-3) / PEAK[main_dt],
                           "kernel": kernel_name(k, main_dt, C, K),
                           "traffic": trk["traffic_bytes"] if trk else None}

    cpu = cpu_baseline(prob, X, W, dO, args.cpu_seconds) if with_cpu else None
    accuracy = accuracy_vs_fp64(prob, X, W, dO, gpu_outputs)
    # MAC efficiency of the mask tiles: valid pairs / (16 rows x offsets of each tile's mask)
    mac_eff = pc.mac_efficiency(prob.mask_sorted, 16, P)

    line = {
        "metric": "sparse-conv TFLOP/s + % MFMA peak per dtype vs plain FP32; serving p50/p99 ms",
        "value": flops_step / wall / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": main_dt, "data": "synthetic",
        "config": {"workload": workload, "ops": list(ops), "voxels": prob.n, "pairs": P, "channels_in": C,
                   "channels_out": K, "kernel_volume": prob.kv, "flops_per_step": flops_step,
                   "mac_efficiency": mac_eff, "dense_equivalent_flops_per_op": 2 * prob.n * prob.kv * C * K},
        "per_dtype": per_dtype,
        "roofline": roofline,
        "roofline_ops": roofline_ops,
        "cpu_baseline": cpu,
        "accuracy_vs_fp64": accuracy,
    }
    return line


if __name__ == "__main__":
    main()
