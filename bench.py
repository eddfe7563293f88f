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
``roofline`` covers the step's dominant kernel, timed with HIP events on its
stream.  ``cpu_baseline`` times the reference's CPU path (numpy FP32
gather-GEMM-scatter over compact pairs) on a bounded sample on the host, and
records the FP64-oracle error of every GPU result.
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
    ap.add_argument("--voxels", type=int, default=120000)
    ap.add_argument("--channels", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--dtypes", default="f32,f16,bf16")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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

    def __init__(self, prob, X, W, dO, dt: str, stream: int):
        import torch
        from cumm_amd import tensorview as tv
        from cumm_amd.conv.main import ConvMainUnitTest, select_conv_algo
        from cumm_amd.tensorview.gemm import ConvOpType, ConvParams

        tdt = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[dt]
        tvdt = {"f32": tv.float32, "f16": tv.float16, "bf16": tv.bfloat16}[dt]
        n, C, K, kv = prob.n, X.shape[1], W.shape[0], prob.kv
        dev = "cuda"
        self.x = torch.from_numpy(X).to(dev, tdt)
        self.w = torch.from_numpy(W).to(dev, tdt)
        self.dout = torch.from_numpy(dO).to(dev, tdt)
        self.out = torch.empty((n, K), dtype=tdt, device=dev)
        self.din = torch.zeros((n, C), dtype=tdt, device=dev)
        self.dw = torch.empty((K, kv, C), dtype=tdt, device=dev)
        pf = torch.from_numpy(prob.pair_fwd).to(dev)
        pb = torch.from_numpy(prob.pair_bwd).to(dev)
        ms = torch.from_numpy(prob.mask_sorted.view(np.int32)).to(dev)
        am = torch.from_numpy(prob.argsort).to(dev)
        fd = select_conv_algo(0, tvdt, tvdt, tvdt, C, K)
        dd = select_conv_algo(1, tvdt, tvdt, tvdt, C, K)
        wd = select_conv_algo(2, tvdt, tvdt, tvdt, C, K)
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
            elif op == 1:
                p.input, p.weight, p.output = tv.from_torch(self.din), tv.from_torch(self.w), tv.from_torch(self.dout)
                p.reverse_mask = True
            else:
                p.input, p.weight, p.output = tv.from_torch(self.x), tv.from_torch(self.dw), tv.from_torch(self.dout)
                p.mask_width = tile
            ws = ConvMainUnitTest.query_workspace_size(p)
            if ws > 0:
                p.workspace = tv.from_torch(torch.empty((ws,), dtype=torch.uint8, device=dev))
            return p

        self.params = {"fwd": params(0, fd), "dgrad": params(1, dd), "wgrad": params(2, wd)}
        self.names = {k: repr(v.conv_algo_desp) for k, v in self.params.items()}
        self.run = ConvMainUnitTest.implicit_gemm2

    def step(self):
        self.run(self.params["fwd"])
        self.run(self.params["dgrad"])
        self.run(self.params["wgrad"])


def time_dtype(layer: Layer, steps: int, warmup: int, stream):
    import torch

    for _ in range(warmup):
        layer.step()
    torch.cuda.synchronize()
    # per-op device time with HIP events on the launch stream
    ev = {k: [] for k in ("fwd", "dgrad", "wgrad")}
    t0 = time.perf_counter()
    for _ in range(steps):
        for k in ("fwd", "dgrad", "wgrad"):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            layer.run(layer.params[k])
            b.record(stream)
            ev[k].append((a, b))
    torch.cuda.synchronize()
    wall_ev = time.perf_counter() - t0
    per_op = {k: float(np.median([a.elapsed_time(b) for a, b in v])) for k, v in ev.items()}
    # whole-step wall clock, synchronize on both sides, no events inside
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        layer.step()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall, per_op


def cpu_baseline(prob, X, W, dO, budget_s: float, gpu_outputs):
    """The reference's CPU path (numpy FP32 gather-GEMM-scatter, compact pairs)
    timed on the host, plus the FP64-oracle error of each GPU result."""
    sys.path.insert(0, str(ROOT))
    from oracle import spconv_ref as ref

    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
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
    # accuracy of the GPU results against the FP64 oracle (original FP32 inputs)
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
    return {
        "value": flops_step / dt / 1e12, "unit": "TFLOP/s", "cores": int(threads), "kind": "port",
        "sample": f"{steps} full fwd+dgrad+wgrad steps of the same 120k-voxel 64->64 layer "
                  f"(numpy FP32, compact pairs), {dt * 1e3:.0f} ms/step",
        "ms_per_step": dt * 1e3,
    }, acc


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
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

    stream = torch.cuda.current_stream()
    sid = stream.cuda_stream
    results = {}
    gpu_outputs = {}
    for dt in args.dtypes.split(","):
        layer = Layer(prob, X, W, dO, dt, sid)
        wall, per_op = time_dtype(layer, args.steps, args.warmup, stream)
        results[dt] = (wall, per_op, layer.names)
        torch.cuda.synchronize()
        gpu_outputs[dt] = (layer.out.double().cpu().numpy(), layer.din.double().cpu().numpy(),
                           layer.dw.double().cpu().numpy(), layer.mask_out.cpu().numpy().view(np.uint32))
        log(f"{dt}: step {wall * 1e3:.3f} ms  " + " ".join(f"{k}={v * 1e3:.1f}us" for k, v in per_op.items()))
        del layer

    main_dt = "f16" if "f16" in results else list(results)[0]
    wall, per_op, names = results[main_dt]
    f32_tf = flops_step / results["f32"][0] / 1e12 if "f32" in results else None
    per_dtype = {}
    for dt, (w, po, nm) in results.items():
        tf = flops_step / w / 1e12
        per_dtype[dt] = {
            "tflops": tf, "ms_per_step": w * 1e3, "frac_mfma_peak": tf * 1e12 / PEAK[dt],
            "speedup_vs_f32": (tf / f32_tf) if f32_tf else None,
            "op_us": {k: v * 1e3 for k, v in po.items()},
            "op_tflops": {k: flops_op / (v * 1e-3) / 1e12 for k, v in po.items()},
        }
    # dominant kernel of the reported step
    dom = max(per_op, key=per_op.get)
    s = DT_BYTES[main_dt]
    byt = op_bytes(dom, prob.n, prob.kv, C, K, s, s)
    t_dom = per_op[dom] * 1e-3
    ach_bw = byt / t_dom
    roofline = {"kernel": names[dom], "op": dom, "bound": "hbm", "achieved": ach_bw / 1e9,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": ach_bw / HBM_PEAK, "traffic": None,
                "algorithmic_bytes": byt, "avg_us": per_op[dom] * 1e3,
                "mfma_tflops": flops_op / t_dom / 1e12}

    cpu = None
    accuracy = None
    if not args.no_cpu_baseline:
        cpu, accuracy = cpu_baseline(prob, X, W, dO, args.cpu_seconds, gpu_outputs)

    line = {
        "metric": "sparse-conv TFLOP/s + % MFMA peak per dtype vs plain FP32; serving p50/p99 ms",
        "value": flops_step / wall / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": main_dt, "data": "synthetic",
        "config": {"workload": f"subm3x3x3 fwd+dgrad+wgrad, {prob.n} voxels on 8 planes in 41x1600x1408, "
                               f"{C}->{K} ch", "voxels": prob.n, "pairs": P, "channels_in": C,
                   "channels_out": K, "kernel_volume": prob.kv, "flops_per_step": flops_step},
        "per_dtype": per_dtype,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy_vs_fp64": accuracy,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
