#!/usr/bin/env python3
"""Batched BytecodeTape fwd+grad throughput on one MI355X.

Default workload: scalar-op tanh MLP f(x) = sum_j w2_j tanh(sum_i W1_ji x_i + b_j),
n = 32 inputs, 64 hidden units (weights from seed 1 baked in as tape
constants, 4288 computed entries), B = 2^20 points x ~ N(0, I), f32,
forward value + reverse gradient per point.  One step = one
gradient_batch over the whole batch with inputs already resident in HBM.

Prints one JSON line.  `roofline` prices the algorithmic bytes of the
slot-per-entry model, S*(6n + 5N + m) per point, against 8 TB/s; the kernel's
average duration comes from HIP events recorded on the launch stream around
every timed launch.  `cpu_baseline` times the CPU oracle (a C++ restatement of
the reference's gradient_batch_par) on a bounded sample of the same points.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--batch-log2", type=int, default=20)
    ap.add_argument("--wg", type=int, default=0)
    ap.add_argument("--lds", type=int, default=0)
    ap.add_argument("--no-fuse", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=32768)
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one 1 GiB device-to-device copy (calibrates HBM byte counters)")
    return ap.parse_args()


def cpu_baseline(tape, x: np.ndarray, sample: int):
    """Oracle fwd+grad on the first `sample` points, all host threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    threads = max(1, min(os.cpu_count() or 1, 16))
    xs = x[:sample].astype(np.float64)
    pyoracle.gradient(tape, xs[:64], nthreads=1)  # warm
    t0 = time.perf_counter()
    pyoracle.gradient(tape, xs, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": sample / dt, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": f"{sample} of the {x.shape[0]} points, f64 oracle fwd+grad, {dt:.2f} s"}


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(points shard by contiguous ranges across GPUs with no exchange)"}))
        return 2
    import torch

    from echidna_amd import workloads as wl
    from echidna_amd.backend import HipContext
    from echidna_amd.tape import GpuTapeData

    f64 = args.dtype == "f64"
    B = 1 << args.batch_log2
    n, H = 32, 64
    tape = wl.tanh_mlp(n, H, 1)
    counts = tape.counts()
    x = wl.points_normal(1, B, n)

    ctx = HipContext.new(0)
    if ctx is None:
        print(json.dumps({"error": "no gfx950 device"}))
        return 1
    if args.wg or args.lds:
        ctx.set_launch(args.wg, args.lds)
    if args.no_fuse:
        ctx.set_fusion(False)
    data = GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape)
    bufs = ctx.upload_tape(data)
    tdt = torch.float64 if f64 else torch.float32
    dx = torch.from_numpy(x.astype(np.float64 if f64 else np.float32)).cuda()
    dy = torch.empty(B, dtype=tdt, device="cuda")
    dg = torch.empty(B, n, dtype=tdt, device="cuda")
    # A dedicated stream: torch's default stream has handle 0, which the
    # library reads as "its own stream"; events must sit on the launch stream.
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    torch.cuda.synchronize()

    if args.copy_probe:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        del src, dst
    tuned = None
    if not (args.wg or args.lds):
        # launch shape from the autotuner, chosen before the timed region
        from echidna_amd.backend import TUNE_GRADIENT
        tuned = ctx.autotune_dev(bufs, TUNE_GRADIENT, B, dx, d_y=dy, d_g=dg)
        torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        for _ in range(args.warmup):
            ctx.gradient_batch_dev(bufs, dx, B, dy, dg, stream=sh)
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        ctx.gradient_batch_dev(bufs, dx, B, dy, dg, stream=sh)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))

    S = 8 if f64 else 4
    bytes_per_point = S * (6 * counts["n"] + 5 * counts["N"] + counts["m"])
    achieved = bytes_per_point * B / (kernel_ms * 1e-3)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_mlp32_fwdgrad_{args.dtype}.json")
    if os.path.exists(tpath) and args.batch_log2 == 20:
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": "fwd+grad point-evals/s",
        "value": B * args.steps / wall,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic: Philox points x~N(0,I), random-init weights baked into the tape",
        "config": {"workload": f"tanh-MLP tape n={n} H={H} seed 1 (N={counts['N']} computed entries), "
                               f"B=2^{args.batch_log2} points, fwd+grad",
                   "points": B, "inputs": n, "hidden": H, "computed_entries": counts["N"],
                   "program": bufs.program_info(1), "parallelism": "none (1 GPU)",
                   "launch": tuned or {"wg": args.wg, "lds_budget": args.lds}},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel_ms": kernel_ms, "algorithmic_bytes_per_point": bytes_per_point},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(tape, x, min(args.cpu_sample, B))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
