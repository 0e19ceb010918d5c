#!/usr/bin/env python3
"""Batched BytecodeTape evaluation on one MI355X.

Default line (`python bench.py`): the headline is forward value + reverse
gradient of the scalar-op tanh MLP f(x) = sum_j w2_j tanh(sum_i W1_ji x_i
+ b_j), n = 32 inputs, 64 hidden units (weights from seed 1 baked in as tape
constants, 4,288 computed entries), B = 2^20 points x ~ N(0, I), f32, inputs
resident in HBM; one step = one gradient_batch over the whole batch.  The
same line carries the rest of the metric, each measured in this run with
its error against the FP64 oracle and a CPU baseline:

  fanin_f64  the headline evaluation with binary64 adjoint accumulation
           (FP32 forward and partials, each adjoint's fan-in summed in
           binary64), its rate and its error against the oracle;
  f64      the same fwd+grad in f64;
  hvp      one Hessian-vector product per point on the same tape, f32;
  stde     plain GPU STDE on the n=256 H=256 MLP (seed 2), B = 65,536,
           S = 64 Rademacher and Gaussian, f32 jets, f64 reduction;
           Laplacians/s at 1e-3 relative standard error against exact
           f64 traces;
  cv       the learned control variate against plain STDE on the n=1024
           2x512 MLP (seed 4), B = 16,384, S = 16 Gaussian, trained on
           2^20 tuples from the tape.

`roofline` is the headline kernel's: its algorithmic arithmetic per point
(echidna_amd.costmodel, counted from the uploaded tape) x points / the
kernel's average duration (HIP events on its launch stream) against the
FP32 vector peak; HBM traffic from committed counter profiles where the
launch shape matches, and the compulsory HBM rate beside it.

`cpu_baseline` times the CPU oracle (a C++ restatement of the reference's
gradient_batch_par, src/bytecode_tape/parallel.rs:240-253) on the host:
one thread and a pool of threads over points, each the median of 5 warm
runs on a bounded sample of the same points.  The same leg compares the GPU
results of those points with the oracle's (`error_vs_oracle`).
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
FP32_PEAK = 157.3e12
FP64_PEAK = 78.6e12
# host threads for the CPU baseline: the host's CPU share for one GPU
CPU_POOL_MAX = 16


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--settle-s", type=float, default=0.25,
                    help="untimed launches of the timed evaluation for this long before the warm-up steps, so the "
                         "GPU's clocks have ramped (the autotuner's trials did this when no cached shape was used)")
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--batch-log2", type=int, default=20)
    ap.add_argument("--wg", type=int, default=0)
    ap.add_argument("--lds", type=int, default=0)
    ap.add_argument("--fetch", type=int, choices=[0, 1, 2], default=0,
                    help="with --wg/--lds: program fetch, 0 LDS windows, 1 scalar cache, "
                         "2 tape-specialised kernel (mlp)")
    ap.add_argument("--ppl", type=int, default=0,
                    help="with --wg and --fetch 2: points per lane of the tape-specialised kernel")
    ap.add_argument("--no-spec", action="store_true",
                    help="mlp, hvp: leave the tape-specialised kernel out (interpreters only)")
    ap.add_argument("--no-fuse", action="store_true")
    ap.add_argument("--no-jet-mma", action="store_true",
                    help="stde, cv: order-3 f32 jets on the exact per-term DENSE path instead of FP32 MFMA")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle error check")
    ap.add_argument("--headline-only", action="store_true",
                    help="default workload: only the fwd+grad headline (no f64/hvp/stde/cv keys)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run one 1 GiB device-to-device copy (calibrates HBM byte counters)")
    ap.add_argument("--workload", choices=["mlp", "hvp", "stde", "cv", "dags"], default="mlp",
                    help="mlp: the default line; hvp: one Hessian-vector product per point on the "
                         "same MLP; stde: plain GPU STDE on the n=256 H=256 MLP; "
                         "cv: learned control variate vs plain on the n=1024 2x512 MLP; dags: 64 uneven "
                         "random DAGs in one persistent launch vs one launch per tape")
    ap.add_argument("--points", type=int, default=0, help="override the workload's point count")
    ap.add_argument("--tuples-log2", type=int, default=20, help="cv: training tuples = 2^k")
    ap.add_argument("--epochs", type=int, default=10, help="cv: training epochs")
    ap.add_argument("--cv-reps", type=int, default=5,
                    help="cv: timed passes of each estimator after training (the best is kept for both)")
    ap.add_argument("--dag-scale", type=float, default=1.0, help="dags: shrink lengths and point counts")
    ap.add_argument("--tune-cache", default=os.path.join(ROOT, "echidna_amd", "tune_cache.json"),
                    help="persisted autotuner shapes to restore (keyed by tape, dtype, evaluation, batch)")
    ap.add_argument("--tune-cache-out", default="", help="write the cache here after the run (default: not written)")
    ap.add_argument("--no-tune-cache", action="store_true", help="autotune every launch in this run")
    return ap.parse_args()


CACHE = None  # echidna_amd.tunecache.TuneCache, set by main


def tune(ctx, bufs, data, f64: bool, what: int, batch: int, d_x, order: int = 0, extra: str = "", **kw) -> dict:
    """The launch shape for one evaluation: restored from the tune cache, or
    autotuned (and recorded in it)."""
    if CACHE is None:
        return dict(ctx.autotune_dev(bufs, what, batch, d_x, order=order, **kw), cached=False)
    key = CACHE.key(data, f64, what, order, batch) + extra
    return CACHE.autotune(ctx, bufs, key, what, batch, d_x, order=order, **kw)


def log(msg: str) -> None:
    print(f"# {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# CPU baseline (the oracle on the host) and the oracle error check
# ---------------------------------------------------------------------------

def host_info() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "affinity": share, "cpu_model": model}


def median_rate(fn, units: int, runs: int = 5) -> float:
    """Median of `runs` warm runs of fn() (one warm-up first), in units/s."""
    fn()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return units / statistics.median(times)


def cpu_rates(make_fn, single_units: int, pool_units: int, unit: str, kind_sample: str) -> dict:
    """Single-thread and thread-pool rates of the oracle: make_fn(threads,
    units) returns a callable doing `units` of work."""
    info = host_info()
    threads = max(1, min(info["affinity"], CPU_POOL_MAX))
    one = median_rate(make_fn(1, single_units), single_units)
    pool = median_rate(make_fn(threads, pool_units), pool_units) if threads > 1 else one
    return {"value": pool, "unit": unit, "cores": threads, "kind": "port", "single_thread_value": one,
            "statistic": "median of 5 warm runs", **info,
            "sample": f"{kind_sample}: {single_units} per run on 1 thread, {pool_units} per run on {threads} "
                      f"threads (the host's CPU share for one GPU, of nproc {info['nproc']})"}


def fwdgrad_error(tape, x: np.ndarray, y: np.ndarray, g: np.ndarray, f64: bool) -> dict:
    """The GPU values and gradients of a strided sample of points against
    the f64 oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    B, n = x.shape
    idx = np.unique(np.concatenate([np.arange(0, B, max(1, B // 4096)), [B - 1]]))
    xs = x[idx].astype(np.float64)
    y_ref, g_ref = pyoracle.gradient(tape, xs, nthreads=CPU_POOL_MAX)
    mag = pyoracle.magnitudes(tape, xs, nthreads=CPU_POOL_MAX)
    gy, gg = np.asarray(y, np.float64)[idx], np.asarray(g, np.float64)[idx]
    err = {"points": int(len(idx)), "bar": 1e-12 if f64 else 2e-5,
           "max_rel_err_vs_oracle": float(max(np.max(np.abs(gy - y_ref[:, 0]) / np.maximum(np.abs(y_ref[:, 0]), 1e-300)),
                                              np.max(np.abs(gg - g_ref) / np.maximum(np.abs(g_ref), 1e-300)))),
           "max_err_over_term_magnitude": float(max(
               np.max(np.abs(gy - y_ref[:, 0]) / np.maximum(np.abs(y_ref[:, 0]), mag["value"][:, 0])),
               np.max(np.abs(gg - g_ref) / np.maximum(np.abs(g_ref), mag["grad"]))))}
    err["within_bar"] = bool(err["max_err_over_term_magnitude"] <= err["bar"])
    return err


def oracle_leg_fwdgrad(tape, x: np.ndarray, y: np.ndarray, g: np.ndarray, f64: bool) -> tuple:
    """CPU baseline of fwd+grad and the GPU results of a strided sample of
    points against the f64 oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    err = fwdgrad_error(tape, x, y, g, f64)
    xb = np.ascontiguousarray(x[:1 << 18].astype(np.float64))

    def make(threads, units):
        return lambda: pyoracle.gradient(tape, xb[:units], nthreads=threads)

    cpu = cpu_rates(make, 8192, min(len(xb), 1 << 17), "points/s", "f64 oracle fwd+grad over the same points")
    return cpu, err


# ---------------------------------------------------------------------------
# the headline: fwd+grad on the n=32 H=64 tanh MLP, B = 2^20
# ---------------------------------------------------------------------------

def run_fwdgrad(args, dtype: str, headline: bool) -> dict:
    import torch

    from echidna_amd import costmodel
    from echidna_amd import workloads as wl
    from echidna_amd.backend import SPEC_GRAD_ACC64, TUNE_GRADIENT, HipContext
    from echidna_amd.tape import GpuError, GpuTapeData

    f64 = dtype == "f64"
    B = 1 << args.batch_log2
    n, H = 32, 64
    tape = wl.tanh_mlp(n, H, 1)
    counts = tape.counts()
    x = wl.points_normal(1, B, n)
    ctx = HipContext.new(0)
    if ctx is None:
        raise RuntimeError("no gfx950 device")
    if args.wg or args.lds:
        ctx.set_launch(args.wg, args.lds)
        if args.fetch != 2:
            ctx.set_fetch(args.fetch)
    if args.no_fuse:
        ctx.set_fusion(False)
    data = GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape)
    bufs = ctx.upload_tape(data)
    spec = None
    if not args.no_spec:
        # straight-line kernel for this tape (hipRTC), built before any
        # timing; should the compile fail the interpreters run and the line
        # says why
        try:
            spec = ctx.specialize(bufs, TUNE_GRADIENT)
        except GpuError as e:
            spec = {"error": str(e)}
    if args.fetch == 2 and args.wg:
        ctx.set_tuned(bufs, TUNE_GRADIENT, args.wg, 0, fetch=2, points_per_lane=args.ppl)
    tdt = torch.float64 if f64 else torch.float32
    dx = torch.from_numpy(x.astype(np.float64 if f64 else np.float32)).cuda()
    dy = torch.empty(B, dtype=tdt, device="cuda")
    dg = torch.empty(B, n, dtype=tdt, device="cuda")
    # a stream of its own: torch's default stream has handle 0, which the
    # library reads as "its own stream"; events must sit on the launch stream
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    torch.cuda.synchronize()
    if args.copy_probe and headline:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        del src, dst
    tuned = None
    if not (args.wg or args.lds):
        tuned = tune(ctx, bufs, data, f64, TUNE_GRADIENT, B, dx, extra="" if spec and "error" not in spec else ":nospec",
                     d_y=dy, d_g=dg)
        torch.cuda.synchronize()
    settle = 0
    t_settle = time.perf_counter()
    with torch.cuda.stream(stream):
        while time.perf_counter() - t_settle < args.settle_s:
            for _ in range(10):
                ctx.gradient_batch_dev(bufs, dx, B, dy, dg, stream=sh)
            settle += 10
            torch.cuda.synchronize()
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
    out = {"value": B * args.steps / wall, "ms_per_step": wall * 1e3 / args.steps, "kernel_ms": kernel_ms,
           "settle_launches": settle}
    if headline:
        # the same steps replayed from one captured hipGraph (reported beside)
        graph = ctx.graph_capture(bufs, TUNE_GRADIENT, B, dx, d_y=dy, d_g=dg, repeats=args.steps)
        graph.launch()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        graph.launch()
        torch.cuda.synchronize()
        out["graph_ms_per_step"] = (time.perf_counter() - t1) * 1e3 / args.steps
        graph.close()
    S = 8 if f64 else 4
    flop = costmodel.fwdgrad_flops(tape)
    peak = FP64_PEAK if f64 else FP32_PEAK
    achieved = flop * B / (kernel_ms * 1e-3)
    compulsory = costmodel.compulsory_bytes(tape, "gradient", S)
    launch = tuned or {"wg": args.wg, "lds_budget": args.lds, "fetch": args.fetch,
                       "points_per_lane": max(1, args.ppl) if args.fetch == 2 else 1}
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_mlp32_fwdgrad_{dtype}.json")
    if os.path.exists(tpath) and args.batch_log2 == 20:
        with open(tpath) as f:
            prof = json.load(f)
        # counters hold for the launch shape (and build) they were taken with
        shape = {k: launch.get(k) for k in ("wg", "lds_budget", "fetch")}
        shape["points_per_lane"] = max(1, launch.get("points_per_lane") or 1)
        for entry in (prof if isinstance(prof, list) else [prof]):
            want = dict(entry.get("launch", {}))
            want.setdefault("points_per_lane", 1)
            if want == shape:
                traffic = entry.get("hbm_bytes_per_launch")
    out["roofline"] = {"bound": "fp64 valu" if f64 else "fp32 valu", "achieved": achieved / 1e12,
                       "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                       "kernel_ms": kernel_ms, "flop_per_point": flop,
                       "hbm": {"compulsory_bytes_per_point": compulsory,
                               "compulsory_GBps": compulsory * B / (kernel_ms * 1e-3) / 1e9,
                               "frac_of_8TBps": compulsory * B / (kernel_ms * 1e-3) / HBM_PEAK,
                               "traffic_over_compulsory": (traffic / (compulsory * B)) if traffic else None}}
    out["config"] = {"workload": f"tanh-MLP tape n={n} H={H} seed 1 (N={counts['N']} computed entries), "
                                 f"B=2^{args.batch_log2} points, fwd+grad",
                     "points": B, "inputs": n, "hidden": H, "computed_entries": counts["N"],
                     "program": bufs.program_info(1), "parallelism": "none (1 GPU)", "launch": launch,
                     "specialized": spec}
    if not args.no_cpu:
        torch.cuda.synchronize()
        cpu, err = oracle_leg_fwdgrad(tape, x.astype(np.float64 if f64 else np.float32).astype(np.float64),
                                      dy.cpu().numpy(), dg.cpu().numpy(), f64)
        out["cpu_baseline"] = cpu
        out["error_vs_oracle"] = err
    if headline and not f64 and spec and "error" not in spec:
        # the same evaluation with binary64 adjoint accumulation (the FP32
        # forward sweep and partials, each adjoint's fan-in summed in
        # binary64), at the tuned launch shape
        out["fanin_f64"] = {"specialized": ctx.specialize(bufs, SPEC_GRAD_ACC64)}
        with torch.cuda.stream(stream):
            for _ in range(args.warmup):
                ctx.gradient_batch_dev(bufs, dx, B, dy, dg, stream=sh)
        torch.cuda.synchronize()
        for k in range(args.steps):
            starts[k].record(stream)
            ctx.gradient_batch_dev(bufs, dx, B, dy, dg, stream=sh)
            ends[k].record(stream)
        torch.cuda.synchronize()
        kms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
        out["fanin_f64"].update({"value": B / (kms * 1e-3), "unit": "points/s", "kernel_ms": kms})
        if not args.no_cpu:
            out["fanin_f64"]["error_vs_oracle"] = fwdgrad_error(
                tape, x.astype(np.float32).astype(np.float64), dy.cpu().numpy(), dg.cpu().numpy(), False)
    ctx.close()
    return out


def run_hvp(args, compact: bool = False) -> dict:
    """Forward-over-reverse HVP, one (x, v) per point: tanh MLP n=32 H=64
    seed 1, B=2^20, x, v ~ N(0, I) (config 2's third evaluation)."""
    import torch

    from echidna_amd import costmodel, philox
    from echidna_amd import workloads as wl
    from echidna_amd.backend import TUNE_HVP, HipContext
    from echidna_amd.tape import GpuError, GpuTapeData

    n, H = 32, 64
    B = 1 << args.batch_log2
    f64 = args.dtype == "f64" and not compact
    tape = wl.tanh_mlp(n, H, 1)
    ctx = HipContext.new(0)
    data = GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape)
    bufs = ctx.upload_tape(data)
    dt = np.float64 if f64 else np.float32
    xh = wl.points_normal(1, B, n).astype(dt)
    vh = philox.normal(1, philox.STREAM_DIRS, np.arange(B * n)).reshape(B, n).astype(dt)
    x = torch.from_numpy(xh).cuda()
    v = torch.from_numpy(vh).cuda()
    g = torch.empty_like(x)
    hv = torch.empty_like(x)
    spec = None
    if not args.no_spec:
        try:
            spec = ctx.specialize(bufs, TUNE_HVP)
        except GpuError as e:
            spec = {"error": str(e)}
    tuned = tune(ctx, bufs, data, f64, TUNE_HVP, B, x, extra="" if spec and "error" not in spec else ":nospec",
                 d_g=g, d_h=hv, d_dirs=v)
    dt_s, _ = timed(lambda: ctx.hvp_batch_dev(bufs, x, n, v, B, g, hv), max(1, args.steps))
    flop = costmodel.hvp_flops(tape)
    peak = FP64_PEAK if f64 else FP32_PEAK
    out = {"value": B / dt_s, "unit": "points/s", "dtype": "f64" if f64 else "f32", "ms_per_step": dt_s * 1e3,
           "launch": tuned, "specialized": spec,
           "valu": {"flop_per_point": flop, "achieved_TFLOPs": flop * B / dt_s / 1e12,
                    "frac": flop * B / dt_s / peak}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle

        idx = np.arange(0, B, B // 1024)
        xs, vs = xh[idx].astype(np.float64), vh[idx].astype(np.float64)
        g_ref, h_ref = pyoracle.hvp(tape, xs, vs, nthreads=CPU_POOL_MAX)
        mag = pyoracle.magnitudes(tape, xs, vs, nthreads=CPU_POOL_MAX)
        hg = hv.cpu().numpy()[idx].astype(np.float64)
        out["error_vs_oracle"] = {"points": len(idx), "bar": 1e-12 if f64 else 2e-5,
                                  "max_err_over_term_magnitude": float(np.max(
                                      np.abs(hg - h_ref) / np.maximum(np.abs(h_ref), mag["hv"]))),
                                  "max_rel_err_vs_oracle": float(np.max(np.abs(hg - h_ref)
                                                                        / np.maximum(np.abs(h_ref), 1e-300)))}
        xb, vb = xh[:1 << 16].astype(np.float64), vh[:1 << 16].astype(np.float64)

        def make(threads, units):
            return lambda: pyoracle.hvp(tape, xb[:units], vb[:units], nthreads=threads)

        out["cpu_baseline"] = cpu_rates(make, 2048, 1 << 15, "points/s", "f64 oracle HVP (Dual forward-over-reverse)")
    if compact:
        ctx.close()
        return out
    out.update({"metric": "HVP point-evals/s", "n_gpus": 1, "higher_is_better": True, "vs_baseline": None,
                "data": "synthetic: Philox points and directions, random-init weights baked into the tape",
                "config": {"workload": f"tanh-MLP tape n={n} H={H} seed 1, B=2^{args.batch_log2} points, one HVP "
                                       f"each", "points": B}})
    ctx.close()
    return out


# ---------------------------------------------------------------------------
# STDE and the learned control variate
# ---------------------------------------------------------------------------

def s_req(var: np.ndarray, lap: np.ndarray, rel: float = 1e-3) -> np.ndarray:
    """Directions each point needs for a standard error of rel * |Laplacian|:
    sigma^2 / (rel * Laplacian)^2, with the exact f64 Laplacian as the
    denominator (an estimate there would put each method's own noise into
    the other's comparison, worst where the Laplacian is near zero)."""
    return var / np.maximum((rel * np.abs(lap)) ** 2, 1e-300)


def lap_rate(var: np.ndarray, lap: np.ndarray, t_sample: float) -> dict:
    """Laplacians/s at 1e-3 relative standard error over a set of points:
    one Laplacian costs S_req * t_sample; the mean over points is the
    throughput of the whole set, the median that of a typical point."""
    sr = s_req(var, lap)
    return {"mean_S_req": float(np.mean(sr)), "median_S_req": float(np.median(sr)),
            "laplacians_per_s": float(1.0 / (np.mean(sr) * t_sample)),
            "median_point_laplacians_per_s": float(1.0 / (np.median(sr) * t_sample))}


def tune_jets(ctx, bufs, data, X, lanes: int, mma: bool) -> dict:
    """Autotune the order-3 jet launch (the STDE samples) on `lanes` lanes:
    rows of X repeated, Gaussian directions."""
    import torch

    from echidna_amd.backend import TUNE_TAYLOR

    n = X.shape[1]
    rows = X[torch.arange(lanes, device=X.device) % X.shape[0]].contiguous()
    dirs = torch.randn(lanes, n, dtype=X.dtype, device=X.device)
    coeff = torch.empty(lanes, bufs.num_outputs * 3, dtype=X.dtype, device=X.device)
    res = tune(ctx, bufs, data, X.dtype == torch.float64, TUNE_TAYLOR, lanes, rows, order=3,
               extra=":mma" if mma else "", d_y=coeff, d_dirs=dirs)
    torch.cuda.synchronize()
    return res


def timed(fn, reps: int):
    import torch

    fn()
    torch.cuda.synchronize()
    best = float("inf")
    out = None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best, out


def stde_oracle_leg(tape, X: np.ndarray, res: np.ndarray, S: int, seed: int, kind: int, points) -> tuple:
    """The GPU estimates of `points` against the oracle's
    laplacian_with_stats on the same directions (regenerated from Philox),
    and the oracle's order-3 jets/s on the host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    from echidna_amd import cv as lcv
    from echidna_amd import philox

    n = X.shape[1]
    worst = 0.0
    for p in points:
        idx = np.arange(p * S * n, (p + 1) * S * n, dtype=np.uint64)
        dirs = philox.direction(kind, seed, idx).reshape(S, n).astype(np.float32).astype(np.float64)
        ref = pyoracle.laplacian_with_stats(tape, X[p].astype(np.float64), dirs)
        worst = max(worst, abs(res[p, 1] - ref["estimate"]) / max(abs(ref["estimate"]), ref["standard_error"]))
    err = {"points": len(points), "max_err_over_max_estimate_se": float(worst), "bar": 1e-4}
    xs = np.repeat(X[:64].astype(np.float64), S, axis=0)
    vs = philox.direction(kind, seed, np.arange(xs.size, dtype=np.uint64)).reshape(-1, n)

    def make(threads, units):
        return lambda: pyoracle.taylor(tape, xs[:units], vs[:units], K=3, nthreads=threads)

    big = tape.counts()["N"] > 500000
    cpu = cpu_rates(make, 8 if big else 64, 128 if big else 1024, "jets/s",
                    "f64 oracle order-3 Taylor jets (the reference's taylor_jet_2nd per direction)")
    return cpu, err


def run_stde(args, compact: bool = False) -> dict:
    """Plain GPU STDE: tanh MLP n=256 H=256 (seed 2), x ~ N(0, I), S = 64
    Rademacher and Gaussian directions per point, f32 jets, f64 reduction;
    error against the exact trace (f64 basis jets) on a 1024-point subset."""
    import torch

    from echidna_amd import costmodel
    from echidna_amd import cv as lcv
    from echidna_amd import workloads as wl
    from echidna_amd.backend import HipContext
    from echidna_amd.tape import GpuTapeData

    n, H, S = 256, 256, 64
    P = args.points or 65536
    tape = wl.tanh_mlp(n, H, 2)
    ctx = HipContext.new(0)
    ctx.set_jet_mma(not args.no_jet_mma)
    data = GpuTapeData.from_tape(tape)
    bufs = ctx.upload_tape(data)
    Xh = wl.points_normal(2, P, n).astype(np.float32)
    X = torch.from_numpy(Xh).cuda()
    tuned = tune_jets(ctx, bufs, data, X, 1 << 20, not args.no_jet_mma)
    # jet arithmetic per direction: the per-direction coefficients plus the
    # point's primal shared by its S directions (costmodel.jet_flops_split)
    f_point, f_lane = costmodel.jet_flops_split(tape, 3)
    flop = f_lane + f_point / S
    out = {"value": None, "unit": "Laplacians/s", "dtype": "f32 jets, f64 reduction",
           "jet_dense": "exact per-term" if args.no_jet_mma else "fp32 mfma",
           "config": {"workload": f"plain GPU STDE, tanh-MLP tape n={n} H={H} seed 2 ({tape.counts()['N']} "
                                  f"computed entries), B={P} points, S={S}", "points": P, "S": S,
                      "launch": tuned}}
    # exact traces on a subset: f64 tape, n basis jets per point
    k = min(P, 1024)
    b64 = ctx.upload_tape(GpuTapeData.from_tape_f64(tape))
    X64 = torch.from_numpy(Xh[:k].astype(np.float64)).cuda()
    t0 = time.perf_counter()
    exact = lcv.hessian_trace_batch(ctx, b64, X64)
    t_exact = time.perf_counter() - t0
    results = {}
    for name, kind in (("rademacher", lcv.DIR_RADEMACHER), ("gaussian", lcv.DIR_GAUSSIAN)):
        dt, res = timed(lambda: lcv.laplacian_batch(ctx, bufs, X, S, kind, 2), max(1, args.steps // 10))
        results[name] = res
        t_sample = dt / (P * S)
        out[name] = {"seconds": dt, "jets_per_s": P * S / dt, "t_sample_ns": t_sample * 1e9,
                     "fp32_frac": flop * P * S / dt / FP32_PEAK,
                     **lap_rate(res[:k, 2], exact, t_sample),
                     "mean_rel_se_at_S": float(np.mean(res[:k, 3] / np.abs(exact)))}
    rad = results["rademacher"]
    z = (rad[:k, 1] - exact) / np.maximum(rad[:k, 3], 1e-300)
    out["exact_check"] = {"points": k, "seconds": t_exact, "mean_z": float(np.mean(z)),
                          "mean_abs_z": float(np.mean(np.abs(z))), "frac_within_3se": float(np.mean(np.abs(z) < 3.0)),
                          "mean_rel_err": float(np.mean(np.abs(rad[:k, 1] - exact) / np.abs(exact)))}
    out["value"] = out["rademacher"]["laplacians_per_s"]
    out["flop_per_jet"] = {"per_direction": f_lane, "per_point": f_point, "charged": flop,
                           "tape_every_coefficient": costmodel.jet_flops(tape, 3)}
    if not args.no_cpu:
        out["cpu_baseline"], out["error_vs_oracle"] = stde_oracle_leg(tape, Xh, rad, S, 2, lcv.DIR_RADEMACHER,
                                                                      (0, P // 2, P - 1))
    ctx.close()
    if not compact:
        out.update({"metric": "Laplacians/s at 1e-3 rel err", "n_gpus": 1, "higher_is_better": True,
                    "vs_baseline": None, "data": "synthetic: Philox points, random-init weights baked into the tape"})
    return out


def run_cv(args, compact: bool = False) -> dict:
    """Learned control variate vs plain STDE: tanh MLP n=1024, 2x512 hidden
    (seed 4), S = 16 Gaussian directions, x ~ N(0, I); CV network
    1024 -> 256 -> 256 -> 1024*16 trained on 2^k tuples from the tape."""
    import torch

    from echidna_amd import cv as lcv
    from echidna_amd import workloads as wl
    from echidna_amd.backend import HipContext
    from echidna_amd.tape import GpuTapeData

    n, H, S, r = 1024, 512, 16, 16
    P = args.points or 16384
    t0 = time.perf_counter()
    tape = wl.tanh_mlp2(n, H, H, 4)
    t_record = time.perf_counter() - t0
    ctx = HipContext.new(0)
    ctx.set_jet_mma(not args.no_jet_mma)
    t0 = time.perf_counter()
    data = GpuTapeData.from_tape(tape)
    bufs = ctx.upload_tape(data)
    t_upload = time.perf_counter() - t0
    Xh = wl.points_normal(4, P, n).astype(np.float32)
    X = torch.from_numpy(Xh).cuda()
    tuned = tune_jets(ctx, bufs, data, X, 1 << 16, not args.no_jet_mma)
    log(f"recorded {tape.counts()['N']} entries in {t_record:.1f} s, uploaded in {t_upload:.1f} s")
    dt_plain, plain = timed(lambda: lcv.laplacian_batch(ctx, bufs, X, S, lcv.DIR_GAUSSIAN, 5), 1)
    log(f"plain STDE {dt_plain:.2f} s")
    T = 1 << args.tuples_log2
    # signed factor (8 columns +1, 8 columns -1): the MLP's Hessian is indefinite
    est = lcv.LearnedStde(ctx, bufs, rank=r, hidden=256, seed=4, npos=r // 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, v, hv = lcv.make_tuples(ctx, bufs, T, 5)
    t_tuples = time.perf_counter() - t0
    del x, v, hv
    t0 = time.perf_counter()
    # start near C = 0 (an untrained network neither adds nor removes
    # variance); cosine-decayed Adam over the tuples
    est.fit(ntuples=T, epochs=args.epochs, batch=1024, lr=2e-3, val_points=256, val_S=S, out_scale=0.03,
            schedule="cosine")
    t_fit = time.perf_counter() - t0 - t_tuples  # fit regenerates the tuples
    log(f"trained in {t_fit:.2f} s, validation variance ratio {est.variance_ratio:.4f}")
    # both estimators timed after training, in the same clock state, best of
    # the same number of passes
    reps = max(1, args.cv_reps)
    dt_plain, _ = timed(lambda: lcv.laplacian_batch(ctx, bufs, X, S, lcv.DIR_GAUSSIAN, 5), reps)
    dt_cv, both = timed(lambda: lcv.laplacian_batch(ctx, bufs, X, S, lcv.DIR_GAUSSIAN, 5, est.cv, with_plain=True),
                        reps)
    log(f"plain pass {dt_plain * 1e3:.2f} ms, pass with the control variate {dt_cv * 1e3:.2f} ms (best of {reps})")
    res_cv, res_pl = both
    t_s_plain = dt_plain / (P * S)
    t_s_cv = dt_cv / (P * S)
    k = min(P, 1024)
    b64 = ctx.upload_tape(GpuTapeData.from_tape_f64(tape))
    t0 = time.perf_counter()
    exact = lcv.hessian_trace_batch(ctx, b64, X[:k].double())
    t_exact = time.perf_counter() - t0
    rate_plain = lap_rate(res_pl[:k, 2], exact, t_s_plain)
    rate_cv = lap_rate(res_cv[:k, 2], exact, t_s_cv)
    lps_plain = rate_plain["laplacians_per_s"]
    lps_cv = rate_cv["laplacians_per_s"] if est.use_cv else lps_plain
    t0 = time.perf_counter()
    bound = rank_bound(ctx, b64, est.cv, X[:16], r, r // 2)
    bound["seconds"] = time.perf_counter() - t0
    log(f"rank-{r} bound: best variance ratio {bound['best_ratio']:.4f} (fixed signs "
        f"{bound['best_ratio_signed']:.4f}), trained network {bound['network_ratio']:.4f}")
    zc = (res_cv[:k, 1] - exact) / np.maximum(res_cv[:k, 3], 1e-300)
    zp = (res_pl[:k, 1] - exact) / np.maximum(res_pl[:k, 3], 1e-300)
    out = {"value": float(lps_cv), "unit": "Laplacians/s", "dtype": "f32 jets, f64 reduction, bf16 MFMA CV",
           "jet_dense": "exact per-term" if args.no_jet_mma else "fp32 mfma",
           "config": {"workload": f"tanh-MLP tape n={n} 2x{H} seed 4 ({tape.counts()['N']} computed entries), "
                                  f"B={P} points, S={S} Gaussian, CV {n}->256->256->{n}x{r} signed {r // 2}+{r // 2}",
                      "points": P, "S": S, "rank": r, "tuples": T, "epochs": args.epochs, "launch": tuned,
                      "timing": f"best of {max(1, args.cv_reps)} passes each, both after training"},
           "plain": {**rate_plain, "seconds": dt_plain, "t_sample_ns": t_s_plain * 1e9,
                     "jets_per_s": P * S / dt_plain, "mean_var": float(np.mean(res_pl[:, 2]))},
           "learned_cv": {**rate_cv, "seconds": dt_cv, "t_sample_ns": t_s_cv * 1e9,
                          "mean_var": float(np.mean(res_cv[:, 2])),
                          "variance_ratio": float(np.mean(res_cv[:, 2]) / np.mean(res_pl[:, 2])),
                          "validation_variance_ratio": est.variance_ratio, "used": est.use_cv,
                          "speedup": float(lps_cv / lps_plain),
                          "median_point_speedup": float(rate_cv["median_point_laplacians_per_s"]
                                                        / rate_plain["median_point_laplacians_per_s"])},
           "training": {"tuple_seconds": t_tuples, "train_seconds": t_fit, "loss_first": float(est.history[0]),
                        "loss_last": float(est.history[-1])},
           "exact_check": {"points": k, "seconds": t_exact, "mean_z_cv": float(np.mean(zc)),
                           "mean_abs_z_cv": float(np.mean(np.abs(zc))),
                           "mean_abs_z_plain": float(np.mean(np.abs(zp)))},
           "rank_bound": bound, "setup": {"record_s": t_record, "upload_s": t_upload}}
    if not args.no_cpu:
        out["cpu_baseline"], out["error_vs_oracle"] = stde_oracle_leg(tape, Xh, res_pl, S, 5, lcv.DIR_GAUSSIAN,
                                                                      (0, P - 1))
    ctx.close()
    if not compact:
        out.update({"metric": "Laplacians/s at 1e-3 rel err, learned CV vs plain", "n_gpus": 1,
                    "higher_is_better": True, "vs_baseline": None,
                    "data": "synthetic: Philox points and tuples"})
    return out


def rank_bound(ctx, b64, cvnet, X, r: int, npos: int) -> dict:
    """What any rank-r control variate could remove on these points.

    The exact f64 Hessian of each point comes from n GPU HVPs along the basis
    vectors.  For Gaussian directions Var(v^T A v) = 2 |A|_F^2, so the best
    rank-r C = L S L^T leaves |H - C|_F^2 = sum of all but the r largest
    squared eigenvalues (Eckart-Young); with S fixed to npos positive and
    r - npos negative signs, the npos largest positive and r - npos most
    negative eigenvalues are removed.  The trained network's own C(x) is
    scored the same way."""
    k, n = X.shape
    xs = np.repeat(X.double().cpu().numpy(), n, axis=0)
    dirs = np.tile(np.eye(n), (k, 1))
    _, hv = ctx.hvp_batch_points(b64, xs, dirs, k * n)
    Hs = hv.reshape(k, n, n)
    Hs = 0.5 * (Hs + Hs.transpose(0, 2, 1))
    L = cvnet.forward(X.float()).double().cpu().numpy().reshape(k, n, r)
    sgn = np.where(np.arange(r) < npos, 1.0, -1.0)
    fro2 = res_any = res_signed = res_net = 0.0
    for p in range(k):
        lam = np.linalg.eigvalsh(Hs[p])
        sq = np.sort(lam * lam)[::-1]
        pos = np.sort(lam[lam > 0])[::-1]
        neg = np.sort(-lam[lam < 0])[::-1]
        f2 = float(np.sum(sq))
        fro2 += f2
        res_any += f2 - float(np.sum(sq[:r]))
        res_signed += f2 - float(np.sum(pos[:npos] ** 2) + np.sum(neg[:r - npos] ** 2))
        C = (L[p] * sgn) @ L[p].T
        res_net += float(np.sum((Hs[p] - C) ** 2))
    return {"points": k, "best_ratio": res_any / fro2, "best_ratio_signed": res_signed / fro2,
            "network_ratio": res_net / fro2}


def run_dags(args) -> dict:
    """64 random DAGs (seed 3): lengths log-U[200, 20000], n ~ U[4, 128],
    B_i log-U[1e3, 1e5], x ~ U[-1, 1]^n, fwd+grad in --dtype: one
    persistent launch over the (tape, point-tile) queue vs one launch per
    tape."""
    import torch

    from echidna_amd import workloads as wl
    from echidna_amd.backend import HipContext
    from echidna_amd.tape import GpuTapeData

    specs = wl.random_dag_suite(3, 64, args.dag_scale)
    f64 = args.dtype == "f64"
    dt, tdt = (np.float64, torch.float64) if f64 else (np.float32, torch.float32)
    ctx = HipContext.new(0)
    jobs, ops = [], 0
    t0 = time.perf_counter()
    for k, (length, n, B) in enumerate(specs):
        tape = wl.random_dag(3, k, length, n)
        bufs = ctx.upload_tape(GpuTapeData.from_tape_f64(tape) if f64 else GpuTapeData.from_tape(tape))
        x = torch.from_numpy(wl.points_uniform(300 + k, B, n, -1.0, 1.0).astype(dt)).cuda()
        jobs.append((bufs, 1, x, B, torch.empty(B, dtype=tdt, device="cuda"),
                     torch.empty(B, n, dtype=tdt, device="cuda")))
        ops += B * tape.counts()["N"]
    t_setup = time.perf_counter() - t0
    points = sum(j[3] for j in jobs)
    reps = max(1, args.steps // 4)
    # launch shape of the queue, tuned before timing: program fetch (LDS
    # windows or the scalar cache) x tile width (workgroup size) x LDS budget
    # for the register file; the per-tape launches it is compared with run
    # the same shape
    shapes = {}
    for mode in (0, 1):
        ctx.set_fetch(mode)
        for wg in (64, 128, 256):
            for lds in (8192, 16384, 24576, 32768, 49152, 65536):
                shapes[(mode, wg, lds)], _ = timed(lambda: ctx.multi_batch_dev(jobs, wg=wg, lds_budget=lds), 2)
    fetch, wg_q, lds_q = min(shapes, key=shapes.get)
    ctx.set_fetch(fetch)
    ctx.set_launch(wg_q, lds_q)
    dt_q, _ = timed(lambda: ctx.multi_batch_dev(jobs, wg=wg_q, lds_budget=lds_q), reps)

    def per_tape():
        for b, _, x, B, y, g in jobs:
            ctx.gradient_batch_dev(b, x, B, y, g)

    dt_s, _ = timed(per_tape, reps)
    return {"metric": "fwd+grad point-evals/s", "value": points / dt_q, "unit": "points/s", "n_gpus": 1,
            "higher_is_better": True, "vs_baseline": None, "dtype": args.dtype, "data": "synthetic random DAGs",
            "config": {"workload": f"64 random DAGs seed 3 scale {args.dag_scale}, {points} points, "
                                   f"{ops:.3e} entry-evals", "points": points,
                       "fetch": "scalar" if fetch else "lds", "wg": wg_q, "lds_budget": lds_q},
            "queue": {"seconds": dt_q, "entry_evals_per_s": ops / dt_q},
            "per_tape_launches": {"seconds": dt_s, "points_per_s": points / dt_s},
            "speedup": dt_s / dt_q, "setup_s": t_setup}


def main() -> int:
    global CACHE
    args = parse()
    if not args.no_tune_cache:
        from echidna_amd.tunecache import TuneCache

        CACHE = TuneCache(args.tune_cache)
    rc = run(args)
    if CACHE is not None and args.tune_cache_out:
        CACHE.path = args.tune_cache_out
        CACHE.save()
    return rc


def run(args) -> int:
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 "
                                   "(points shard by contiguous ranges across GPUs with no exchange)"}))
        return 2
    if args.workload != "mlp":
        fn = {"hvp": run_hvp, "stde": run_stde, "cv": run_cv, "dags": run_dags}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return 0
    t_start = time.perf_counter()
    head = run_fwdgrad(args, args.dtype, True)
    log(f"headline {head['value'] / 1e6:.1f} M points/s ({time.perf_counter() - t_start:.0f} s)")
    out = {
        "metric": "fwd+grad point-evals/s",
        "value": head["value"],
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic: Philox points x~N(0,I), random-init weights baked into the tape",
        "config": head["config"],
        "graph_ms_per_step": head["graph_ms_per_step"],
        "kernel_ms": head["kernel_ms"],
        "settle_launches": head["settle_launches"],
        "roofline": head["roofline"],
    }
    for key in ("cpu_baseline", "error_vs_oracle", "fanin_f64"):
        if key in head:
            out[key] = head[key]
    if not args.headline_only:
        other = "f64" if args.dtype == "f32" else "f32"
        sub = run_fwdgrad(args, other, False)
        out[other] = {k: sub[k] for k in ("value", "ms_per_step", "kernel_ms", "roofline", "cpu_baseline",
                                          "error_vs_oracle") if k in sub}
        out[other]["launch"] = sub["config"]["launch"]
        log(f"{other} {sub['value'] / 1e6:.1f} M points/s ({time.perf_counter() - t_start:.0f} s)")
        out["hvp"] = run_hvp(args, compact=True)
        log(f"hvp {out['hvp']['value'] / 1e6:.1f} M points/s ({time.perf_counter() - t_start:.0f} s)")
        out["stde"] = run_stde(args, compact=True)
        log(f"stde {out['stde']['rademacher']['jets_per_s'] / 1e6:.1f} M jets/s "
            f"({time.perf_counter() - t_start:.0f} s)")
        out["cv"] = run_cv(args, compact=True)
        log(f"cv speedup {out['cv']['learned_cv']['speedup']:.3f} ({time.perf_counter() - t_start:.0f} s)")
    out["bench_seconds"] = time.perf_counter() - t_start
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
