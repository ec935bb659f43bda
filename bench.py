#!/usr/bin/env python3
"""Benchmark of the MI355X PLOM SMC filter on one GPU.

Workload (one "step" = one full filter over the series): SIRS with seasonal forcing
(amplitude 0.3, period 365 d) and multiplicative white noise on beta (sigma 0.1),
4 coupled populations of 5e5, N = 520 weekly intervals, J = 100000 particles,
dt = 0.25 d (28 Euler substeps per interval), systematic resampling every interval.
Inputs are resident in HBM before the timed region; the filter runs in FP64 (the
oracle-exact mode).  `value` is particle-steps/s (J * N per step).

Extra legs on the same line: the FP32-propagation run of the same workload, the
roofline of the dominant kernel (k_propagate) from HIP events around every launch on
the library's stream, and a CPU baseline: the C oracle (this project's restatement of
the reference filter) timed on the host cores on a bounded sample.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "tests"))
sys.path.insert(0, HERE)

from pkgload import load_package  # noqa: E402

HBM_PEAK_GBPS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def bytes_per_particle_step(S, real_bytes):
    """Algorithmic HBM bytes per particle-step: state read + written (2 S b), log-weight
    written and read (16), ancestor index (4)."""
    return 2 * S * real_bytes + 20


def read_traffic(name):
    """HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by
    the copy-kernel calibration, as committed in profiles/<name> ('traffic_bytes_per_launch: N')."""
    path = os.path.join(HERE, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def run_gpu(smc, configs, precision, steps, warmup, J, N):
    model = smc.Model.get("sirs4")
    times, data = model.dataset(N)
    pv = model.pack(configs.TRUTH["sirs4"])
    x0 = model.x0(configs.X0_PROPS["sirs4"])
    ctx = smc.Context(0)
    f = smc.Filter(ctx, model, J, times, data, pv, x0, precision=precision, seed=configs.FILTER_SEED, stream=0)
    for _ in range(warmup):
        f.run()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        f.run()
    ctx.sync()
    t1 = time.perf_counter()
    res = f.read()
    # per-launch timing pass (events around every launch), not part of the timed region
    ctx.set_kernel_timing(True)
    f.run()
    total_ms, prop_ms, scan_ms, nint = ctx.last_timing()
    ctx.set_kernel_timing(False)
    f.close()
    ctx.close()
    wall = (t1 - t0) / steps
    return {
        "ms_per_step": wall * 1e3,
        "particle_steps_per_s": J * N / wall,
        "loglik": res.loglik,
        "n_all_fail": res.n_all_fail,
        "propagate_ms_per_launch": prop_ms / max(nint, 1),
        "scan_ms_per_launch": scan_ms / max(nint, 1),
        "launches": nint,
        "S": model.S,
    }


def cpu_baseline(configs, smc, J, N, threads):
    from oracle import oracle
    model = smc.Model.get("sirs4")
    times, data = model.dataset(N)
    r = oracle.smc(model.id, J, times, data, model.pack(configs.TRUTH["sirs4"]),
                   model.x0(configs.X0_PROPS["sirs4"]), seed=configs.FILTER_SEED, nthreads=threads)
    return J * N / r["elapsed_s"], r["elapsed_s"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--J", type=int, default=100000)
    ap.add_argument("--N", type=int, default=520)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-J", type=int, default=20000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB device copy (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each filter or request "
              "runs on one device; see DESIGN.md for how it would shard). Run with --gpus 1.", file=sys.stderr)
        return 2
    load_package()
    from iip_amd import smc, configs

    if args.copy_calibration:
        with smc.Context(0) as ctx:
            gbps, ms = smc.copy_bandwidth_gbps(ctx, 1 << 30, 1)
        print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30, "GBps": gbps, "ms": ms}))
        return 0

    with smc.Context(0) as ctx:
        copy_gbps, _ = smc.copy_bandwidth_gbps(ctx, 1 << 30, 5)

    J, N = args.J, args.N
    fp64 = run_gpu(smc, configs, "fp64", args.steps, args.warmup, J, N)
    fp32 = run_gpu(smc, configs, "fp32", args.steps, args.warmup, J, N)

    B = bytes_per_particle_step(fp64["S"], 8)
    per_launch = B * J
    achieved = per_launch / (fp64["propagate_ms_per_launch"] * 1e-3) / 1e9
    traffic = read_traffic("k_propagate_sirs4_fp64_J100000.md")
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 5), "traffic": traffic,
                "kernel": "k_propagate<sirs4,double,full>",
                "algorithmic_bytes_per_launch": per_launch,
                "avg_launch_ms": round(fp64["propagate_ms_per_launch"], 5),
                "measured_copy_GBps": round(copy_gbps, 1)}

    cpu = None
    if not args.no_cpu:
        threads = max(1, min(16, os.cpu_count() or 1))
        rate, secs = cpu_baseline(configs, smc, args.cpu_J, N, threads)
        cpu = {"value": round(rate, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": "oracle/smc_oracle.c (canonical FP64) on the same workload with J=%d instead of %d "
                         "(particle-steps/s is linear in J), N=%d, %.1f s on %d threads" % (args.cpu_J, J, N, secs, threads)}

    out = {
        "metric": "particle-steps/s & p50/p99 loglik latency, batched+served vs plain SMC",
        "value": round(fp64["particle_steps_per_s"], 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fp64["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (series simulated from the model at the benchmark truth, seed 20261203)",
        "config": {"workload": "SIRS 4x5e5 coupled, seasonal forcing 0.3/365d, env noise 0.1; weekly incidence; "
                               "J=%d, N=%d, dt=0.25 d, FP64" % (J, N), "J": J, "N": N, "dt": 0.25,
                   "particle_steps_per_step": J * N, "parallelism": "1 GPU"},
        "loglik": fp64["loglik"],
        "fp32": {"value": round(fp32["particle_steps_per_s"], 1), "ms_per_step": round(fp32["ms_per_step"], 3),
                 "loglik": fp32["loglik"], "avg_launch_ms": round(fp32["propagate_ms_per_launch"], 5)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
