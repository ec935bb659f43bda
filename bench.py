#!/usr/bin/env python3
"""Benchmark of the MI355X PLOM SMC filter on one GPU: batched and served vs plain.

Headline workload (one "step" = one batched pass over all filters): the pMCMC batch, 256
likelihood evaluations of the SIRS model (seasonal forcing amplitude 0.3 / period 365 d,
white noise on beta sigma 0.1, 4 coupled populations of 5e5, weekly incidence), with
parameter sets drawn around the truth, J ~ U{1000..20000} and N ~ U{104..520} per filter,
dt = 0.25 d (28 Euler substeps per interval), systematic resampling every interval.
All 256 filters advance together (one batched launch round per interval), replayed from
a captured hipGraph.  It runs in FP64, where every log-likelihood is bit-identical to the
filter run alone (and to the CPU oracle), so the error against FP64 is exactly 0.
`value` is particle-steps/s (sum of J * N over the filters / time per pass), inputs
resident in HBM.

Legs on the same JSON line:
  plain     the same 256 filters run one at a time (one filter per launch sequence,
            result read before the next one starts), FP64: the baseline of the speedup
  batched   the headline numbers, the FP32-propagation pass of the same batch and its
            deviation from FP64
  served    a Poisson stream (200 requests/s for 60 s) of SIR / SIRS requests, J log-uniform
            in [500, 5e4], N ~ U{52..520}, served with continuous batching and admission by
            free HBM: sustained particle-steps/s and p50 / p99 latency (submit -> result);
            then the same mix offered at 400, 600 and 800 requests/s (10 s each): the capacity is
            the highest sustained rate
  single    one large filter (SIRS, J = 1e5, N = 520), FP64 and FP32, and FP64 with the
            outputs the smc command writes by default (hat and prediction residuals)
  psr       the same filter with the psr implementation (Euler-multinomial step, FP64),
            and the C oracle's psr on a bounded sample of it
  mif       one iterated-filtering pass of the same filter with per-particle parameters
            (28 per-particle slots, random walk after every observation, theta moments)
  roofline  of the dominant kernel (k_propagate of the batch), from HIP events around
            every propagate launch on the stream it runs on (the device time the lanes'
            launches occupy)
  cpu_baseline  the C oracle (this project's restatement of the reference filter) on a
            bounded sample of the headline workload, timed on the host cores

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs a,b,...] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from pkgload import load_package  # noqa: E402

HBM_PEAK_GBPS = 8000.0
FP64_PEAK_TFLOPS = 78.6
METRIC = "particle-steps/s & p50/p99 loglik latency, batched+served vs plain SMC"
LEGS = ("batched", "plain", "served", "single", "psr", "mif", "memory", "cpu")
# further legs named only on request: memory_alone (the memory-heavy filter without the
# served stream), fp64_only (the batched leg without its FP32 pass)


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def bytes_per_particle_step(S, real_bytes):
    """Algorithmic HBM bytes per particle-step: state read + written (2 S b), log-weight
    written and read (16), ancestor index (4)."""
    return 2 * S * real_bytes + 20


def read_profile_value(name, key):
    """A figure measured by rocprofv3 and committed in profiles/<name> as a line 'key: value'
    (traffic_bytes_per_launch: FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel
    calibration; valu_issue_frac: SQ_INSTS_VALU x 4 cycles over the SIMD cycles of the
    dispatch; fp64_flops_per_particle_step: from the SQ_INSTS_VALU_*_F64 counters)."""
    path = os.path.join(HERE, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def cpu_model():
    """The host CPU's model name (what lscpu reports), from /proc/cpuinfo."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class Inputs:
    """Datasets, parameter vectors and initial states, built once per (model, N)."""

    def __init__(self, smc, configs):
        self.smc, self.configs = smc, configs
        self.models, self.data, self.x0 = {}, {}, {}

    def model(self, name):
        if name not in self.models:
            self.models[name] = self.smc.Model.get(name)
            self.x0[name] = self.models[name].x0(self.configs.X0_PROPS[name])
        return self.models[name]

    def get(self, name, N, params):
        m = self.model(name)
        if (name, N) not in self.data:
            self.data[(name, N)] = m.dataset(N)
        times, data = self.data[(name, N)]
        return m, times, data, m.pack(params), self.x0[name]


def make_filters(smc, ctx, inp, draws, precision):
    fs = []
    for d in draws:
        m, times, data, pv, x0 = inp.get(d["model"], d["N"], d["params"])
        fs.append(smc.Filter(ctx, m, d["J"], times, data, pv, x0, precision=precision,
                             seed=inp.configs.FILTER_SEED, stream=d["stream"]))
    return fs


def batched_leg(smc, ctx, filters, steps, warmup, lanes=0):
    """All filters advance together; the launch sequence is captured once as a hipGraph
    (first launch) and replayed.  lanes = 0: the autotuner picks the number of lanes
    (concurrent chains) from the first intervals."""
    b = smc.Batch(ctx, filters)
    tuned = None
    if lanes:
        b.configure(lanes)
    else:
        tuned = b.tune(rounds=32)
        lanes = tuned["lanes"]
    b.launch(use_graph=True)              # capture + first replay
    for _ in range(warmup):
        b.launch(use_graph=True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.launch(use_graph=True)
    ctx.sync()
    wall = (time.perf_counter() - t0) / steps
    ll = np.array([f.read().loglik for f in filters])
    # per-launch timing pass, outside the timed region, in the timed configuration (the
    # same lanes): every lane brackets its launches with events on its own stream, so a
    # launch's duration is the one it has beside the other lanes' launches
    ctx.set_kernel_timing(True)
    b.launch(use_graph=False)
    total_ms, prop_ms, scan_ms, nint, nprop = ctx.last_timing()
    lane_prop, lane_scan = ctx.last_timing_lanes()
    prop_busy, scan_busy = ctx.last_timing_busy()
    ctx.set_kernel_timing(False)
    b.close()
    return {"wall_s": wall, "loglik": ll, "prop_ms": prop_ms, "scan_ms": scan_ms, "total_ms": total_ms,
            "launches": nprop, "rounds": nint, "lanes": lanes, "tuning": tuned,
            "lane_prop_ms": lane_prop, "lane_scan_ms": lane_scan, "prop_busy_ms": prop_busy,
            "scan_busy_ms": scan_busy}


def plain_leg(filters):
    """One filter at a time: its launch sequence, then its result, then the next filter."""
    t0 = time.perf_counter()
    ll = []
    for f in filters:
        f.run()
        ll.append(f.read().loglik)
    return {"wall_s": time.perf_counter() - t0, "loglik": np.array(ll)}


def served_leg(smc, ctx, inp, configs, seconds, rate, check_n=16, heavy=None, window_s=None):
    """heavy: an extra request (dict with model, J, N, stream) submitted with the first
    arrival, with lineage, served alongside the stream.  window_s: also the p99 latency of
    the requests that arrived in the first window_s seconds (the stream of that length is
    a prefix of this one)."""
    reqs = configs.served_mix(rate=rate, duration_s=seconds)
    prepared = [inp.get(r["model"], r["N"], r["params"]) for r in reqs]
    srv = smc.Server(ctx)
    # warm the server's pools and every kernel it launches (not measured)
    warm = []
    for k in range(min(4, len(reqs))):
        m, times, data, pv, x0 = prepared[k]
        warm.append(srv.submit(m, reqs[k]["J"], times, data, pv, x0, seed=configs.FILTER_SEED, stream=reqs[k]["stream"]))
    while len(srv.done()) < len(warm):
        srv.step()
    for rid in warm:
        srv.result(rid)
    sub_t, lat, res = {}, {}, {}
    i, peak_active = 0, 0
    heavy_id, heavy_out = None, {}
    t0 = time.perf_counter()
    if heavy is not None:
        m, times, data, pv, x0 = inp.get(heavy["model"], heavy["N"], configs.TRUTH[heavy["model"]])
        heavy_id = srv.submit(m, heavy["J"], times, data, pv, x0, seed=configs.FILTER_SEED, stream=heavy["stream"],
                              lineage=True)
    while True:
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i]["t"] <= now:
            m, times, data, pv, x0 = prepared[i]
            rid = srv.submit(m, reqs[i]["J"], times, data, pv, x0, seed=configs.FILTER_SEED, stream=reqs[i]["stream"])
            sub_t[rid] = (time.perf_counter() - t0, i)
            i += 1
        a, q, d = srv.step()
        peak_active = max(peak_active, a)
        if d:
            for rid in srv.done():
                t_done = time.perf_counter() - t0
                if rid == heavy_id:
                    r = srv.result(rid)
                    heavy_out = {"latency_s": t_done, "loglik": r.loglik, "lineage_rows": int(r.lineage.shape[0])}
                    continue
                ts, k = sub_t.pop(rid)
                lat[k] = t_done - ts
                res[k] = srv.result(rid).loglik
        if i == len(reqs) and not sub_t and (heavy_id is None or heavy_out):
            break
        if a == 0 and q == 0 and d == 0 and not sub_t and i < len(reqs):
            time.sleep(max(0.0, min(0.001, reqs[i]["t"] - (time.perf_counter() - t0))))
    makespan = time.perf_counter() - t0
    ps = sum(r["J"] * r["N"] for r in reqs)
    lat_ms = np.array([lat[k] for k in range(len(reqs))]) * 1e3
    # a sample of requests run alone: identical results (FP64)
    rng = np.random.default_rng(7)
    sample = rng.choice(len(reqs), size=min(check_n, len(reqs)), replace=False)
    max_rel = 0.0
    for k in sample:
        m, times, data, pv, x0 = prepared[k]
        alone = smc.run(ctx, m, reqs[k]["J"], times, data, pv, x0, seed=configs.FILTER_SEED, stream=reqs[k]["stream"]).loglik
        max_rel = max(max_rel, abs(res[k] - alone) / abs(alone))
    srv.close()
    out = {"requests": len(reqs), "offered_particle_steps_per_s": ps / seconds, "makespan_s": makespan,
            "particle_steps": ps, "particle_steps_per_s": ps / makespan,
            "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
            "mean_ms": float(lat_ms.mean()), "max_ms": float(lat_ms.max()), "peak_active": peak_active,
            "checked_alone": len(sample), "max_rel_err_vs_alone": max_rel}
    if heavy is not None:
        out["heavy"] = heavy_out
    if window_s is not None:
        w = np.array([lat_ms[k] for k in range(len(reqs)) if reqs[k]["t"] < window_s])
        out["window"] = {"seconds": window_s, "requests": int(w.size),
                         "p99_ms": float(np.percentile(w, 99)) if w.size else None}
    return out


def memory_leg(smc, ctx, inp, configs, J, N):
    """The memory-heavy filter: SEIR, 10 gravity-coupled populations of 1e6 (S = 40), one
    run in FP64 keeping the ancestry [N][J] on the device and returning one lineage."""
    m, times, data, pv, x0 = inp.get("seir10", N, configs.TRUTH["seir10"])
    f = smc.Filter(ctx, m, J, times, data, pv, x0, flags=smc.FLAG_LINEAGE, seed=configs.FILTER_SEED, stream=999)
    dev = f.device_bytes()
    f.run()                              # warm
    ctx.sync()
    t0 = time.perf_counter()
    f.run()
    ctx.sync()
    wall = time.perf_counter() - t0
    r = f.read(lineage=True)
    f.close()
    return {"workload": "seir10_j%d_n%d: SEIR 10x1e6 gravity-coupled, env noise, ancestry kept, lineage, FP64" % (J, N),
            "value": round(J * N / wall, 1), "s_per_filter": round(wall, 3), "device_bytes": dev,
            "loglik": r.loglik, "lineage_final_slot": int(r.lineage_slots[-1])}


def pick_tiles_per_block(ctx, f):
    """The tiles per propagate block (1, 2, 4) under which one run of the filter is fastest
    (a warm run first); returns (chosen, {tpb: ms})."""
    ms = {}
    for tpb in (1, 2, 4):
        ctx.set_tiles_per_block(tpb)
        f.run()
        ctx.sync()
        t0 = time.perf_counter()
        f.run()
        ctx.sync()
        ms[tpb] = round((time.perf_counter() - t0) * 1e3, 3)
    best = min(ms, key=ms.get)
    ctx.set_tiles_per_block(best)
    return best, ms


def single_leg(smc, ctx, inp, configs, precision, steps, warmup, J, N):
    m, times, data, pv, x0 = inp.get("sirs4", N, configs.TRUTH["sirs4"])
    f = smc.Filter(ctx, m, J, times, data, pv, x0, precision=precision, seed=configs.FILTER_SEED, stream=0)
    tpb, tpb_ms = pick_tiles_per_block(ctx, f)
    for _ in range(warmup):
        f.run()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        f.run()
    ctx.sync()
    wall = (time.perf_counter() - t0) / steps
    res = f.read()
    ctx.set_kernel_timing(True)
    f.run()
    _, prop_ms, _, nint, _ = ctx.last_timing()
    ctx.set_kernel_timing(False)
    ctx.set_tiles_per_block(1)
    f.close()
    return {"value": round(J * N / wall, 1), "ms_per_filter": round(wall * 1e3, 3), "loglik": res.loglik,
            "propagate_ms_per_launch": round(prop_ms / max(nint, 1), 5), "tiles_per_block": tpb,
            "ms_per_filter_by_tiles_per_block": tpb_ms}


def hat_leg(smc, ctx, inp, configs, steps, warmup, J, N):
    """The same filter with the outputs the smc command writes by default: the hat (weighted
    mean and 95 % interval of every state and observation per interval) and the
    prediction residuals."""
    m, times, data, pv, x0 = inp.get("sirs4", N, configs.TRUTH["sirs4"])
    f = smc.Filter(ctx, m, J, times, data, pv, x0, flags=smc.FLAG_HAT, seed=configs.FILTER_SEED, stream=0)
    for _ in range(max(warmup, 1)):
        f.run()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        f.run()
    ctx.sync()
    wall = (time.perf_counter() - t0) / steps
    f.close()
    return {"value": round(J * N / wall, 1), "ms_per_filter": round(wall * 1e3, 3)}


def psr_leg(smc, ctx, inp, configs, steps, warmup, J, N, cpu, cpu_J, cpu_N):
    m, times, data, pv, x0 = inp.get("sirs4", N, configs.TRUTH["sirs4"])
    x0 = np.round(x0)
    f = smc.Filter(ctx, m, J, times, data, pv, x0, flags=smc.FLAG_PSR, seed=configs.FILTER_SEED, stream=0)
    for _ in range(warmup):
        f.run()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        f.run()
    ctx.sync()
    wall = (time.perf_counter() - t0) / steps
    res = f.read()
    f.close()
    out = {"value": round(J * N / wall, 1), "ms_per_filter": round(wall * 1e3, 3), "loglik": res.loglik}
    if cpu:
        from oracle import oracle
        mc, tc, dc, pvc, x0c = inp.get("sirs4", cpu_N, configs.TRUTH["sirs4"])
        r = oracle.smc(mc.id, cpu_J, tc, dc, pvc, np.round(x0c), seed=configs.FILTER_SEED, stream=0, psr=True,
                       nthreads=min(16, os.cpu_count() or 1))
        out["cpu_oracle"] = {"value": round(cpu_J * cpu_N / r["elapsed_s"], 1), "cores": min(16, os.cpu_count() or 1),
                             "sample": "SIRS psr, J=%d, N=%d" % (cpu_J, cpu_N)}
    return out


def mif_leg(smc, ctx, inp, configs, steps, warmup, J, N):
    """theta = log of every per-particle pv slot (logit of the reporting rates), jump sizes
    0.02, the walk factor of a weekly interval at the first iteration"""
    m, times, data, pv, x0 = inp.get("sirs4", N, configs.TRUTH["sirs4"])
    lay = m.layout
    rep = set(range(lay["obs"]["rep"], lay["obs"]["rep"] + m.n_ts))
    tr = [2 if k in rep else 1 for k in range(m.n_pp)]
    th = np.array([np.log(pv[k] / (1 - pv[k])) if k in rep else np.log(pv[k]) for k in range(m.n_pp)])
    nth = m.n_pp
    mp = {"pv_src": list(range(nth)), "tr": tr, "scale": [0] * nth, "duration": [0] * nth, "min": [0.0] * nth,
          "max": [0.0] * nth, "multiplier": [1.0] * nth, "sd": [0.02] * nth}
    rng = np.random.default_rng(7)
    theta = th[:, None] + rng.normal(0.0, 0.02, (nth, J))
    xs = np.repeat(x0[:, None], J, axis=1)
    mut = np.full(N, np.sqrt(7.0))
    f = smc.Filter(ctx, m, J, times, data, pv, x0, flags=smc.FLAG_MIF, mif=mp, seed=configs.FILTER_SEED)
    for _ in range(warmup):
        f.mif_start(theta, xs, mut, seed=1, stream=0)
        f.run()
    ctx.sync()
    wall = 0.0
    for it in range(steps):
        f.mif_start(theta, xs, mut, seed=1, stream=it)   # the upload of the starting cloud is not timed
        t0 = time.perf_counter()
        f.run()
        ctx.sync()
        wall += time.perf_counter() - t0
    wall /= steps
    res = f.read(mif=True)
    f.close()
    return {"value": round(J * N / wall, 1), "ms_per_pass": round(wall * 1e3, 3), "loglik": res.loglik,
            "n_theta": nth}


def cpu_baseline(smc, inp, configs, draws, threads, reference=False):
    from oracle import oracle
    secs, ps = 0.0, 0
    mode = oracle.REFERENCE if reference else oracle.CANONICAL
    for d in draws:
        m, times, data, pv, x0 = inp.get(d["model"], d["N"], d["params"])
        r = oracle.smc(m.id, d["J"], times, data, pv, x0, seed=configs.FILTER_SEED, stream=d["stream"], nthreads=threads,
                       mode=mode)
        secs += r["elapsed_s"]
        ps += d["J"] * d["N"]
    return ps / secs, secs, ps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--legs", default=",".join(LEGS), help="comma list of " + ",".join(LEGS))
    ap.add_argument("--filters", type=int, default=256)
    ap.add_argument("--lanes", type=int, default=0, help="concurrent chains of the batch (0: autotuned)")
    ap.add_argument("--served-seconds", type=float, default=60.0)
    ap.add_argument("--alongside-seconds", type=float, default=30.0,
                    help="length of the stream served alongside the memory-heavy filter")
    ap.add_argument("--served-rate", type=float, default=200.0)
    ap.add_argument("--sweep-rates", type=lambda x: [float(v) for v in x.split(",") if v], default=[400.0, 600.0, 800.0],
                    help="further arrival rates of the served capacity sweep (comma list; empty: none)")
    ap.add_argument("--sweep-seconds", type=float, default=10.0)
    ap.add_argument("--single-J", type=int, default=100000)
    ap.add_argument("--single-N", type=int, default=520)
    ap.add_argument("--cpu-filters", type=int, default=8)
    ap.add_argument("--memory-J", type=int, default=1000000)
    ap.add_argument("--memory-N", type=int, default=1040)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB device copy (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each filter or request runs "
              "on one device). Run with --gpus 1.", file=sys.stderr)
        return 2
    legs = set(x for x in args.legs.split(",") if x)
    if args.no_cpu:
        legs.discard("cpu")
    load_package()
    from iip_amd import smc, configs

    if args.copy_calibration:
        with smc.Context(0) as ctx:
            gbps, ms = smc.copy_bandwidth_gbps(ctx, 1 << 30, 1)
        print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30, "GBps": gbps, "ms": ms}))
        return 0

    ctx = smc.Context(0)
    inp = Inputs(smc, configs)
    copy_gbps, _ = smc.copy_bandwidth_gbps(ctx, 1 << 30, 5)
    draws = configs.pmcmc_batch(args.filters)
    ps_batch = sum(d["J"] * d["N"] for d in draws)
    S = inp.model("sirs4").S
    out = {"metric": METRIC, "value": None, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic: series simulated from each model at its truth (seed 20261203); parameter sets, "
                   "sizes and arrivals drawn from fixed seeds",
           "config": {"workload": "pmcmc_batch: %d SIRS 4x5e5 filters (forcing 0.3/365 d, env noise 0.1), "
                                  "J~U{1000..20000}, N~U{104..520}, dt 0.25 d, FP64, one batched pass per "
                                  "step replayed from a hipGraph" % args.filters,
                      "filters": args.filters, "particle_steps_per_step": ps_batch, "parallelism": "1 GPU"}}

    fp64 = None
    if "batched" in legs or "plain" in legs:
        log("creating %d FP64 filters" % args.filters)
        f64 = make_filters(smc, ctx, inp, draws, "fp64")
    if "batched" in legs:
        log("batched FP64")
        fp64 = batched_leg(smc, ctx, f64, args.steps, args.warmup, args.lanes)
        out["value"] = round(ps_batch / fp64["wall_s"], 1)
        out["ms_per_step"] = round(fp64["wall_s"] * 1e3, 3)
        out["batched"] = {
            "fp64": {"value": out["value"], "ms_per_pass": out["ms_per_step"], "graph": True,
                     "lanes": fp64["lanes"], "tuning": fp64["tuning"],
                     "propagate_busy_ms": round(fp64["prop_busy_ms"], 3),
                     "weights_scan_busy_ms": round(fp64["scan_busy_ms"], 3),
                     "launch_rounds": fp64["rounds"], "propagate_launches": fp64["launches"]}}
        B = bytes_per_particle_step(S, 8)
        nl = max(fp64["launches"], 1)
        per_launch = B * ps_batch / nl
        # the device time per launch: the union of the lanes' propagate spans over the
        # launches (launches that run beside each other count once), so that the kernel
        # time of a pass never exceeds the pass; each launch's own span, beside the other
        # lanes' launches, is reported too
        avg_ms = fp64["prop_busy_ms"] / nl
        span_ms = fp64["prop_ms"] / nl
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        prof = "k_propagate_sirs4_fp64_pmcmc_batch.md"
        valu = read_profile_value(prof, "valu_issue_frac")
        fpps = read_profile_value(prof, "fp64_flops_per_particle_step")
        rl = {"bound": "valu", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
              "frac": round(achieved / HBM_PEAK_GBPS, 5),
              "traffic": read_profile_value(prof, "traffic_bytes_per_launch"),
              "kernel": "k_propagate<sirs4,double,full> (batched, %d lanes)" % fp64["lanes"],
              "binding": "VALU issue, in it the FP64 operations: the kernel's FP64 arithmetic per byte is about "
                         "9x the chip's FP64 ridge, so the frac above is its HBM share and valu_issue_frac "
                         "(and fp64.frac) its compute use",
              "valu_issue_frac": valu,
              "algorithmic_bytes_per_particle_step": B,
              "algorithmic_bytes_per_launch": round(per_launch, 1),
              "avg_launch_ms": round(avg_ms, 5), "launches": fp64["launches"],
              "avg_launch_span_ms": round(span_ms, 5),
              "propagate_busy_ms_per_pass": round(fp64["prop_busy_ms"], 3),
              "timing_pass_ms": round(fp64["total_ms"], 3),
              "propagate_span_ms_per_pass_by_lane": [round(x, 3) for x in fp64["lane_prop_ms"]],
              "pass_wall_GBps": round(B * ps_batch / fp64["wall_s"] / 1e9, 2),
              "measured_copy_GBps": round(copy_gbps, 1)}
        if fpps:
            tf = fpps * ps_batch / nl / (avg_ms * 1e-3) / 1e12
            rl["fp64"] = {"flops_per_particle_step": fpps, "achieved_TFLOPs": round(tf, 3),
                          "peak_TFLOPs": FP64_PEAK_TFLOPS, "frac": round(tf / FP64_PEAK_TFLOPS, 4)}
        out["roofline"] = rl
    if "batched" in legs and "fp64_only" not in legs:
        log("batched FP32")
        f32 = make_filters(smc, ctx, inp, draws, "fp32")
        fp32 = batched_leg(smc, ctx, f32, args.steps, args.warmup, args.lanes)
        for f in f32:
            f.close()
        rel32 = np.abs(fp32["loglik"] - fp64["loglik"]) / np.abs(fp64["loglik"])
        out["batched"]["fp32"] = {
            "value": round(ps_batch / fp32["wall_s"], 1), "ms_per_pass": round(fp32["wall_s"] * 1e3, 3),
            "lanes": fp32["lanes"], "tuning": fp32["tuning"],
            "mean_rel_dev_vs_fp64": float(rel32.mean()), "max_rel_dev_vs_fp64": float(rel32.max()),
            "note": "FP32 propagation with FP64 weights: the FP64 filter's Gaussians rounded to FP32 (the same "
                    "words and inverse-CDF segments); the deviation is FP32 rounding, amplified where it flips "
                    "a resampling choice"}
    if "plain" in legs:
        log("plain FP64 (one filter at a time)")
        pl = plain_leg(f64)
        out["plain"] = {"value": round(ps_batch / pl["wall_s"], 1), "s_per_pass": round(pl["wall_s"], 3),
                        "what": "the same %d FP64 filters, one filter per launch sequence, result read before "
                                "the next starts" % args.filters}
        if fp64 is not None:
            rel = np.abs(fp64["loglik"] - pl["loglik"]) / np.abs(pl["loglik"])
            out["batched"]["fp64"]["speedup_vs_plain"] = round(out["value"] / out["plain"]["value"], 2)
            out["batched"]["fp64"]["max_rel_err_vs_plain"] = float(rel.max())
    if "batched" in legs or "plain" in legs:
        for f in f64:
            f.close()
    if "served" in legs:
        log("served stream, %.0f s at %.0f requests/s" % (args.served_seconds, args.served_rate))
        sv = served_leg(smc, ctx, inp, configs, args.served_seconds, args.served_rate,
                        window_s=args.alongside_seconds if "memory" in legs else None)
        sv = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in sv.items()}
        sv["workload"] = ("served_mix: Poisson %.0f req/s for %.0f s, 50/50 SIR / SIRS4, J log-uniform [500, 5e4], "
                          "N~U{52..520}, FP64" % (args.served_rate, args.served_seconds))
        if "plain" in out:
            sv["vs_plain"] = round(sv["particle_steps_per_s"] / out["plain"]["value"], 2)
        # capacity: the same mix offered faster, until the server no longer keeps up
        points = [{"rate": args.served_rate, "seconds": args.served_seconds,
                   **{k: sv[k] for k in ("requests", "offered_particle_steps_per_s", "particle_steps_per_s",
                                         "makespan_s", "p50_ms", "p99_ms", "peak_active")}}]
        for rate in (r for r in args.sweep_rates if r != args.served_rate):
            log("served capacity sweep: %.0f s at %.0f requests/s" % (args.sweep_seconds, rate))
            p = served_leg(smc, ctx, inp, configs, args.sweep_seconds, rate, check_n=4)
            points.append({"rate": rate, "seconds": args.sweep_seconds,
                           **{k: round(p[k], 3) if isinstance(p[k], float) else p[k]
                              for k in ("requests", "offered_particle_steps_per_s", "particle_steps_per_s",
                                        "makespan_s", "p50_ms", "p99_ms", "peak_active", "max_rel_err_vs_alone")}})
        best = max(points, key=lambda q: q["particle_steps_per_s"])
        sv["sweep"] = points
        sv["capacity"] = {"rate": best["rate"], "particle_steps_per_s": round(best["particle_steps_per_s"], 1),
                          "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"],
                          "keeps_up": best["particle_steps_per_s"] >= 0.95 * best["offered_particle_steps_per_s"]}
        if "plain" in out:
            sv["capacity"]["vs_plain"] = round(best["particle_steps_per_s"] / out["plain"]["value"], 2)
        out["served"] = sv
    if "memory" in legs or "memory_alone" in legs:
        log("memory-heavy filter alone")
        mem = memory_leg(smc, ctx, inp, configs, args.memory_J, args.memory_N)
        out["memory_heavy"] = mem
    if "memory" in legs:
        log("memory-heavy filter served alongside the stream")
        sv = served_leg(smc, ctx, inp, configs, args.alongside_seconds, args.served_rate, check_n=4,
                        heavy={"model": "seir10", "J": args.memory_J, "N": args.memory_N, "stream": 999})
        mem["served_alongside"] = {"heavy_latency_s": round(sv["heavy"]["latency_s"], 3),
                                   "heavy_loglik_equals_alone": sv["heavy"]["loglik"] == mem["loglik"],
                                   "p50_ms": round(sv["p50_ms"], 3), "p99_ms": round(sv["p99_ms"], 3),
                                   "particle_steps_per_s": round((sv["particle_steps"] + args.memory_J * args.memory_N)
                                                                 / sv["makespan_s"], 1),
                                   "max_rel_err_vs_alone": sv["max_rel_err_vs_alone"],
                                   "heavy_latency_vs_alone": round(sv["heavy"]["latency_s"] / mem["s_per_filter"], 2)}
        base = out.get("served", {}).get("window")
        if base and base.get("p99_ms"):
            # the same requests (the served stream's first alongside_seconds) without the
            # heavy request
            mem["served_alongside"]["stream_seconds"] = args.alongside_seconds
            mem["served_alongside"]["small_p99_without_heavy_ms"] = round(base["p99_ms"], 3)
            mem["served_alongside"]["small_p99_vs_without_heavy"] = round(sv["p99_ms"] / base["p99_ms"], 2)
        out["memory_heavy"] = mem
    if "single" in legs:
        log("single large filter")
        out["single"] = {"workload": "SIRS 4x5e5, J=%d, N=%d" % (args.single_J, args.single_N),
                         "fp64": single_leg(smc, ctx, inp, configs, "fp64", args.steps, args.warmup, args.single_J, args.single_N),
                         "fp32": single_leg(smc, ctx, inp, configs, "fp32", args.steps, args.warmup, args.single_J, args.single_N),
                         "fp64_hat": hat_leg(smc, ctx, inp, configs, args.steps, args.warmup, args.single_J, args.single_N)}
    if "psr" in legs:
        log("psr filter")
        out["psr"] = dict(workload="SIRS 4x5e5 psr (Euler-multinomial), J=%d, N=%d, FP64" % (args.single_J, args.single_N),
                          **psr_leg(smc, ctx, inp, configs, args.steps, args.warmup, args.single_J, args.single_N,
                                    "cpu" in legs, 2000, 104))
    if "mif" in legs:
        log("iterated-filtering pass")
        out["mif"] = dict(workload="SIRS 4x5e5, J=%d, N=%d, FP64, per-particle parameters" % (args.single_J, args.single_N),
                          **mif_leg(smc, ctx, inp, configs, args.steps, args.warmup, args.single_J, args.single_N))
    ctx.close()
    if "cpu" in legs:
        host = os.cpu_count() or 1
        threads = min(16, host)        # the CPU share of one GPU on the measuring host
        sample = draws[:args.cpu_filters]
        log("CPU oracle on %d filters, %d threads" % (len(sample), threads))
        rate, secs, ps = cpu_baseline(smc, inp, configs, sample, threads)
        ref_sample = sample[:2]
        log("CPU oracle, the reference's sequential arithmetic, on %d filters" % len(ref_sample))
        rate_ref, secs_ref, ps_ref = cpu_baseline(smc, inp, configs, ref_sample, threads, reference=True)
        out["cpu_baseline"] = {"value": round(rate, 1), "unit": "particle-steps/s", "cores": threads, "kind": "port",
                               "host_cpus": host, "cpu_model": cpu_model(),
                               "sample": "oracle/smc_oracle.c (canonical FP64) on the first %d filters of the "
                                         "pmcmc_batch workload (%d particle-steps), %.1f s on %d threads"
                                         % (len(sample), ps, secs, threads),
                               "reference_arithmetic": {
                                   "value": round(rate_ref, 1), "cores": threads,
                                   "sample": "the same oracle in the reference's sequential weight / resampling "
                                             "arithmetic (smc.c:27-112 as written) on the first %d filters "
                                             "(%d particle-steps), %.1f s" % (len(ref_sample), ps_ref, secs_ref)}}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
