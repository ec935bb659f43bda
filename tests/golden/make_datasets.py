#!/usr/bin/env python3
"""Write the synthetic observation series into the example contexts.

For each benchmark model: one trajectory of the stochastic model (the oracle's
propagation, seed DATA_SEED, stream 0xDA7A) at weekly observation times, then
y = max(0, round(mu + sd * z)) with the model's observation mean and variance
(the role of the reference's `observation`, C/templates/core_template.c:396-412, and of
its simulator, model_builder/C/simulation), z from numpy's PCG64 seeded with DATA_SEED.
The series are written as context.json data (dates, values), the same format as
model_builder/example/noise/context.json.

Run from the repository root:  python tests/golden/make_datasets.py
"""
import datetime
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "tests"))
from pkgload import load_package  # noqa: E402

pkg = load_package()
sys.path.insert(0, ROOT)
from oracle import oracle  # noqa: E402
from iip_amd import configs  # noqa: E402
from iip_amd.smc import Model  # noqa: E402
from iip_amd.codegen import expr as E  # noqa: E402


def obs_mean_var(model, x, pv, ts):
    lay = model.layout
    env = {"x": x, "t": 0.0}
    for p, off in lay["obs"].items():
        env[p] = pv[off + ts]
    for p, off in lay["fixed_obs"].items():
        env[p] = pv[off + ts]
    funcs = {}
    mu = E.evaluate(E.parse(model.desc.obs_model["mean"]), env, funcs)
    var = E.evaluate(E.parse(model.desc.obs_model["var"]), env, funcs)
    return mu, var


def main():
    rng = np.random.default_rng(configs.DATA_SEED)
    for name in ("sir", "sirs4", "seir10"):
        model = Model.get(name)
        N = configs.DATA_N[name]
        times = np.arange(N + 1, dtype=np.float64) * 7.0
        pv = model.pack(configs.TRUTH[name])
        x0 = model.x0(configs.X0_PROPS[name])
        traj = oracle.trajectory(model.id, times, pv, x0, seed=configs.DATA_SEED, stream=0xDA7A, dt=0.25)
        obs = np.zeros(model.n_ts)
        data = np.zeros((N, model.n_ts))
        for n in range(N):
            obs, _, _ = oracle.obs_loglik(model.id, traj[n + 1], np.full(model.n_ts, np.nan), pv)
            for ts in range(model.n_ts):
                mu, var = obs_mean_var(model, obs[ts], pv, ts)
                data[n, ts] = max(0.0, float(np.rint(mu + np.sqrt(max(var, 0.0)) * rng.standard_normal())))
        path = os.path.join(ROOT, "iterated-inference-pipeline_amd", "examples", name, "context.json")
        ctx = json.load(open(path))
        t0 = datetime.date(2010, 1, 3)
        dates = [(t0 + datetime.timedelta(days=int(t))).isoformat() for t in times[1:]]
        # data columns are in the model's sorted time-series order
        for k, ts in enumerate(model.desc.ts_id):
            ctx["data"]["source"][ts] = {"t0": t0.isoformat(),
                                         "value": [[d, float(v)] for d, v in zip(dates, data[:, k])]}
        for meta in ctx.get("metadata", []):
            for cac, src in meta["source"].items():
                v = src["value"][0][1]
                src["value"] = [[t0.isoformat(), v], [dates[len(dates) // 2], v], [dates[-1], v]]
        ctx["description"] = ("synthetic context: series simulated from the model at the benchmark truth "
                              "(tests/golden/make_datasets.py, seed %d)" % configs.DATA_SEED)
        with open(path, "w") as f:
            json.dump(ctx, f, indent=1)
        print(name, "N =", N, "total incidence", data.sum(axis=0)[:4], "max", data.max())


if __name__ == "__main__":
    main()
