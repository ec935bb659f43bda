"""The `pmcmc` command (particle marginal Metropolis-Hastings, pmcmc/pmcmc.c:40-266) on the
GPU against the CPU restatement of the chain (oracle/pmcmc_oracle.py, pinned in
test_pmcmc_oracle.py) driven by the CPU filter oracle: with the same random keys the
chains are identical -- every trace row (index, natural parameters, log-likelihood, 17
digits), the acceptance decisions, the empirical covariance and the sampled
trajectories of the X file."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle
from oracle import pmcmc_oracle as po
from iip_amd import smc
from test_cli import PKG, model_dir

pytestmark = pytest.mark.gpu

PMCMC = os.path.join(PKG, "cli", "pmcmc")
THETA = os.path.join(PKG, "examples", "sir", "theta.json")


def run_pmcmc(d, args):
    import subprocess
    with open(THETA) as f:
        return subprocess.run([PMCMC] + args, cwd=str(d), stdin=f, capture_output=True, text=True, timeout=600)


def sir_problem(dry):
    """The SIR example's chain: S, I (initial proportions, p_S + p_I <= 1), beta, gamma
    (a duration: rate 1 / (1e-5 + d)), rep, phi, all uniform priors (logit_ab scale)."""
    comps = dry["components"]
    nat = lambda th, o: po.f_inv(comps[o]["tr"], th[o], comps[o]["min"], comps[o]["max"])
    prob = {"theta": dry["theta"], "var": dry["var"], "estimated": dry["estimated"], "components": comps,
            "constrain": lambda th: int(nat(th, 0) + nat(th, 1) > 1.0)}

    def inputs(th):
        N = 1e6
        pv = np.array([nat(th, 2), 1.0 / (0.00001 + nat(th, 3)), nat(th, 4), nat(th, 5), N])
        x0 = np.array([nat(th, 0) * N, nat(th, 1) * N, 0.0])
        return pv, x0
    return prob, inputs


def read_csv(path):
    lines = open(path).read().splitlines()
    return lines[0].split(","), [[float(v) for v in line.split(",")] for line in lines[1:]]


@pytest.mark.parametrize("mode", ["sequential", "full"])
def test_pmcmc_chains_equal_oracle(tmp_path, mode):
    d = model_dir(tmp_path, "sir")
    J, N, M, K = 400, 20, 31, 2
    args = ["sde", "-J", str(J), "-o", str(N), "-M", str(M), "--trace_17g", "-n", "6", "--chains", str(K), "-i", "4"]
    if mode == "full":
        args += ["--full", "-E", "5", "-S", "20"]
    dry = json.loads(run_pmcmc(d, args + ["--dry_run"]).stdout)
    prob, inputs = sir_problem(dry)
    pv0, x00 = inputs(dry["theta"])
    assert list(pv0) == dry["pv"] and list(x00) == dry["x0"]
    print("pmcmc", mode, flush=True)
    r = run_pmcmc(d, args)
    assert r.returncode == 0, r.stderr
    m = smc.Model.get("sir")
    times, data = m.dataset(N)
    for k in range(K):
        seed = 2 + 4 + k

        def loglik(th, it):
            pv, x0 = inputs(th)
            return oracle.smc(m.id, J, times, data, pv, x0, seed=seed, stream=it, nthreads=8)["loglik"]

        o = po.run(prob, loglik, M, seed, full=(mode == "full"), m_eps=5 if mode == "full" else 50,
                   m_switch=20 if mode == "full" else -1)
        head, rows = read_csv(tmp_path / "sir" / ("trace_%d.csv" % (4 + k)))
        assert head == ["index", "S:all", "I:all", "beta:all", "gamma:all", "rep:all", "phi:all", "log_like"]
        assert len(rows) == len(o["trace"])
        for got, (idx, vals, ll) in zip(rows, o["trace"]):
            assert got[0] == idx
            assert got[1:-1] == vals
            assert got[-1] == ll
        assert o["accept"] > 0
        _, cov = read_csv(tmp_path / "sir" / ("covariance_%d.csv" % (4 + k)))
        np.testing.assert_allclose(np.array(cov), o["var_sampling"], rtol=1e-5, atol=1e-12)
        # the X file: at every printed iteration, the lineage of that iteration's filter
        head, xrows = read_csv(tmp_path / "sir" / ("X_%d.csv" % (4 + k)))
        assert head == ["index", "time", "S:city1__all", "I:city1__all", "R:city1__all", "obs_mean:city1__CDC__inc",
                        "obs_real:city1__CDC__inc"]
        printed = sorted({int(row[0]) for row in xrows})
        assert printed[0] == 0 and len(xrows) == len(printed) * (N + 1)
        for idx in printed[:3]:
            it = idx if mode == "full" else 6 * idx
            th = dry["theta"] if it == 0 else o["proposals"][it - 1]
            pv, x0 = inputs(th)
            lo = oracle.smc(m.id, J, times, data, pv, x0, seed=seed, stream=it, lineage=True, nthreads=8)
            block = np.array([row for row in xrows if int(row[0]) == idx])
            np.testing.assert_allclose(block[:, 1], times[::-1], rtol=1e-6)
            lin = lo["lineage"][::-1]
            np.testing.assert_allclose(block[:, 2:4], lin[:, :2], rtol=1e-5)
            np.testing.assert_allclose(block[:, 4], 1e6 - lin[:, 0] - lin[:, 1], rtol=1e-5, atol=1e-3)
            np.testing.assert_allclose(block[:, 5:7], lo["lineage_obs"][::-1, 0, :], rtol=1e-5, atol=1e-9)
        done = json.load(open(tmp_path / "sir" / (".done_%d.json" % (4 + k))))
        final = dict(zip(["S", "I", "beta", "gamma", "rep", "phi"], o["trace"][-1][1]))
        for name, v in final.items():
            assert done["parameter"][name]["group"]["all"]["guess"]["value"] == pytest.approx(v, rel=1e-15)


def test_pmcmc_first_iteration_is_the_smc_filter(tmp_path, ctx):
    """Iteration 0 of chain 0 is the filter of the initial theta with the key (2 + id, 0),
    the `smc` command's key: bit for bit with the inputs of the Bayesian transforms, and
    to their rounding against `smc` (which moves uniform-prior parameters on their own
    transforms instead of logit_ab)."""
    d = model_dir(tmp_path, "sir")
    args = ["sde", "-J", "1000", "-o", "30", "-M", "1", "--trace_17g", "-i", "2", "-n", "0"]
    dry = json.loads(run_pmcmc(d, args + ["--dry_run"]).stdout)
    r = run_pmcmc(d, args)
    assert r.returncode == 0, r.stderr
    _, rows = read_csv(tmp_path / "sir" / "trace_2.csv")
    m = smc.Model.get("sir")
    times, data = m.dataset(30)
    g = smc.run(ctx, m, 1000, times, data, np.array(dry["pv"]), np.array(dry["x0"]), seed=4, stream=0)
    assert rows[0][-1] == g.loglik
    smc_bin = os.path.join(PKG, "cli", "smc")
    import subprocess
    with open(THETA) as f:
        s = subprocess.run([smc_bin, "sde", "-J", "1000", "-o", "30", "-i", "2", "-e", "-d", "-b"], cwd=str(d), stdin=f,
                           capture_output=True, text=True, timeout=600)
    ll = float([l for l in s.stdout.splitlines() if l.startswith("logV_17g:")][0].split(":")[1])
    assert rows[0][-1] == pytest.approx(ll, rel=1e-12)
    assert not os.path.exists(tmp_path / "sir" / "X_2.csv")
