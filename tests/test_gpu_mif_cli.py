"""The `mif` command (iterated filtering, C/mif/mif.c:21-247) on the GPU against the CPU
restatement of its driver (oracle/mif_oracle.py) driven by the CPU filter oracle with the
same keys: identical trace rows (17 digits) in both update regimes (m <= SWITCH and
after), for two replicates run as one batch, and the mif file of -r."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle, mif_oracle
from iip_amd import smc
from test_cli import PKG, model_dir

pytestmark = pytest.mark.gpu

MIF = os.path.join(PKG, "cli", "mif")
THETA = os.path.join(PKG, "examples", "sir", "theta.json")


def run_mif(d, args):
    with open(THETA) as f:
        return subprocess.run([MIF] + args, cwd=str(d), stdin=f, capture_output=True, text=True, timeout=600)


def read_csv(path):
    lines = open(path).read().splitlines()
    return lines[0].split(","), [[float(v) for v in line.split(",")] for line in lines[1:]]


def test_mif_command_equals_restatement(tmp_path):
    d = model_dir(tmp_path, "sir")
    J, M, K = 400, 4, 2
    args = ["sde", "-J", str(J), "-M", str(M), "-S", "2", "-a", "0.9", "-r", "--trace_17g", "--chains", str(K), "-i", "3"]
    dry = json.loads(run_mif(d, args + ["--dry_run"]).stdout)
    print("mif command", flush=True)
    r = run_mif(d, args)
    assert r.returncode == 0, r.stderr
    m = smc.Model.get("sir")
    N = dry["N"]
    times, data = m.dataset(N)
    mp = {k: dry[k] for k in ("pv_src", "tr", "scale", "duration", "min", "max", "multiplier", "sd")}
    pv = np.array(dry["pv"])

    def pass_fn(th, x0, mut, seed, stream):
        return oracle.mif_pass(m.id, th.shape[1], times, data, pv, mp, th, x0, mut, seed=seed, stream=stream, nthreads=8)

    def x0_of(th):
        return [mif_oracle.f_inv(2, th[0], 0, 0) * 1e6, mif_oracle.f_inv(2, th[1], 0, 0) * 1e6, 0.0]

    def constrain(th):
        return int(mif_oracle.f_inv(2, th[0], 0, 0) + mif_oracle.f_inv(2, th[1], 0, 0) > 1.0)

    for k in range(K):
        o = mif_oracle.run(dry, pass_fn, x0_of, constrain, M=M, J=J, seed=2 + 3 + k, a=0.9, switch=2, print_traj=True)
        head, rows = read_csv(tmp_path / "sir" / ("trace_%d.csv" % (3 + k)))
        assert head == ["index", "S:all", "I:all", "beta:all", "gamma:all", "rep:all", "phi:all", "log_like"]
        assert len(rows) == M + 1
        assert math.isnan(rows[0][-1])
        for got, (idx, vals, ll) in zip(rows, o["trace"]):
            assert got[0] == idx
            assert got[1:-1] == vals
            if idx > 0:
                assert got[-1] == ll
        head, mrows = read_csv(tmp_path / "sir" / ("mif_%d.csv" % (3 + k)))
        assert head[:4] == ["index", "time", "mean:S:all", "var:S:all"] and head[-1] == "ess"
        assert len(mrows) == M * len(dry["data_rows"])
        done = json.load(open(tmp_path / "sir" / (".done_%d.json" % (3 + k))))
        assert done["parameter"]["beta"]["group"]["all"]["guess"]["value"] == pytest.approx(o["trace"][-1][1][2], rel=1e-15)


def test_mif_command_with_a_covariance(tmp_path):
    """A "covariance" in theta: multivariate starting proposals and a correlated random walk
    (mif.c:84-121, methods.c:275-330); the command's trace equals the restatement's."""
    d = model_dir(tmp_path, "sir")
    th = json.load(open(THETA))
    names = ["S", "I", "beta", "gamma", "rep", "phi"]
    sd = [0.01, 0.01, 0.02, 0.02, 0.02, 0.02]
    cov = [[sd[a] * sd[b] * (1.0 if a == b else (0.6 if {a, b} == {2, 3} else (0.3 if {a, b} == {2, 4} else 0.0)))
            for b in range(6)] for a in range(6)]
    th["covariance"] = [names] + cov
    tpath = tmp_path / "theta_cov.json"
    tpath.write_text(json.dumps(th) + "\n")
    J, M = 300, 3
    args = ["sde", "-J", str(J), "-M", str(M), "-S", "1", "--trace_17g", "-i", "0"]

    def run(a):
        with open(tpath) as f:
            return subprocess.run([MIF] + a, cwd=str(d), stdin=f, capture_output=True, text=True, timeout=600)
    dry = json.loads(run(args + ["--dry_run"]).stdout)
    assert len(dry["chol"]) == 36 and dry["chol"][2 * 6 + 3] == 0.0 and dry["chol"][3 * 6 + 2] != 0.0
    print("mif command mvn", flush=True)
    r = run(args)
    assert r.returncode == 0, r.stderr
    m = smc.Model.get("sir")
    times, data = m.dataset(dry["N"])
    mp = {k: dry[k] for k in ("pv_src", "tr", "scale", "duration", "min", "max", "multiplier", "sd")}
    mp["chol"] = np.array(dry["chol"]).reshape(6, 6)
    pv = np.array(dry["pv"])

    def pass_fn(thp, x0, mut, seed, stream):
        return oracle.mif_pass(m.id, thp.shape[1], times, data, pv, mp, thp, x0, mut, seed=seed, stream=stream, nthreads=8)

    def x0_of(t):
        return [mif_oracle.f_inv(2, t[0], 0, 0) * 1e6, mif_oracle.f_inv(2, t[1], 0, 0) * 1e6, 0.0]

    def constrain(t):
        return int(mif_oracle.f_inv(2, t[0], 0, 0) + mif_oracle.f_inv(2, t[1], 0, 0) > 1.0)

    o = mif_oracle.run(dry, pass_fn, x0_of, constrain, M=M, J=J, seed=2, switch=1)
    head, rows = read_csv(tmp_path / "sir" / "trace_0.csv")
    assert len(rows) == M + 1
    for (idx, vals, ll), row in zip(o["trace"], rows):
        assert row[0] == idx
        np.testing.assert_array_equal(row[1:-1], vals)
        if idx:
            assert row[-1] == ll


def test_mif_command_with_a_prior(tmp_path):
    """mif --prior: logit_ab for the uniform and pseudo-uniform priors, the prior patch of the
    weights (patch_likelihood_prior, methods.c:469-523) on the device; the command's trace
    equals the restatement's with the same map."""
    d = model_dir(tmp_path, "sir")
    th = json.load(open(THETA))
    th["parameter"]["beta"]["group"]["all"]["prior"]["value"] = "normal"
    th["parameter"]["rep"]["group"]["all"]["prior"]["value"] = "pseudo_uniform"
    tpath = tmp_path / "theta_prior.json"
    tpath.write_text(json.dumps(th) + "\n")
    J, M = 400, 3
    args = ["sde", "-J", str(J), "-M", str(M), "-S", "1", "--trace_17g", "-i", "0", "--prior"]

    def run(a):
        with open(tpath) as f:
            return subprocess.run([MIF] + a, cwd=str(d), stdin=f, capture_output=True, text=True, timeout=600)
    dry = json.loads(run(args + ["--dry_run"]).stdout)
    assert dry["prior"] == [0, 0, 1, 0, 2, 0] and dry["tr"] == [3, 3, 1, 3, 3, 3]
    print("mif command prior", flush=True)
    r = run(args)
    assert r.returncode == 0, r.stderr
    m = smc.Model.get("sir")
    times, data = m.dataset(dry["N"])
    mp = {k: dry[k] for k in ("pv_src", "tr", "scale", "duration", "min", "max", "multiplier", "sd", "prior",
                              "prior_pow", "prior_until")}
    pv = np.array(dry["pv"])
    tr, mn, mx = dry["tr"], dry["min"], dry["max"]

    def pass_fn(thp, x0, mut, seed, stream):
        return oracle.mif_pass(m.id, thp.shape[1], times, data, pv, mp, thp, x0, mut, seed=seed, stream=stream, nthreads=8)

    def x0_of(t):
        return [mif_oracle.f_inv(tr[o], t[o], mn[o], mx[o]) * 1e6 for o in (0, 1)] + [0.0]

    def constrain(t):
        return int(mif_oracle.f_inv(tr[0], t[0], mn[0], mx[0]) + mif_oracle.f_inv(tr[1], t[1], mn[1], mx[1]) > 1.0)

    o = mif_oracle.run(dry, pass_fn, x0_of, constrain, M=M, J=J, seed=2, switch=1)
    head, rows = read_csv(tmp_path / "sir" / "trace_0.csv")
    assert len(rows) == M + 1
    for (idx, vals, ll), row in zip(o["trace"], rows):
        assert row[0] == idx
        np.testing.assert_array_equal(row[1:-1], vals)
        if idx:
            assert row[-1] == ll
