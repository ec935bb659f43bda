"""The ode implementation (PLOM_FLAG_ODE: adaptive RKF45 per particle, csrc/plom_ode.h) on
the GPU against the CPU oracle, which is pinned in test_ode.py: bit-identical FP64
log-likelihoods, ESS, hat and lineages; per-particle parameters (iterated filtering), where
every lane integrates its own system with its own step sizes; batched with stochastic
filters; time-varying covariates evaluated as splines at the integrator's times (alone,
iterated filtering, batched, and the `smc ode` command); the refusals (FP32, PSR)."""
import numpy as np
import pytest

from oracle import oracle
from iip_amd import configs, smc
from iip_amd.smc import Model
from test_gpu_mif import check_equal, mif_setup

pytestmark = pytest.mark.gpu


def setup(name, N):
    m = Model.get(name)
    times, data = m.dataset(N)
    return m, times, data, m.pack(configs.TRUTH[name]), m.x0(configs.X0_PROPS[name])


@pytest.mark.parametrize("name,J,N,eps", [("sir", 1, 52, (0.0, 0.0)), ("sirs4", 1000, 60, (0.0, 0.0)),
                                          ("seir10", 300, 20, (1e-8, 1e-8)), ("noise", 257, 104, (0.0, 0.0)),
                                          ("sir", 64, 30, (1e-3, 1e-2))])
def test_ode_filter_equals_oracle(ctx, name, J, N, eps):
    m, times, data, pv, x0 = setup(name, N)
    print("ode", name, J, N, eps, flush=True)
    f = smc.Filter(ctx, m, J, times, data, pv, x0, flags=smc.FLAG_ODE | smc.FLAG_HAT | smc.FLAG_LINEAGE,
                   eps_abs=eps[0], eps_rel=eps[1], seed=3)
    f.run()
    g = f.read(hat=True, lineage=True)
    f.close()
    o = oracle.smc(m.id, J, times, data, pv, x0, ode=True, eps_abs=eps[0], eps_rel=eps[1], seed=3, hat=True,
                   lineage=True, nthreads=8)
    np.testing.assert_array_equal(g.loglik_n, o["loglik_n"])
    np.testing.assert_array_equal(g.ess_n, o["ess_n"])
    np.testing.assert_array_equal(g.hat, o["hat"])
    np.testing.assert_array_equal(g.lineage, o["lineage"])
    assert g.loglik == o["loglik"]
    np.testing.assert_array_equal(g.ess_n[~np.isnan(data).all(axis=1)], float(J))


def test_ode_mif_pass_equals_oracle(ctx):
    """Per-particle parameters: every lane its own ODE, step sizes and step counts."""
    m, times, data, pv, mp, theta0, x0, mut = mif_setup("sirs4", 30, 1500, 5)
    print("ode mif", flush=True)
    f = smc.Filter(ctx, m, 1500, times, data, pv, x0[:, 0], flags=smc.FLAG_MIF | smc.FLAG_ODE, mif=mp)
    f.mif_start(theta0, x0, mut, seed=4, stream=9)
    f.run()
    g = f.read(mif=True)
    f.close()
    o = oracle.mif_pass(m.id, 1500, times, data, pv, mp, theta0, x0, mut, seed=4, stream=9, nthreads=8, ode=True)
    check_equal(g, o)


@pytest.mark.parametrize("lanes", [1, 2])
def test_ode_batched_with_stochastic_filters(ctx, lanes):
    cases = [("sir", 500, 40, smc.FLAG_ODE), ("sirs4", 2000, 30, 0), ("noise", 800, 50, smc.FLAG_ODE),
             ("sirs4", 700, 25, smc.FLAG_PSR)]
    fs, alone = [], []
    for k, (name, J, N, flags) in enumerate(cases):
        m, times, data, pv, x0 = setup(name, N)
        if flags & smc.FLAG_PSR:
            x0 = np.round(x0)
        print("ode batch case", name, J, N, flags, flush=True)
        fs.append(smc.Filter(ctx, m, J, times, data, pv, x0, flags=flags, seed=7, stream=k))
        alone.append(smc.run(ctx, m, J, times, data, pv, x0, flags=flags, seed=7, stream=k))
    b = smc.Batch(ctx, fs)
    b.configure(lanes)
    b.launch(use_graph=True)
    for f, a in zip(fs, alone):
        r = f.read()
        np.testing.assert_array_equal(r.loglik_n, a.loglik_n)
        assert r.loglik == a.loglik
    b.close()
    for f in fs:
        f.close()


def test_ode_refusals(ctx):
    m, times, data, pv, x0 = setup("noise", 10)
    with pytest.raises(smc.PlomError):
        smc.Filter(ctx, m, 10, times, data, pv, x0, precision="fp32", flags=smc.FLAG_ODE)
    with pytest.raises(smc.PlomError):
        smc.Filter(ctx, m, 10, times, data, pv, x0, flags=smc.FLAG_ODE | smc.FLAG_PSR)


@pytest.mark.parametrize("J,N,eps", [(257, 104, (0.0, 0.0)), (1, 60, (1e-8, 1e-8))])
def test_ode_with_covariates_equals_oracle(ctx, J, N, eps):
    m, times, data, pv, x0 = setup("noise", N)
    cov = m.covariates()
    print("ode cov", J, N, eps, flush=True)
    f = smc.Filter(ctx, m, J, times, data, pv, x0, flags=smc.FLAG_ODE | smc.FLAG_HAT | smc.FLAG_LINEAGE,
                   eps_abs=eps[0], eps_rel=eps[1], seed=3, cov=cov)
    f.run()
    g = f.read(hat=True, lineage=True)
    f.close()
    o = oracle.smc(m.id, J, times, data, pv, x0, ode=True, eps_abs=eps[0], eps_rel=eps[1], seed=3, hat=True,
                   lineage=True, nthreads=8, cov=cov)
    np.testing.assert_array_equal(g.loglik_n, o["loglik_n"])
    np.testing.assert_array_equal(g.hat, o["hat"])
    np.testing.assert_array_equal(g.lineage, o["lineage"])
    assert g.loglik == o["loglik"]
    # the splines matter: constant covariates give another path
    assert g.loglik != oracle.smc(m.id, J, times, data, pv, x0, ode=True, eps_abs=eps[0], eps_rel=eps[1], seed=3)["loglik"]


def test_ode_mif_pass_with_covariates_equals_oracle(ctx):
    m, times, data, pv, mp, theta0, x0, mut = mif_setup("noise", 30, 700, 2)
    cov = m.covariates()
    print("ode mif cov", flush=True)
    f = smc.Filter(ctx, m, 700, times, data, pv, x0[:, 0], flags=smc.FLAG_MIF | smc.FLAG_ODE, mif=mp, cov=cov)
    f.mif_start(theta0, x0, mut, seed=4, stream=9)
    f.run()
    g = f.read(mif=True)
    f.close()
    o = oracle.mif_pass(m.id, 700, times, data, pv, mp, theta0, x0, mut, seed=4, stream=9, nthreads=8, ode=True,
                        cov=cov)
    check_equal(g, o)


def test_ode_with_covariates_batched(ctx):
    m, times, data, pv, x0 = setup("noise", 50)
    cov = m.covariates()
    cases = [(smc.FLAG_ODE, cov, 300), (0, cov, 1200), (smc.FLAG_ODE, None, 200)]
    fs, alone = [], []
    for k, (flags, cv, J) in enumerate(cases):
        print("ode cov batch", k, flush=True)
        fs.append(smc.Filter(ctx, m, J, times, data, pv, x0, flags=flags, seed=5, stream=k, cov=cv))
        alone.append(oracle.smc(m.id, J, times, data, pv, x0, ode=bool(flags), seed=5, stream=k, cov=cv, nthreads=8))
    b = smc.Batch(ctx, fs)
    b.configure(1)
    b.launch(use_graph=True)
    for f, o in zip(fs, alone):
        r = f.read()
        np.testing.assert_array_equal(r.loglik_n, o["loglik_n"])
    b.close()
    for f in fs:
        f.close()


def test_smc_ode_command_with_covariates(tmp_path, ctx):
    """`smc ode` on the noise example (time-varying N, mu_b, mu_d, prop): the command's
    log-likelihood is the library's and the oracle's with the model's covariate series."""
    import json
    import os
    from test_cli import PKG, model_dir, run_cli
    d = model_dir(tmp_path, "noise")
    theta = os.path.join(PKG, "examples", "noise", "theta.json")
    m = Model.get("noise")
    times, data = m.dataset()
    cov = m.covariates()
    dry = json.loads(run_cli(d, ["ode", "-J", "1", "--dry_run"], theta).stdout)
    pv, x0 = np.array(dry["pv"]), np.array(dry["x0"])
    print("smc ode cov command", flush=True)
    r = run_cli(d, ["ode", "-J", "1", "-b", "-e", "-d"], theta)
    assert r.returncode == 0, r.stderr
    ll = float([line for line in r.stdout.splitlines() if line.startswith("logV_17g:")][0].split(":")[1])
    g = smc.run(ctx, m, 1, times, data, pv, x0, flags=smc.FLAG_ODE, seed=2, stream=0, cov=cov)
    o = oracle.smc(m.id, 1, times, data, pv, x0, ode=True, seed=2, stream=0, cov=cov)
    assert ll == g.loglik == o["loglik"]
