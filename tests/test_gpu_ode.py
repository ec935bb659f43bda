"""The ode implementation (PLOM_FLAG_ODE: adaptive RKF45 per particle, csrc/plom_ode.h) on
the GPU against the CPU oracle, which is pinned in test_ode.py: bit-identical FP64
log-likelihoods, ESS, hat and lineages; per-particle parameters (iterated filtering), where
every lane integrates its own system with its own step sizes; batched with stochastic
filters; the refusals (FP32, PSR, time-varying covariates)."""
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
    with pytest.raises(smc.PlomError):
        smc.Filter(ctx, m, 10, times, data, pv, x0, flags=smc.FLAG_ODE, cov=m.covariates())
