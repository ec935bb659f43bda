"""Time-varying covariates on the GPU (the covariate table read with scalar loads at every
substep and observation time) against the CPU oracle, which is pinned in
test_covariates.py.  Model: the noise example (population sizes, birth and death rates and
proportions under surveillance that change with time).  FP64: bit-exact log-likelihoods,
ancestors, hat, residuals, lineages and MIF moments, alone and batched; FP32: the table
path of constant series equals the constants of pv bit for bit, and time-varying series
give the FP64 log-likelihood within Monte Carlo error."""
import numpy as np
import pytest

from oracle import oracle
from iip_amd import configs, smc
from iip_amd.smc import Model
from test_gpu_mif import check_equal, mif_setup

pytestmark = pytest.mark.gpu


def setup(N=104):
    m = Model.get("noise")
    times, data = m.dataset(N)
    return m, times, data, m.pack(configs.TRUTH["noise"]), m.x0(configs.X0_PROPS["noise"]), m.covariates()


@pytest.mark.parametrize("J,N,noises_off,flags", [(2000, 104, 0, 0), (257, 40, 1, 0), (1000, 60, 0, smc.FLAG_PSR),
                                                  (700, 30, 2, smc.FLAG_NO_FILTER)])
def test_filter_equals_oracle(ctx, J, N, noises_off, flags):
    m, times, data, pv, x0, cov = setup(N)
    if flags & smc.FLAG_PSR:
        x0 = np.round(x0)
    print("cov filter", J, N, noises_off, flags, flush=True)
    g = smc.run(ctx, m, J, times, data, pv, x0, ancestry=True, final_state=True, noises_off=noises_off, flags=flags,
                seed=4, stream=2, cov=cov)
    o = oracle.smc(m.id, J, times, data, pv, x0, noises_off=noises_off, flags=flags, seed=4, stream=2, ancestry=True,
                   final_state=True, nthreads=8, cov=cov, psr=bool(flags & smc.FLAG_PSR))
    np.testing.assert_array_equal(g.ancestry, o["ancestry"])
    np.testing.assert_array_equal(g.loglik_n, o["loglik_n"])
    np.testing.assert_array_equal(g.ess_n, o["ess_n"])
    np.testing.assert_array_equal(g.final_state, o["final_state"])
    assert g.loglik == o["loglik"]
    # the covariates matter: the constants at t = 0 give another likelihood
    if not flags & smc.FLAG_NO_FILTER:
        c = oracle.smc(m.id, J, times, data, pv, x0, noises_off=noises_off, flags=flags, seed=4, stream=2, nthreads=8,
                       psr=bool(flags & smc.FLAG_PSR))
        assert c["loglik"] != o["loglik"]


def test_hat_and_lineage_equal_oracle(ctx):
    m, times, data, pv, x0, cov = setup(52)
    print("cov hat + lineage", flush=True)
    f = smc.Filter(ctx, m, 1500, times, data, pv, x0, flags=smc.FLAG_HAT | smc.FLAG_LINEAGE, seed=6, stream=1,
                   cov=cov)
    f.run()
    g = f.read(hat=True, lineage=True)
    f.close()
    o = oracle.smc(m.id, 1500, times, data, pv, x0, seed=6, stream=1, hat=True, lineage=True, nthreads=8, cov=cov)
    np.testing.assert_array_equal(g.hat, o["hat"])
    np.testing.assert_array_equal(g.pred_res, o["pred_res"])
    np.testing.assert_array_equal(g.lineage, o["lineage"])
    np.testing.assert_array_equal(g.lineage_obs, o["lineage_obs"])
    assert g.loglik == o["loglik"]


def test_mif_pass_equals_oracle(ctx):
    m, times, data, pv, mp, theta0, x0, mut = mif_setup("noise", 40, 1200, 3)
    cov = m.covariates()
    print("cov mif", flush=True)
    f = smc.Filter(ctx, m, 1200, times, data, pv, x0[:, 0], flags=smc.FLAG_MIF, mif=mp, cov=cov)
    f.mif_start(theta0, x0, mut, seed=8, stream=5)
    f.run()
    g = f.read(mif=True)
    f.close()
    o = oracle.mif_pass(m.id, 1200, times, data, pv, mp, theta0, x0, mut, seed=8, stream=5, nthreads=8, cov=cov)
    check_equal(g, o)


@pytest.mark.parametrize("lanes", [1, 2])
def test_batched_with_constant_filters_equals_alone(ctx, lanes):
    """A plan mixing filters with and without covariate tables: each equals its run alone."""
    m, times, data, pv, x0, cov = setup(60)
    s = Model.get("sirs4")
    st, sd = s.dataset(40)
    cases = [(m, times, data, pv, x0, cov, 1500, 1), (s, st, sd, s.pack(configs.TRUTH["sirs4"]),
                                                      s.x0(configs.X0_PROPS["sirs4"]), None, 2100, 2),
             (m, times, data, pv, x0, None, 900, 3), (m, times[:31], data[:30], pv, x0, cov, 3000, 4)]
    fs, alone = [], []
    for mm, t, d, p, x, c, J, stream in cases:
        print("cov batch case", mm.name, J, stream, c is not None, flush=True)
        fs.append(smc.Filter(ctx, mm, J, t, d, p, x, seed=11, stream=stream, cov=c))
        alone.append(smc.run(ctx, mm, J, t, d, p, x, seed=11, stream=stream, cov=c))
    b = smc.Batch(ctx, fs)
    b.configure(lanes)
    for rep in range(2):
        b.launch(use_graph=True)
        for f, a in zip(fs, alone):
            r = f.read()
            np.testing.assert_array_equal(r.loglik_n, a.loglik_n)
            assert r.loglik == a.loglik
    b.close()
    for f in fs:
        f.close()


def test_fp32_table_path(ctx):
    m, times, data, pv, x0, cov = setup(52)
    fo = m.fix_off
    const = [(np.array([0.0, 200.0, 400.0]), np.full(3, pv[fo + i])) for i in range(m.n_fix)]
    print("cov fp32 constant table", flush=True)
    a = smc.run(ctx, m, 3000, times, data, pv, x0, precision="fp32", seed=2, stream=7)
    b = smc.run(ctx, m, 3000, times, data, pv, x0, precision="fp32", seed=2, stream=7, cov=const)
    np.testing.assert_array_equal(a.loglik_n, b.loglik_n)
    # time-varying: FP32 propagation is another Monte Carlo estimate of the FP64 likelihood
    l32, l64 = [], []
    for r in range(8):
        l32.append(smc.run(ctx, m, 20000, times, data, pv, x0, precision="fp32", seed=40 + r, cov=cov).loglik)
        l64.append(smc.run(ctx, m, 20000, times, data, pv, x0, seed=80 + r, cov=cov).loglik)
    se = np.sqrt(np.var(l32, ddof=1) / 8 + np.var(l64, ddof=1) / 8)
    assert abs(np.mean(l32) - np.mean(l64)) < 4 * se + 0.05, (l32, l64)


def test_served_requests_with_covariates_equal_alone(ctx):
    """Served requests take time-varying covariates (sde, psr and ode) and drift maps,
    copied at submission; each result equals the same filter run alone, bit for bit, mixed
    with requests without them.  Covariates that do not cover the times are refused at
    submission."""
    from iip_amd import configs
    m, times, data, pv, x0, cov = setup(60)
    dm = Model.get("drift")
    dtimes, ddata = dm.dataset(40)
    dpv = dm.pack(configs.TRUTH["drift"])
    dx0, drift = configs.drift_setup(dm, dm.x0(configs.X0_PROPS["drift"]), configs.TRUTH["drift"])
    s = Model.get("sirs4")
    st, sd = s.dataset(30)
    reqs = [dict(model=m, J=900, times=times, data=data, pv=pv, x0=x0, cov=cov),
            dict(model=s, J=2000, times=st, data=sd, pv=s.pack(configs.TRUTH["sirs4"]), x0=s.x0(configs.X0_PROPS["sirs4"])),
            dict(model=m, J=300, times=times, data=data, pv=pv, x0=x0, cov=cov, flags=smc.FLAG_ODE),
            dict(model=m, J=700, times=times[:31], data=data[:30], pv=pv, x0=np.round(x0), cov=cov, flags=smc.FLAG_PSR),
            dict(model=dm, J=1200, times=dtimes, data=ddata, pv=dpv, x0=dx0, cov=dm.covariates(), drift=drift),
            dict(model=m, J=500, times=times, data=data, pv=pv, x0=x0)]
    srv = smc.Server(ctx)
    try:
        ids = []
        for k, r in enumerate(reqs):
            print("served cov", k, r["model"].name, r.get("flags", 0), flush=True)
            ids.append(srv.submit(r["model"], r["J"], r["times"], r["data"], r["pv"], r["x0"], seed=4, stream=k,
                                  flags=r.get("flags", 0), cov=r.get("cov"), drift=r.get("drift")))
        for _ in range(100000):
            srv.step()
            if all(srv.status(i) == smc.SRV_DONE for i in ids):
                break
        for k, (r, i) in enumerate(zip(reqs, ids)):
            got = srv.result(i)
            a = smc.run(ctx, r["model"], r["J"], r["times"], r["data"], r["pv"], r["x0"], seed=4, stream=k,
                        flags=r.get("flags", 0), cov=r.get("cov"), drift=r.get("drift"))
            np.testing.assert_array_equal(got.loglik_n, a.loglik_n)
            assert got.loglik == a.loglik
        short = [(x[:2], y[:2]) if len(x) > 1 else (x, y) for x, y in cov]
        with pytest.raises(smc.PlomError):
            srv.submit(m, 100, times, data, pv, x0, cov=short)
    finally:
        srv.close()
