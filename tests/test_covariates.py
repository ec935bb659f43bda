"""Time-varying covariates (par_fixed_values splines, C/core/build.c:1075-1211, evaluated by
the generated code at model_builder/Ccoder.py:39-43), pinned without the GPU:

* the spline of csrc/plom_cov.h against scipy: natural cubic spline above 3 knots
  (CubicSpline, bc_type="natural"), linear interpolation for 2 or 3 (numpy.interp), the value
  for 1; the library's plom_covariate_eval equals the oracle build bit for bit;
* the table layout: row n kmax + k holds the value at the start of substep k of interval n
  (t accumulated as the generated code does), rows N kmax + r the values at times[r];
  times outside the knots are refused;
* constant series give the filter of the constants in pv bit for bit;
* with every noise off, one particle of the noise model (births, deaths, white noise on
  transmission, prevalence and incidence series, covariates N, mu_b, mu_d and prop) is the
  deterministic Euler scheme of its equations with the covariates at every substep start,
  brute-forced here in numpy from scipy splines; its log-likelihood per interval is the
  discretised-normal likelihood (scipy's normal CDF) with prop at the observation time,
  and the remainder of hat is N(t) - S - I.
"""
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.stats import norm

from oracle import oracle
from iip_amd import configs, smc
from iip_amd.smc import Model


def test_natural_cubic_spline_against_scipy():
    rng = np.random.default_rng(11)
    for n in (4, 5, 9, 40):
        x = np.cumsum(rng.uniform(0.5, 10.0, n))
        y = rng.uniform(-3.0, 5.0, n) * 10.0 ** rng.integers(-4, 6)
        t = np.concatenate([rng.uniform(x[0], x[-1], 300), x, [x[0], x[-1]]])
        ref = CubicSpline(x, y, bc_type="natural")(t)
        got = oracle.spline(x, y, t)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(y).max())
        # a knot that starts a segment is interpolated exactly (dx = 0); the last one ends the
        # last segment and is within the tolerance above
        np.testing.assert_array_equal(got[300:300 + n - 1], y[:-1])
        np.testing.assert_array_equal(smc.covariate_eval(x, y, t), got)


def test_linear_and_constant_series():
    rng = np.random.default_rng(12)
    for n in (2, 3):
        x = np.cumsum(rng.uniform(1.0, 30.0, n))
        y = rng.uniform(0.0, 1e6, n)
        t = np.concatenate([rng.uniform(x[0], x[-1], 200), x])
        got = oracle.spline(x, y, t)
        np.testing.assert_allclose(got, np.interp(t, x, y), rtol=0, atol=2e-15 * np.abs(y).max())
        np.testing.assert_array_equal(smc.covariate_eval(x, y, t), got)
    np.testing.assert_array_equal(oracle.spline([3.0], [7.5], [-100.0, 0.0, 1e9]), [7.5, 7.5, 7.5])


def test_eval_refuses_times_outside_the_knots():
    with pytest.raises(smc.PlomError):
        smc.covariate_eval([0.0, 7.0, 14.0], [1.0, 2.0, 3.0], [14.5])
    with pytest.raises(smc.PlomError):
        smc.covariate_eval([0.0, 7.0, 7.0], [1.0, 2.0, 3.0], [1.0])   # knots must increase strictly


def noise_cov(m, scale=None):
    cov = m.covariates()
    if scale is not None:
        cov = [(x, y * scale) for x, y in cov]
    return cov


def test_table_layout():
    m = Model.get("noise")
    cov = noise_cov(m)
    times = np.array([0.0, 7.0, 15.0, 21.5, 30.0, 31.0])
    dt = 0.25
    tab, kmax = oracle.cov_table(m.id, times, dt, cov)
    N = len(times) - 1
    Ks = []
    for n in range(N):
        t, K = times[n], 0
        while t < times[n + 1]:
            t += dt
            K += 1
        Ks.append(K)
    assert kmax == max(Ks) and tab.shape == (N * kmax + N + 1, m.n_fix)
    for i, (x, y) in enumerate(cov):
        sp = (lambda tt, x=x, y=y: oracle.spline(x, y, tt))
        for n in range(N):
            ts, t = [], times[n]
            for k in range(Ks[n]):
                ts.append(t)
                t += dt
            np.testing.assert_array_equal(tab[n * kmax:n * kmax + Ks[n], i], sp(np.array(ts)))
            np.testing.assert_array_equal(tab[n * kmax + Ks[n]:(n + 1) * kmax, i], 0.0)   # padding
        np.testing.assert_array_equal(tab[N * kmax:, i], sp(times))
    # the knots end at 728 days: a later observation is refused
    with pytest.raises(ValueError):
        oracle.cov_table(m.id, np.array([0.0, 700.0, 735.0]), dt, cov)


def test_constant_series_equal_the_constants_of_pv():
    m = Model.get("sir")
    times, data = m.dataset(30)
    pv, x0 = m.pack(configs.TRUTH["sir"]), m.x0(configs.X0_PROPS["sir"])
    fo = m.layout["fix_off"]
    cov = [(np.array([0.0, 100.0, 300.0]), np.full(3, pv[fo + i])) for i in range(m.n_fix)]
    a = oracle.smc(m.id, 500, times, data, pv, x0, hat=True, lineage=True, ancestry=True)
    b = oracle.smc(m.id, 500, times, data, pv, x0, hat=True, lineage=True, ancestry=True, cov=cov)
    for k in ("loglik_n", "ess_n", "ancestry", "hat", "pred_res", "lineage", "lineage_obs"):
        np.testing.assert_array_equal(a[k], b[k])
    assert a["loglik"] == b["loglik"]


def euler_noise(m, pv, x0, times, dt, cov):
    """Deterministic Euler of the noise model with the covariates at every substep start,
    from scipy splines (model_builder/example/noise/process.json: births mu_b N, infection
    r0/N v I S, recovery v I, deaths mu_d S and mu_d I; inc_out = recoveries + deaths of I
    over both cities, inc_in = infections in city 2)."""
    lay = m.layout
    fo = lay["fix_off"]
    sp = [CubicSpline(x, y, bc_type="natural") if len(x) >= 4 else
          (lambda t, x=x, y=y: np.interp(t, x, y)) if len(x) > 1 else (lambda t, y=y: y[0]) for x, y in cov]
    c = lambda name, cac, t: float(sp[lay["fixed"][name] + cac - fo](t))
    r0, v = pv[lay["proc"]["r0"]], pv[lay["proc"]["v"]]
    S, I = [x0[0], x0[1]], [x0[2], x0[3]]
    rows = [np.concatenate([S, I, [0.0, 0.0]])]
    for n in range(len(times) - 1):
        inc_out, inc_in = 0.0, 0.0
        t = times[n]
        while t < times[n + 1]:
            for cac in (0, 1):
                N, mb, md = c("N", cac, t), c("mu_b", cac, t), c("mu_d", cac, t)
                inf = r0 / N * v * I[cac] * S[cac]
                dS = mb * N - inf - md * S[cac]
                dI = inf - v * I[cac] - md * I[cac]
                inc_out += (v * I[cac] + md * I[cac]) * dt
                if cac == 1:
                    inc_in += inf * dt
                S[cac] = max(S[cac] + dS * dt, 0.0)
                I[cac] = max(I[cac] + dI * dt, 0.0)
            t += dt
        rows.append(np.concatenate([S, I, [inc_out, inc_in]]))
    return np.array(rows), sp


def test_deterministic_path_and_likelihood_with_covariates():
    m = Model.get("noise")
    N = 30
    times, data = m.dataset(N)
    cov = noise_cov(m)
    pv, x0 = m.pack(configs.TRUTH["noise"]), m.x0(configs.X0_PROPS["noise"])
    ref, sp = euler_noise(m, pv, x0, times, 0.25, cov)
    got = oracle.trajectory(m.id, times, pv, x0, noises_off=3, cov=cov)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-6)
    # one particle, no noise: loglik_n is the log-likelihood of its path
    o = oracle.smc(m.id, 1, times, data, pv, x0, noises_off=3, cov=cov, hat=True)
    lay, fo = m.layout, m.layout["fix_off"]
    rep, phi = pv[lay["obs"]["rep"]], pv[lay["obs"]["phi"]]
    for n in range(N):
        x = ref[n + 1]
        obs = [x[4], x[4], x[5], x[2]]               # inc_out (CDC, google), inc_in, prevalence city 1
        ll = 0.0
        for ts in range(4):
            y = data[n, ts]
            if math.isnan(y):
                continue
            prop = float(sp[lay["fixed_obs"]["prop"] + ts - fo](times[n + 1]))
            mu = rep * prop * obs[ts]
            sd = math.sqrt(rep * (1 - rep) * prop * obs[ts] + (rep * phi * prop * obs[ts]) ** 2)
            like = norm.cdf(y + 0.5, mu, sd) - (norm.cdf(y - 0.5, mu, sd) if y > 0 else 0.0)
            ll += math.log(max(like, 1e-17))
        assert o["loglik_n"][n] == pytest.approx(ll, rel=1e-9, abs=1e-9), n
        # hat of one particle: the remainder of city c is N(t) - S - I
        for cac in (0, 1):
            Nt = float(sp[lay["fixed"]["N"] + cac - fo](times[n + 1]))
            assert o["hat"][n + 1, 4 + cac, 1] == pytest.approx(Nt - x[cac] - x[2 + cac], rel=1e-10)
