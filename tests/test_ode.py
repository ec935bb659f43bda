"""The ode implementation (adaptive Runge-Kutta-Fehlberg 4(5), csrc/plom_ode.h), pinned on
the CPU:

* the integrator on analytic systems: the fifth-order solution integrates y' = p(t) exactly
  for deg p <= 4 (a coefficient typo breaks this); its error estimate vanishes for
  deg p <= 3 (the fourth-order weights are exact there), so the step grows by the maximum
  factor 5 at every step; one step on y' = -y has a local error of order h^6; with a tight
  tolerance, y' = -y and the harmonic oscillator follow their closed forms; the error
  shrinks with the tolerance;
* a model: with tight tolerances the SIR and noise models' ODE paths match scipy's DOP853
  on the same equations (written out here from the model files), incidences included; at
  the reference's default tolerances (1e-6, 1e-3) within 1e-3 relative;
* the filter: every particle follows the same path, so ESS = J and the per-interval
  log-likelihood is the one-path likelihood, exactly;
* time-varying covariates: the right-hand side evaluates the covariates' splines at the
  integrator's own times (the reference's step_ode calls gsl_spline_eval at t); with tight
  tolerances the noise model's path matches scipy's DOP853 with scipy's natural cubic
  splines in the right-hand side, and differs from the path with the covariates frozen at
  their t = 0 values.
"""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.interpolate import CubicSpline

from oracle import oracle
from iip_amd import configs
from iip_amd.smc import Model


def test_polynomials_are_integrated_exactly():
    p = [1.0, -2.0, 0.5, 3.0, -1.0]
    F = lambda t: p[0] * t + p[1] * t ** 2 / 2 + p[2] * t ** 3 / 3 + p[3] * t ** 4 / 4 + p[4] * t ** 5 / 5
    y, n4, rc = oracle.ode_test(1, p, [0.25], 0.0, 2.0, 0.3, 1e-9, 1e-9)
    assert rc == 0 and y[0] == pytest.approx(0.25 + F(2.0), rel=1e-14, abs=1e-14)
    # cubic: zero error estimate -> every step 5 times the previous one (0.01, 0.05, ... then cut at t1)
    q = [1.0, -2.0, 0.5, 3.0, 0.0]
    G = lambda t: q[0] * t + q[1] * t ** 2 / 2 + q[2] * t ** 3 / 3 + q[3] * t ** 4 / 4
    y, n3, rc = oracle.ode_test(1, q, [0.0], 0.0, 10.0, 0.01, 1e-12, 1e-12)
    steps = 0
    h, t = 0.01, 0.0
    while t < 10.0:
        h = min(h, 10.0 - t)
        t += h
        h *= 5.0
        steps += 1
    assert y[0] == pytest.approx(G(10.0), rel=1e-13) and n3 == 6 * steps
    # the quartic term makes the estimate nonzero: more steps at the same tolerance
    y, n4, rc = oracle.ode_test(1, [1.0, -2.0, 0.5, 3.0, -1.0], [0.0], 0.0, 10.0, 0.01, 1e-12, 1e-12)
    assert n4 > n3 and y[0] == pytest.approx(F(10.0), rel=1e-13)


def test_local_error_is_sixth_order():
    """One step (a tolerance no step can fail) of y' = -y: local error ~ C h^6."""
    errs = []
    for h in (0.4, 0.2, 0.1):
        y, n, rc = oracle.ode_test(0, [-1.0], [1.0], 0.0, h, h, 1e30, 1e30)
        assert n == 6
        errs.append(abs(y[0] - math.exp(-h)))
    assert 50 < errs[0] / errs[1] < 80 and 55 < errs[1] / errs[2] < 75


def test_closed_forms_and_tolerance():
    y, n, rc = oracle.ode_test(0, [-0.7], [2.0], 0.0, 10.0, 0.25, 1e-12, 1e-12)
    assert y[0] == pytest.approx(2.0 * math.exp(-7.0), rel=1e-9)
    y, n, rc = oracle.ode_test(2, [4.0], [1.0, 0.0], 0.0, 5.0, 0.25, 1e-11, 1e-11)
    assert y[0] == pytest.approx(math.cos(10.0), abs=1e-8) and y[1] == pytest.approx(-2.0 * math.sin(10.0), abs=2e-8)
    errs, evals = [], []
    for tol in (1e-3, 1e-6, 1e-9):
        y, n, rc = oracle.ode_test(2, [4.0], [1.0, 0.0], 0.0, 5.0, 0.25, tol, tol)
        errs.append(abs(y[0] - math.cos(10.0)))
        evals.append(n)
    assert errs[0] > errs[1] > errs[2] and evals[0] < evals[1] < evals[2]
    assert errs[1] < 1e-4 and errs[2] < 1e-7


def sir_rhs(pv, m):
    lay = m.layout
    beta, gamma = pv[lay["proc"]["beta"]], pv[lay["proc"]["gamma"]]
    N = pv[lay["fixed"]["N"]]

    def f(t, y):
        S, I, inc = y
        inf = beta * I / N * S
        return [-inf, inf - gamma * I, inf]
    return f


def noise_rhs(pv, m):
    """dS/dt = mu_b N - r0/N v I S - mu_d S, dI/dt = r0/N v I S - v I - mu_d I per city;
    inc_out = sum over cities of v I + mu_d I, inc_in = infections in city 2."""
    lay, fo = m.layout, m.layout["fix_off"]
    r0, v = pv[lay["proc"]["r0"]], pv[lay["proc"]["v"]]
    c = lambda name, cac: pv[lay["fixed"][name] + cac]

    def f(t, y):
        S, I = y[0:2], y[2:4]
        out = np.zeros(6)
        for k in (0, 1):
            N, mb, md = c("N", k), c("mu_b", k), c("mu_d", k)
            inf = r0 / N * v * I[k] * S[k]
            out[k] = mb * N - inf - md * S[k]
            out[2 + k] = inf - v * I[k] - md * I[k]
            out[4] += v * I[k] + md * I[k]
            if k == 1:
                out[5] = inf
        return out
    return f


@pytest.mark.parametrize("name,rhs", [("sir", sir_rhs), ("noise", noise_rhs)])
def test_model_path_against_scipy(name, rhs):
    m = Model.get(name)
    N = 30
    times, _ = m.dataset(N)
    pv, x0 = m.pack(configs.TRUTH[name]), m.x0(configs.X0_PROPS[name])
    f = rhs(pv, m)
    ref = [x0.copy()]
    y = x0.copy()
    for n in range(N):
        y[len(y) - m.desc.N_TS_INC_UNIQUE:] = 0.0      # reset the incidences
        sol = solve_ivp(f, (times[n], times[n + 1]), y, method="DOP853", rtol=1e-13, atol=1e-10)
        y = sol.y[:, -1].copy()
        ref.append(y.copy())
    ref = np.array(ref)
    tight = oracle.trajectory(m.id, times, pv, x0, ode=True, eps_abs=1e-10, eps_rel=1e-11)
    scale = np.abs(ref).max(axis=0) + 1.0
    assert np.all(np.abs(tight[1:] - ref[1:]) <= 2e-8 * scale)
    dflt = oracle.trajectory(m.id, times, pv, x0, ode=True)
    assert np.all(np.abs(dflt[1:] - ref[1:]) <= 1e-3 * scale)
    # the tolerances reach the integrator (the default ones are the reference's 1e-6, 1e-3)
    assert not np.array_equal(tight, dflt)
    np.testing.assert_array_equal(dflt, oracle.trajectory(m.id, times, pv, x0, ode=True, eps_abs=1e-6, eps_rel=1e-3))


def test_filter_is_the_one_path_likelihood():
    m = Model.get("sirs4")
    times, data = m.dataset(40)
    pv, x0 = m.pack(configs.TRUTH["sirs4"]), m.x0(configs.X0_PROPS["sirs4"])
    o = oracle.smc(m.id, 300, times, data, pv, x0, ode=True, hat=True, nthreads=4)
    path = oracle.trajectory(m.id, times, pv, x0, ode=True)
    np.testing.assert_array_equal(o["ess_n"], 300.0)
    for n in range(40):
        obs, ll, nf = oracle.obs_loglik(m.id, path[n + 1], data[n], pv, times[n + 1])
        assert o["loglik_n"][n] == ll
        np.testing.assert_array_equal(o["hat"][n + 1, :, 0], o["hat"][n + 1, :, 2])   # no spread


def spline_fns(cov):
    """scipy's version of each covariate: natural cubic spline (4 knots or more), linear (2 or
    3), constant (1)"""
    out = []
    for x, y in cov:
        if len(x) >= 4:
            out.append(CubicSpline(x, y, bc_type="natural"))
        elif len(x) > 1:
            out.append(lambda t, x=x, y=y: np.interp(t, x, y))
        else:
            out.append(lambda t, y=y: y[0])
    return out


def noise_rhs_cov(pv, m, cov):
    """noise_rhs with N, mu_b and mu_d of each city read from their splines at t"""
    lay, fo = m.layout, m.layout["fix_off"]
    r0, v = pv[lay["proc"]["r0"]], pv[lay["proc"]["v"]]
    sp = spline_fns(cov)
    c = lambda name, k, t: float(sp[lay["fixed"][name] + k - fo](t))

    def f(t, y):
        S, I = y[0:2], y[2:4]
        out = np.zeros(6)
        for k in (0, 1):
            N, mb, md = c("N", k, t), c("mu_b", k, t), c("mu_d", k, t)
            inf = r0 / N * v * I[k] * S[k]
            out[k] = mb * N - inf - md * S[k]
            out[2 + k] = inf - v * I[k] - md * I[k]
            out[4] += v * I[k] + md * I[k]
            if k == 1:
                out[5] = inf
        return out
    return f


def test_covariates_are_splines_of_the_integrators_time():
    m = Model.get("noise")
    N = 60
    times, _ = m.dataset(N)
    pv, x0 = m.pack(configs.TRUTH["noise"]), m.x0(configs.X0_PROPS["noise"])
    cov = m.covariates()
    assert any(len(x) >= 4 and np.ptp(y) > 0 for x, y in cov)
    f = noise_rhs_cov(pv, m, cov)
    ref = [x0.copy()]
    y = x0.copy()
    for n in range(N):
        y[len(y) - m.desc.N_TS_INC_UNIQUE:] = 0.0
        sol = solve_ivp(f, (times[n], times[n + 1]), y, method="DOP853", rtol=1e-13, atol=1e-10)
        y = sol.y[:, -1].copy()
        ref.append(y.copy())
    ref = np.array(ref)
    tight = oracle.trajectory(m.id, times, pv, x0, ode=True, eps_abs=1e-10, eps_rel=1e-11, cov=cov)
    scale = np.abs(ref).max(axis=0) + 1.0
    assert np.all(np.abs(tight[1:] - ref[1:]) <= 2e-8 * scale), np.abs(tight[1:] - ref[1:]).max(axis=0) / scale
    frozen = oracle.trajectory(m.id, times, pv, x0, ode=True, eps_abs=1e-10, eps_rel=1e-11)
    assert np.abs(frozen[1:] - ref[1:]).max() > 1e3 * 2e-8 * scale.max()


def test_filter_with_covariates_is_the_one_path_likelihood():
    m = Model.get("noise")
    times, data = m.dataset(40)
    pv, x0 = m.pack(configs.TRUTH["noise"]), m.x0(configs.X0_PROPS["noise"])
    cov = m.covariates()
    o = oracle.smc(m.id, 200, times, data, pv, x0, ode=True, nthreads=4, cov=cov)
    path = oracle.trajectory(m.id, times, pv, x0, ode=True, cov=cov)
    tab, kmax = oracle.cov_table(m.id, times, 0.25, cov)
    np.testing.assert_array_equal(o["ess_n"], 200.0)
    for n in range(40):
        cvd = tab[40 * kmax + n + 1]
        pvn = pv.copy()
        pvn[m.fix_off:] = cvd
        obs, ll, nf = oracle.obs_loglik(m.id, path[n + 1], data[n], pvn, times[n + 1])
        assert o["loglik_n"][n] == ll
