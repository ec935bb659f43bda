"""Pin the CPU restatement of one iterated-filtering pass (oracle.mif_pass,
C/mif/mif.c:128-200) without the GPU:

* the natural transforms (plom_mif_natural) against 50-digit mpmath;
* with equal parameters in every particle and no random walk, the pass is the plain
  filter of those parameters, bit for bit;
* the moments of theta per observation against numpy (weights exp(l - max), variance
  with divisor J - 1);
* the random walk: theta of slot j after an interval with data is its ancestor's plus
  (mut_fac * sd) * z, z the Box-Muller values (mpmath) of Philox (j, n, 4, o // 2); no move
  after an interval without data.
"""
import math

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle
from iip_amd import configs
from iip_amd.smc import Model

mp.mp.dps = 50

SIR_MAP = {  # theta: S, I (initial proportions), beta, gamma (duration), rep, phi
    "pv_src": [2, 3, 4, 5], "tr": [3, 3, 1, 1, 2, 1], "scale": [0] * 6, "duration": [0, 0, 0, 1, 0, 0],
    "min": [0.99, 0.0, 0, 0, 0, 0], "max": [1.0, 0.01, 0, 0, 0, 0], "multiplier": [1.0] * 6,
    "sd": [0.0, 0.0, 0.02, 0.02, 0.02, 0.02]}


def logit(p):
    return math.log(p / (1 - p))


def sir_theta():
    t = configs.TRUTH["sir"]
    return np.array([math.log((0.999 - 0.99) / (1.0 - 0.999)), math.log(0.001 / (0.01 - 0.001)),
                     math.log(t["beta"]), math.log(1.0 / t["gamma"]), logit(t["rep"]), math.log(t["phi"])])


def natural_pv(m, theta, mp_map=SIR_MAP):
    pv = m.pack(configs.TRUTH[m.name])
    for s, o in enumerate(mp_map["pv_src"]):
        if o >= 0:
            pv[s] = oracle.mif_natural(mp_map["tr"][o], mp_map["scale"][o], mp_map["duration"][o], theta[o],
                                       mp_map["min"][o], mp_map["max"][o], mp_map["multiplier"][o])
    return pv


@pytest.mark.parametrize("tr,sc,dur", [(0, 0, 0), (1, 0, 0), (2, 0, 0), (3, 0, 0), (1, 0, 1), (0, 1, 0), (1, 2, 0),
                                       (3, 1, 0)])
def test_natural_transform_against_mpmath(tr, sc, dur):
    rng = np.random.default_rng(tr * 10 + sc + dur)
    a, b = 0.2, 7.5
    for x in list(rng.uniform(-6, 6, 40)) + [0.0, -0.0, 1e-3, -1e-3, 30.0, -30.0]:
        X = mp.mpf(float(x))
        if tr == 1:
            y = mp.exp(X)
        elif tr == 2:
            y = 1 / (1 + mp.exp(-X))
        elif tr == 3:
            y = (b * mp.exp(X) + a) / (1 + mp.exp(X))
        else:
            y = X
        e = y
        if sc == 1:
            y = mp.power(10, y)
        elif sc == 2:
            y = mp.power(10, -y)
        y = y * 3
        if dur:
            y = 1 / (mp.mpf("0.00001") + y)
        got = oracle.mif_natural(tr, sc, dur, float(x), a, b, 3.0)
        # 10^y is exp(y ln 10): a relative error of y (up to 3 eps from the inverse transform)
        # and the rounding of y ln 10 become a relative error ~ |y ln 10| * 4 eps of 10^y
        tol = 2e-15 if sc == 0 else 2.3e-16 * (4.0 + abs(float(e)) * 2.31 * 4.0)
        assert abs(got - float(y)) <= tol * abs(float(y)) + 1e-300, (x, got, float(y))


def test_pass_with_equal_parameters_is_the_filter():
    m = Model.get("sir")
    times, data = m.dataset(30)
    J = 700
    th = sir_theta()
    theta0 = np.repeat(th[:, None], J, axis=1)
    pv = natural_pv(m, th)
    x0 = m.x0(configs.X0_PROPS["sir"])
    mp0 = dict(SIR_MAP, sd=[0.0] * 6)
    r = oracle.mif_pass(m.id, J, times, data, pv, mp0, theta0, np.repeat(x0[:, None], J, axis=1), np.full(30, 0.7),
                        seed=5, stream=3, nthreads=8)
    o = oracle.smc(m.id, J, times, data, pv, x0, seed=5, stream=3, nthreads=8)
    assert r["loglik"] == o["loglik"]
    np.testing.assert_array_equal(r["loglik_n"], o["loglik_n"])
    np.testing.assert_array_equal(r["ess_n"], o["ess_n"])
    np.testing.assert_allclose(r["mif_mean"][1:], np.repeat(th[None, :], 30, axis=0), rtol=1e-14)
    assert np.all(np.abs(r["mif_var"][1:]) < 1e-25)
    assert np.isnan(r["mif_mean"][0]).all()


def random_start(m, J, seed):
    rng = np.random.default_rng(seed)
    th = sir_theta()
    theta0 = th[:, None] + rng.normal(0.0, 0.05, (6, J)) * np.array([0, 0, 1, 1, 1, 1])[:, None]
    x0 = np.repeat(m.x0(configs.X0_PROPS["sir"])[:, None], J, axis=1)
    return theta0, x0


def test_moments_against_numpy():
    m = Model.get("sir")
    times, data = m.dataset(12)
    data = data.copy()
    data[4] = np.nan
    J = 600
    theta0, x0 = random_start(m, J, 1)
    pv = natural_pv(m, sir_theta())
    r = oracle.mif_pass(m.id, J, times, data, pv, SIR_MAP, theta0, x0, np.full(12, 0.9), seed=2, stream=7,
                        theta_out=True, logw=True)
    for n in range(12):
        if n == 4:
            assert np.isnan(r["mif_mean"][n + 1]).all() and np.isnan(r["mif_var"][n + 1]).all()
            continue
        lw = r["logw"][n]
        w = np.where(np.isfinite(lw), np.exp(lw - lw.max()), 0.0)
        th = r["theta"][n + 1]
        np.testing.assert_allclose(r["mif_mean"][n + 1], (th * w).sum(axis=1) / w.sum(), rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(r["mif_var"][n + 1], th.var(axis=1, ddof=1), rtol=1e-9, atol=1e-30)


def test_random_walk_and_resampling_of_theta():
    m = Model.get("sir")
    times, data = m.dataset(8)
    data = data.copy()
    data[2] = np.nan
    J = 300
    theta0, x0 = random_start(m, J, 4)
    fac = np.linspace(0.5, 1.2, 8)
    pv = natural_pv(m, sir_theta())
    r = oracle.mif_pass(m.id, J, times, data, pv, SIR_MAP, theta0, x0, fac, seed=9, stream=1, theta_out=True,
                        ancestry=True)
    th, anc = r["theta"], r["ancestry"]
    np.testing.assert_array_equal(th[1], theta0)          # interval 0: no move
    sd = np.array(SIR_MAP["sd"])
    for n in range(1, 8):
        prev = th[n][:, anc[n - 1]]                       # the ancestors' theta
        if n - 1 == 2:                                    # no data before interval n: no move
            np.testing.assert_array_equal(th[n + 1], prev)
            continue
        for j in range(0, J, 37):
            for o in range(6):
                if sd[o] == 0.0:
                    assert th[n + 1][o, j] == prev[o, j]
                    continue
                w = oracle.philox([j, n, 4, o // 2], [9, 1])
                u1 = (mp.mpf(int(w[0])) + mp.mpf(1) / 2) / 2 ** 32
                u2 = (mp.mpf(int(w[1])) + mp.mpf(1) / 2) / 2 ** 32
                rr = mp.sqrt(-2 * mp.log(u1))
                z = float(rr * (mp.cos(2 * mp.pi * u2) if o % 2 == 0 else mp.sin(2 * mp.pi * u2)))
                assert th[n + 1][o, j] == pytest.approx(prev[o, j] + (fac[n - 1] * sd[o]) * z, rel=0, abs=1e-14)


def test_correlated_walk_with_a_covariance_factor():
    """With a lower-triangular factor L (a "covariance" in theta, methods.c:275-330): after an
    interval with data, theta_o = ancestor's + mut_fac sum_{q<=o} L[o][q] z_q, the z from the
    same Philox counters as the diagonal walk (mpmath Box-Muller here); rows of zeros (the
    initial conditions S, I here) do not move."""
    m = Model.get("sir")
    times, data = m.dataset(6)
    J = 200
    theta0, x0 = random_start(m, J, 7)
    rng = np.random.default_rng(3)
    L = np.zeros((6, 6))
    for o in range(2, 6):
        L[o, 2:o + 1] = rng.normal(0.0, 0.02, o - 1)
        L[o, o] = abs(L[o, o]) + 0.01
    mp_map = dict(SIR_MAP, chol=L)
    fac = np.linspace(0.6, 1.1, 6)
    pv = natural_pv(m, sir_theta())
    r = oracle.mif_pass(m.id, J, times, data, pv, mp_map, theta0, x0, fac, seed=4, stream=2, theta_out=True,
                        ancestry=True)
    th, anc = r["theta"], r["ancestry"]
    for n in range(1, 6):
        prev = th[n][:, anc[n - 1]]
        for j in range(0, J, 23):
            z = []
            for b in range(3):
                w = oracle.philox([j, n, 4, b], [4, 2])
                u1 = (mp.mpf(int(w[0])) + mp.mpf(1) / 2) / 2 ** 32
                u2 = (mp.mpf(int(w[1])) + mp.mpf(1) / 2) / 2 ** 32
                rr = mp.sqrt(-2 * mp.log(u1))
                z += [float(rr * mp.cos(2 * mp.pi * u2)), float(rr * mp.sin(2 * mp.pi * u2))]
            for o in range(6):
                want = prev[o, j] + fac[n - 1] * sum(L[o, q] * z[q] for q in range(o + 1))
                if o < 2:
                    assert th[n + 1][o, j] == prev[o, j]
                else:
                    assert th[n + 1][o, j] == pytest.approx(want, rel=0, abs=1e-14)
