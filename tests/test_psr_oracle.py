"""Pin the psr implementation (Euler-multinomial step, step_psr,
C/templates/core_template.c:104-209) of the CPU oracle without the GPU:

* exact identities of one SIR path: states stay integers, S never grows, and the weekly
  incidence is the drop of S (the only exit of S is infection);
* the law of large numbers: the mean of 2000 paths (1e6 hosts) follows the deterministic
  recursion of the step's expected flows, S <- S exp(-beta I/N dt), I <- I exp(-gamma dt) +
  S (1 - exp(-beta I/N dt)), within 0.3 % (and a few hosts near extinction);
* white noise: with --no_white_noise the SIRS model's paths differ from the noisy ones,
  and the noisy paths' mean incidence stays within a few percent (the noise has mean 1);
* the filter runs with the psr step (log-likelihood finite, the same for the same key).
"""
import math

import numpy as np
import pytest

from oracle import oracle
from iip_amd import configs
from iip_amd.smc import Model


def sir_setup(N):
    m = Model.get("sir")
    times, data = m.dataset(N)
    pv = m.pack(configs.TRUTH["sir"])
    x0 = np.round(m.x0(configs.X0_PROPS["sir"]))
    return m, times, data, pv, x0


def test_sir_path_identities():
    m, times, data, pv, x0 = sir_setup(40)
    for j in range(20):
        tr = oracle.trajectory(m.id, times, pv, x0, seed=3, j=j, psr=True)
        assert np.all(tr == np.floor(tr))
        S, I, inc = tr[:, 0], tr[:, 1], tr[:, 2]
        assert np.all(np.diff(S) <= 0) and np.all(I >= 0)
        np.testing.assert_array_equal(inc[1:], S[:-1] - S[1:])


def test_sir_mean_follows_expected_flows():
    m, times, data, pv, x0 = sir_setup(30)
    paths = np.array([oracle.trajectory(m.id, times, pv, x0, seed=8, j=j, psr=True) for j in range(2000)])
    mean = paths.mean(axis=0)
    beta, gamma, Npop = pv[0], pv[1], pv[4]
    S, I = x0[0], x0[1]
    dt = 0.25
    det = [(S, I)]
    for n in range(30):
        for k in range(28):
            pS = 1.0 - math.exp(-beta * I / Npop * dt)
            pI = 1.0 - math.exp(-gamma * dt)
            inf, rec = S * pS, I * pI
            S, I = S - inf, I + inf - rec
        det.append((S, I))
    det = np.array(det)
    # relative agreement while the epidemic is large; near extinction (a few hosts) the
    # discrete paths stop where the recursion keeps decaying: a few hosts of slack
    np.testing.assert_allclose(mean[:, 0], det[:, 0], rtol=3e-3, atol=5.0)
    np.testing.assert_allclose(mean[:, 1], det[:, 1], rtol=3e-3, atol=5.0)


def test_white_noise_switch():
    m = Model.get("sirs4")
    times, data = m.dataset(20)
    pv = m.pack(configs.TRUTH["sirs4"])
    x0 = np.round(m.x0(configs.X0_PROPS["sirs4"]))
    noisy = np.array([oracle.trajectory(m.id, times, pv, x0, seed=4, j=j, psr=True) for j in range(300)])
    quiet = np.array([oracle.trajectory(m.id, times, pv, x0, seed=4, j=j, psr=True, noises_off=2) for j in range(300)])
    assert not np.array_equal(noisy, quiet)
    k0 = m.S - m.n_ts   # the incidence slots are the last N_TS states
    rel = noisy[:, 1:, k0:].mean(axis=0).sum() / quiet[:, 1:, k0:].mean(axis=0).sum()
    assert abs(rel - 1.0) < 0.05


def test_psr_filter_runs():
    m, times, data, pv, x0 = sir_setup(20)
    a = oracle.smc(m.id, 500, times, data, pv, x0, seed=3, stream=1, psr=True, nthreads=8)
    b = oracle.smc(m.id, 500, times, data, pv, x0, seed=3, stream=1, psr=True, nthreads=8)
    c = oracle.smc(m.id, 500, times, data, pv, x0, seed=3, stream=1, nthreads=8)
    assert math.isfinite(a["loglik"]) and a["loglik"] == b["loglik"] and a["loglik"] != c["loglik"]
    assert a["n_all_fail"] == 0
