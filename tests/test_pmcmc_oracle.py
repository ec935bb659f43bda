"""Pin the CPU restatement of the pMCMC driver (oracle/pmcmc_oracle.py) against what the
mathematics fixes, with closed-form likelihoods instead of the particle filter:

* the chain's stationary distribution is the posterior prior x likelihood on the natural
  scale (sequential and full-update proposals, flat / normal / pseudo-uniform priors,
  the logit_ab transform of Bayesian methods): sample means agree with quadrature
  within Monte Carlo error, and dropping the change-of-variable factor is detected;
* a sequential cycle updates every estimated component exactly once;
* the empirical covariance is the population covariance of the accepted thetas;
* every proposal satisfies the initial-condition constraint.
"""
import math

import numpy as np
import pytest

from oracle import pmcmc_oracle as po


def logit_ab(x, a, b):
    return math.log((x - a) / (b - x))


def problem_1d(prior="flat", sd=1.0, lo=0.0, hi=1.0, guess=0.5):
    return {"theta": [logit_ab(guess, lo, hi)], "var": [sd * sd], "estimated": [0],
            "components": [{"tr": "logit_ab", "min": lo, "max": hi, "prior": prior}],
            "constrain": lambda th: 0}


def batch_se(x, nb=20):
    x = np.asarray(x)
    b = len(x) // nb
    means = x[:nb * b].reshape(nb, b).mean(axis=1)
    return means.std(ddof=1) / math.sqrt(nb)


def quad_mean(dens, lo, hi, n=200001):
    x = np.linspace(lo, hi, n)[1:-1]
    w = dens(x)
    return float(np.sum(x * w) / np.sum(w))


def gauss_ll(mu, sd):
    return lambda th, m: -0.5 * ((po.f_inv("logit_ab", th[0], 0.0, 1.0) - mu) / sd) ** 2


@pytest.mark.parametrize("prior", ["flat", "normal", "pseudo_uniform"])
def test_posterior_mean_1d(prior):
    mu, sd = 0.3, 0.15
    r = po.run(problem_1d(prior), gauss_ll(mu, sd), 12000, seed=11)
    th = np.array([row[1][0] for row in r["trace"]])[500:]
    pr = {"flat": lambda x: np.ones_like(x),
          "normal": lambda x: np.exp(-0.5 * ((x - 0.5) / 0.25) ** 2),
          "pseudo_uniform": lambda x: np.array([po.prior_pdf("pseudo_uniform", v, 0.0, 1.0) for v in x])}[prior]
    exact = quad_mean(lambda x: pr(x) * np.exp(-0.5 * ((x - mu) / sd) ** 2), 0.0, 1.0, 20001)
    se = batch_se(th)
    assert abs(th.mean() - exact) < 4 * se + 1e-3, (th.mean(), exact, se)


def test_change_of_variable_matters():
    """Without the derivative of the inverse transform the chain targets another
    distribution (flat on the logit scale), and the test above would see it."""
    mu, sd = 0.3, 0.15
    r = po.run(problem_1d(), gauss_ll(mu, sd), 12000, seed=11, jacobian=False)
    th = np.array([row[1][0] for row in r["trace"]])[500:]
    exact = quad_mean(lambda x: np.exp(-0.5 * ((x - mu) / sd) ** 2), 0.0, 1.0, 20001)
    assert abs(th.mean() - exact) > 4 * batch_se(th) + 1e-3


def test_full_update_2d_with_constraint():
    """Two proportions p, q on [0, 1] with p + q <= 1 (the initial-condition constraint),
    correlated Gaussian likelihood; MVN proposals with epsilon adaptation and the switch
    to the empirical covariance."""
    mu = np.array([0.45, 0.35])
    cov = np.array([[0.02, -0.008], [-0.008, 0.015]])
    icov = np.linalg.inv(cov)

    def nat(th):
        return np.array([po.f_inv("logit_ab", th[0], 0.0, 1.0), po.f_inv("logit_ab", th[1], 0.0, 1.0)])

    def ll(th, m):
        d = nat(th) - mu
        return -0.5 * float(d @ icov @ d)

    prob = {"theta": [logit_ab(0.3, 0, 1), logit_ab(0.3, 0, 1)], "var": [0.3, 0.0, 0.0, 0.3], "estimated": [0, 1],
            "components": [{"tr": "logit_ab", "min": 0.0, "max": 1.0, "prior": "flat"}] * 2,
            "constrain": lambda th: int(nat(th).sum() > 1.0)}
    r = po.run(prob, ll, 15000, seed=5, full=True, m_eps=50)
    assert all(nat(p).sum() <= 1.0 for p in r["proposals"])
    th = np.array([row[1] for row in r["trace"]])[1000:]
    g = np.linspace(0, 1, 801)[1:-1]
    P, Q = np.meshgrid(g, g, indexing="ij")
    D = np.stack([P - mu[0], Q - mu[1]], -1)
    w = np.exp(-0.5 * np.einsum("...i,ij,...j->...", D, icov, D)) * (P + Q <= 1.0)
    exact = np.array([np.sum(P * w), np.sum(Q * w)]) / np.sum(w)
    for k in range(2):
        assert abs(th[:, k].mean() - exact[k]) < 4 * batch_se(th[:, k]) + 2e-3, (k, th[:, k].mean(), exact[k])
    assert 0.05 < r["global_ar"] < 0.8


def test_sequential_cycles_update_each_component_once():
    prob = {"theta": [0.0, 0.1, -0.2, 0.3], "var": list(np.diag([0.1, 0.0, 0.2, 0.3]).ravel()), "estimated": [0, 2, 3],
            "components": [{"tr": "logit_ab", "min": 0.0, "max": 1.0, "prior": "flat"}] * 4,
            "constrain": lambda th: 0}
    r = po.run(prob, lambda th, m: 0.0, 1 + 3 * 40, seed=3)
    up = r["updated"]
    assert len(up) == 3 * 40
    orders = [tuple(up[3 * c:3 * c + 3]) for c in range(40)]
    assert all(sorted(o) == [0, 2, 3] for o in orders)
    assert len(set(orders)) > 1          # the order is shuffled
    # the component with zero jump size never moves
    assert all(p[1] == 0.1 for p in r["proposals"])


def test_empirical_covariance_is_population_covariance():
    prob = {"theta": [0.0, 0.0], "var": [0.5, 0.1, 0.1, 0.4], "estimated": [0, 1],
            "components": [{"tr": "logit_ab", "min": 0.0, "max": 1.0, "prior": "flat"}] * 2,
            "constrain": lambda th: 0}
    r = po.run(prob, lambda th, m: -0.5 * (th[0] ** 2 + (th[1] - 0.5) ** 2), 400, seed=9, full=True)
    means = np.array(r["means"][1:])
    np.testing.assert_allclose(r["var_sampling"], np.cov(means.T, bias=True), rtol=1e-9, atol=1e-14)
