"""Pin the discrete samplers of the PSR step (csrc/plom_psr.h) on the CPU against the
distributions they draw from (scipy), with fixed keys (deterministic):

* chi-square goodness of fit of Poisson and Binomial draws in both regimes of each
  sampler (inversion below mean 10, transformed rejection above) and at the symmetry
  switch p = 1/2 of the binomial; the test rejects a 1 % shifted parameter;
* Kolmogorov-Smirnov of Gamma draws for shapes below and above 1 and of the uniforms (two
  words, and the one-word uniforms of the rejection attempts);
* log k! against mpmath; the multinomial's counts sum to N with the right means; edge
  cases (n = 0, p = 0, p = 1, mu = 0).
"""
import mpmath as mp
import numpy as np
import pytest
from scipy import stats

from oracle import oracle

mp.mp.dps = 40


def chi2_pvalue(x, pmf, lo=None):
    """chi-square p-value of integer draws x against pmf, bins with expected >= 5 (tails merged)"""
    x = np.asarray(x, dtype=np.int64)
    n = len(x)
    kmin, kmax = int(x.min()), int(x.max())
    ks = np.arange(max(0, kmin - 50), kmax + 51)
    e = n * pmf(ks)
    obs = np.bincount(x - ks[0], minlength=len(ks))[:len(ks)].astype(float)
    bins_o, bins_e, acc_o, acc_e = [], [], 0.0, 0.0
    for o, ee in zip(obs, e):
        acc_o += o
        acc_e += ee
        if acc_e >= 5.0:
            bins_o.append(acc_o)
            bins_e.append(acc_e)
            acc_o = acc_e = 0.0
    if len(bins_o) < 2:
        return 0.0      # the draws sit where the distribution puts (almost) no mass
    bins_o[-1] += acc_o
    bins_e[-1] += acc_e
    bins_e = np.array(bins_e)
    bins_e *= sum(bins_o) / bins_e.sum()
    return stats.chisquare(bins_o, bins_e, ddof=0).pvalue


@pytest.mark.parametrize("mu", [0.3, 3.0, 9.99, 10.0, 25.0, 300.0, 1e5])
def test_poisson(mu):
    x = oracle.psr_draws(oracle.PSR_POISSON, mu, 60000, seed=11, j=int(mu * 7) % 1000)
    assert np.all(x == np.floor(x)) and x.min() >= 0
    assert chi2_pvalue(x, lambda k: stats.poisson.pmf(k, mu)) > 1e-4
    if mu >= 3.0:   # power: a 1 % shift of the mean is rejected
        assert chi2_pvalue(x, lambda k: stats.poisson.pmf(k, mu * (1.01 if mu > 100 else 1.1))) < 1e-4


@pytest.mark.parametrize("n,p", [(10, 0.3), (100, 0.05), (60, 0.5), (1000, 0.3), (1000000, 1e-5), (1000000, 0.7),
                                 (500000, 0.5), (30, 0.99), (200000, 0.0002)])
def test_binomial(n, p):
    x = oracle.psr_draws(oracle.PSR_BINOMIAL, n, 60000, b=p, seed=5, j=n % 977)
    assert np.all(x == np.floor(x)) and x.min() >= 0 and x.max() <= n
    assert chi2_pvalue(x, lambda k: stats.binom.pmf(k, n, p)) > 1e-4
    if n * p * (1 - p) > 50:
        assert chi2_pvalue(x, lambda k: stats.binom.pmf(k, n, p * 1.01)) < 1e-4


@pytest.mark.parametrize("a", [0.3, 1.0, 2.5, 25.0, 400.0])
def test_gamma(a):
    x = oracle.psr_draws(oracle.PSR_GAMMA, a, 40000, seed=2, j=int(a * 10))
    assert x.min() > 0
    assert stats.kstest(x, stats.gamma(a).cdf).pvalue > 1e-4
    assert stats.kstest(x, stats.gamma(a * 1.03).cdf).pvalue < 1e-4


def test_uniform32():
    """the one-word uniforms of the rejection attempts: (w + 1/2) 2^-32, symmetric about 1/2"""
    u = oracle.psr_draws(oracle.PSR_UNIFORM32, 0.0, 100000, seed=9)
    assert u.min() > 0.0 and u.max() < 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-4
    w = u * 2.0 ** 32 - 0.5
    assert np.all(w == np.floor(w)) and w.min() >= 0 and w.max() <= 2.0 ** 32 - 1


def test_uniform():
    u = oracle.psr_draws(oracle.PSR_UNIFORM, 0.0, 100000, seed=9)
    assert u.min() > 0.0 and u.max() < 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-4
    # 53-bit resolution: the fractional parts below 2^-32 are populated
    assert len(np.unique(np.floor(u * 2.0 ** 40) % 256)) > 250


def test_log_factorial():
    ks = list(range(0, 40)) + [99, 1000, 12345, 10 ** 6, 3 * 10 ** 7]
    for k in ks:
        got = oracle.psr_draws(oracle.PSR_LFACT, k, 1)[0]
        exact = float(mp.log(mp.factorial(k))) if k < 5000 else float(mp.loggamma(k + 1))
        assert abs(got - exact) <= 1e-14 * max(1.0, abs(exact)), (k, got, exact)


def test_log_factorial_pairs():
    """log k1! + log k2! with one shared reciprocal (plom_lfact2, the binomial sampler's
    two terms) against mpmath, on both sides of the table / series switch."""
    pairs = [(0, 0), (5, 29), (29, 30), (30, 31), (40, 10 ** 6), (1000, 12345), (3 * 10 ** 7, 25),
             (10 ** 6, 10 ** 6), (123457, 999983)]
    for k1, k2 in pairs:
        got = oracle.psr_draws(oracle.PSR_LFACT2, k1, 1, b=k2)[0]
        exact = float(mp.loggamma(k1 + 1) + mp.loggamma(k2 + 1))
        assert abs(got - exact) <= 1e-14 * max(1.0, abs(exact)), (k1, k2, got, exact)


def test_multinomial_and_edges():
    p = np.array([0.1, 0.25, 0.05, 0.6])
    tot = np.zeros(4)
    for j in range(3000):
        c = oracle.psr_draws(oracle.PSR_MULTINOMIAL, 1000, 0, p=p, seed=4, j=j)
        assert c.sum() == 1000 and np.all(c >= 0)
        tot += c
    np.testing.assert_allclose(tot / 3000, 1000 * p, rtol=0.01)
    assert np.all(oracle.psr_draws(oracle.PSR_BINOMIAL, 0, 10, b=0.5) == 0)
    assert np.all(oracle.psr_draws(oracle.PSR_BINOMIAL, 77, 10, b=0.0) == 0)
    assert np.all(oracle.psr_draws(oracle.PSR_BINOMIAL, 77, 10, b=1.0) == 77)
    assert np.all(oracle.psr_draws(oracle.PSR_POISSON, 0.0, 10) == 0)
    # a zero-probability category draws nothing and the rest is shared
    c = oracle.psr_draws(oracle.PSR_MULTINOMIAL, 500, 0, p=[0.0, 0.5, 0.5], seed=1)
    assert c[0] == 0 and c.sum() == 500
