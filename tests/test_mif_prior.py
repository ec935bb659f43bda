"""The prior patch of iterated filtering (mif --prior: patch_likelihood_prior,
C/mif/methods.c:469-523), pinned on the CPU:

* the densities (plom_mif_prior_pdf) against scipy: the uniform density of
  gsl_ran_flat_pdf (1/(b-a) on [a, b), 0 elsewhere), the normal prior with mean (a+b)/2 and sd
  (b-a)/4; the pseudo-uniform density through the properties that fix it: zero at a and b,
  flat on [a+d, b-d] (d = (b-a)/100), quadratic and once differentiable at the joins,
  symmetric, integrating to 1;
* the term e log(max(p, LIKE_MIN)), NaN / zero densities counting as LIKE_MIN
  (sanitize_likelihood, C/core/likelihood.c:77-91);
* the pass: with every noise and the random walk off, particle j's path is the deterministic
  path of its root ancestor's theta and initial state, so its weight at every interval
  follows from oracle.trajectory and oracle.obs_loglik; the oracle's log-weights equal
  log(exp(l) prod_o p_o^e_o) of the reference's linear weights (mpmath), and a particle
  fails exactly when that weight is at or below LIKE_MIN^N_TS (weight(), C/core/smc.c:44),
  including the rows where every particle's observation is floored; the initial conditions'
  terms stop at interval L;
* exponents 0 leave the pass bit-identical to the pass without a prior.
"""
import math

import mpmath as mp
import numpy as np
import pytest
from scipy import integrate, stats

from oracle import oracle
from iip_amd import configs
from iip_amd.smc import Model

mp.mp.dps = 40
LIKE_MIN = 1e-17


@pytest.mark.parametrize("a,b", [(0.0, 1.0), (-3.0, 7.5), (0.2, 0.25), (1e3, 5e3)])
def test_densities_against_scipy(a, b):
    rng = np.random.default_rng(int(b * 10))
    xs = np.concatenate([rng.uniform(a - 0.2 * (b - a), b + 0.2 * (b - a), 300), [a, b]])
    for x in xs:
        f = oracle.mif_prior_pdf(oracle.PRIOR_UNIFORM, x, a, b)
        assert f == (1.0 / (b - a) if a <= x < b else 0.0)
        g = oracle.mif_prior_pdf(oracle.PRIOR_NORMAL, x, a, b)
        assert g == pytest.approx(stats.norm.pdf(x, 0.5 * (a + b), (b - a) / 4.0), rel=1e-14, abs=1e-300)
    p = lambda x: oracle.mif_prior_pdf(oracle.PRIOR_PSEUDO_UNIFORM, x, a, b)
    d = 0.01 * (b - a)
    h = p(0.5 * (a + b))
    assert p(a) == 0.0 and p(b) == 0.0 and p(a - d) == 0.0 and p(b + d) == 0.0
    assert p(a + d) == h and p(b - 1.5 * d) == h and p(a + 1.5 * d) == h
    # the shoulders: continuous with the flat part, zero slope at the joins, zero at the ends
    for e in (1e-4, 1e-6):
        assert p(a + d - e * d) == pytest.approx(h, rel=3 * e * e + 1e-12)
        assert p(b - d + e * d) == pytest.approx(h, rel=3 * e * e + 1e-12)
        assert p(a + e * d) == pytest.approx(h * (2 * e - e * e), rel=1e-6)
    # quadratic on a shoulder: third differences vanish
    u = a + d * np.array([0.1, 0.3, 0.5, 0.7])
    v = [p(x) for x in u]
    assert abs(v[3] - 3 * v[2] + 3 * v[1] - v[0]) <= 1e-9 * h
    for t in rng.uniform(0, 1, 50):
        assert p(a + t * (b - a)) == pytest.approx(p(b - t * (b - a)), rel=1e-9, abs=1e-12 * h)
    total = sum(integrate.quad(p, lo, hi, epsabs=0, epsrel=1e-13)[0]
                for lo, hi in ((a, a + d), (a + d, b - d), (b - d, b)))
    assert total == pytest.approx(1.0, rel=1e-11)


def test_prior_term_floors_and_transforms():
    lm = math.log(LIKE_MIN)
    # outside the support, NaN: LIKE_MIN
    assert oracle.mif_prior_term(oracle.PRIOR_UNIFORM, 0.5, 0, 2.0, 0.0, 1.0) == 0.5 * lm
    assert oracle.mif_prior_term(oracle.PRIOR_NORMAL, 1.0, 0, float("nan"), 0.0, 1.0) == lm
    assert oracle.mif_prior_term(oracle.PRIOR_NORMAL, 1.0, 0, 40.0, 0.0, 1.0) == lm      # density ~ 1e-2775
    # the density is taken at f_inv(theta): log, logit, logit_ab
    for tr, th, a, b, x in ((1, math.log(0.3), 0.0, 1.0, 0.3), (2, 0.4, 0.0, 2.0, 1 / (1 + math.exp(-0.4))),
                            (3, -0.7, 2.0, 9.0, (9.0 * math.exp(-0.7) + 2.0) / (1 + math.exp(-0.7)))):
        got = oracle.mif_prior_term(oracle.PRIOR_NORMAL, 0.25, tr, th, a, b)
        assert got == pytest.approx(0.25 * stats.norm.logpdf(x, 0.5 * (a + b), (b - a) / 4.0), rel=1e-13)
    assert oracle.mif_prior_term(oracle.PRIOR_UNIFORM, 1 / 3, 1, math.log(0.5), 0.0, 4.0) == \
        pytest.approx(-math.log(4.0) / 3, rel=1e-15)


# theta of the SIR model: S, I (initial proportions, logit_ab), beta, gamma (a duration), rep, phi
SIR_MAP = {"pv_src": [2, 3, 4, 5], "tr": [3, 3, 1, 1, 2, 1], "scale": [0] * 6, "duration": [0, 0, 0, 1, 0, 0],
           "min": [0.99, 0.0, 0.2, 4.85, -0.4, 0.0], "max": [1.0, 0.01, 0.8, 5.6, 0.8, 1.0],
           "multiplier": [1.0] * 6, "sd": [0.0] * 6}


def prior_setup(N, J, L, seed=3):
    """A prior on every component but phi: pseudo-uniform and uniform on the initial
    proportions (exponent 1/L, intervals n < L), normal on beta (1/N), uniform on the
    duration 1/gamma with exponent 1 (particles outside [4.85, 5.6) fail) and a wide normal
    on rep with exponent 1 whose log-density changes sign across the particles; the
    particles' own theta and initial states; row 6 far from the model (every particle's
    observation floored) and row 9 missing."""
    m = Model.get("sir")
    times, data = m.dataset(N)
    data = data.copy()
    data[6] = 1e9
    data[9] = np.nan
    t = configs.TRUTH["sir"]
    rng = np.random.default_rng(seed)
    th = np.array([math.log((0.999 - 0.99) / (1.0 - 0.999)), math.log(0.001 / (0.01 - 0.001)), math.log(t["beta"]),
                   math.log(1.0 / t["gamma"]), math.log(t["rep"] / (1 - t["rep"])), math.log(t["phi"])])
    theta0 = th[:, None] + rng.normal(0.0, 1.0, (6, J)) * np.array([0.3, 0.3, 0.05, 0.06, 0.8, 0.05])[:, None]
    x0 = np.repeat(m.x0(configs.X0_PROPS["sir"])[:, None], J, axis=1) * (1.0 + rng.normal(0.0, 0.01, (1, J)))
    mp_ = dict(SIR_MAP, prior=[oracle.PRIOR_PSEUDO_UNIFORM, oracle.PRIOR_UNIFORM, oracle.PRIOR_NORMAL,
                               oracle.PRIOR_UNIFORM, oracle.PRIOR_NORMAL, oracle.PRIOR_NONE],
               prior_pow=[1.0 / L, 1.0 / L, 1.0 / N, 1.0, 1.0, 0.0], prior_until=[L, L, N, N, N, N])
    pv = m.pack(t)
    return m, times, data, pv, mp_, theta0, x0


def natural_pv(m, pv, mp_, th):
    pv = pv.copy()
    for s, o in enumerate(mp_["pv_src"]):
        if o >= 0:
            pv[s] = oracle.mif_natural(mp_["tr"][o], mp_["scale"][o], mp_["duration"][o], th[o], mp_["min"][o],
                                       mp_["max"][o], mp_["multiplier"][o])
    return pv


def reference_weight(l, th, mp_, n):
    """exp(l) prod_o prior_o(f_inv(theta_o))^e_o in 40 digits (the reference's linear weight);
    the densities are sanitized at LIKE_MIN as in sanitize_likelihood."""
    w = mp.exp(mp.mpf(l))
    for o in range(6):
        k = mp_["prior"][o]
        if k < 0 or n >= mp_["prior_until"][o]:
            continue
        a, b = mp.mpf(mp_["min"][o]), mp.mpf(mp_["max"][o])
        x, tr = mp.mpf(th[o]), mp_["tr"][o]
        x = mp.exp(x) if tr == 1 else (1 / (1 + mp.exp(-x)) if tr == 2 else ((b * mp.exp(x) + a) / (1 + mp.exp(x))
                                                                            if tr == 3 else x))
        if k == oracle.PRIOR_UNIFORM:
            p = 1 / (b - a) if a <= x < b else mp.mpf(0)
        elif k == oracle.PRIOR_NORMAL:
            p = mp.npdf(x, (a + b) / 2, (b - a) / 4)
        else:
            d = (b - a) / 100
            hh = 1 / (b - a - 2 * d / 3)
            p = mp.mpf(0) if (x <= a or x >= b) else (hh if a + d <= x < b - d else
                                                     hh - hh / d ** 2 * (x - (a + d if x < a + d else b - d)) ** 2)
        p = max(p, mp.mpf(LIKE_MIN))
        w *= p ** mp.mpf(mp_["prior_pow"][o])
    return w


def test_patched_weights_of_deterministic_paths():
    N, J, L = 12, 300, 4
    m, times, data, pv, mp_, theta0, x0 = prior_setup(N, J, L)
    r = oracle.mif_pass(m.id, J, times, data, pv, mp_, theta0, x0, np.full(N, 0.8), noises_off=3, seed=4, stream=1,
                        logw=True, ancestry=True, nthreads=8, like_min=LIKE_MIN)
    anc = r["ancestry"]
    paths = {}
    thr = mp.mpf(LIKE_MIN) ** m.n_ts
    seen = {"fail": 0, "floored_kept": 0, "floored_failed": 0, "kept": 0}
    for n in range(N):
        if n == 9:
            continue
        for j in range(J):
            root = j
            for k in range(n - 1, -1, -1):
                root = anc[k][root]
            th = theta0[:, root]
            pvj = natural_pv(m, pv, mp_, th)
            if root not in paths:
                paths[root] = oracle.trajectory(m.id, times, pvj, x0[:, root], noises_off=3)
            obs, l, nf = oracle.obs_loglik(m.id, paths[root][n + 1], data[n], pvj, times[n + 1], LIKE_MIN)
            w = reference_weight(l, th, mp_, n)
            got = r["logw"][n][j]
            if w <= thr:
                assert got == -math.inf, (n, j, float(mp.log(w)))
                seen["floored_failed" if nf == m.n_ts else "fail"] += 1
            else:
                assert got == pytest.approx(float(mp.log(w)), rel=1e-13, abs=1e-12), (n, j)
                seen["floored_kept" if nf == m.n_ts else "kept"] += 1
    # every branch of the failure test was taken
    assert min(seen.values()) > 0, seen
    # the initial conditions' terms stop at L: their exponents do not reach intervals >= L
    assert r["n_all_fail"] == 0


def test_zero_exponents_are_the_pass_without_prior():
    N, J, L = 14, 500, 5
    m, times, data, pv, mp_, theta0, x0 = prior_setup(N, J, L, seed=8)
    data = m.dataset(N)[1]
    mp_["prior_pow"] = [0.0] * 6
    mut = np.linspace(0.3, 0.9, N)
    walk = dict(mp_, sd=[0.0, 0.0, 0.02, 0.02, 0.02, 0.02])
    plain = {k: v for k, v in walk.items() if not k.startswith("prior")}
    a = oracle.mif_pass(m.id, J, times, data, pv, walk, theta0, x0, mut, seed=2, stream=5, nthreads=8)
    b = oracle.mif_pass(m.id, J, times, data, pv, plain, theta0, x0, mut, seed=2, stream=5, nthreads=8)
    assert a["loglik"] == b["loglik"]
    np.testing.assert_array_equal(a["loglik_n"], b["loglik_n"])
    np.testing.assert_array_equal(a["mif_mean"], b["mif_mean"])
    # with the exponents on, the weights move
    c = oracle.mif_pass(m.id, J, times, data, pv, dict(walk, prior_pow=[0.2] * 5 + [0.0]), theta0, x0, mut, seed=2,
                        stream=5, nthreads=8)
    assert c["loglik"] != a["loglik"]
