"""mif --prior on the GPU (plom_mif_map.prior*: the prior patch of the weights,
patch_likelihood_prior, C/mif/methods.c:469-523) against the CPU oracle, which is pinned in
test_mif_prior.py: bit-identical FP64 log-likelihoods, ESS and theta moments, with the
random walk on and off, rows where every particle's observation is floored, the ode
implementation, batched with passes without a prior; the refusals."""
import numpy as np
import pytest

from oracle import oracle
from iip_amd import smc
from test_gpu_mif import check_equal, mif_setup
from test_mif_prior import LIKE_MIN, prior_setup

pytestmark = pytest.mark.gpu


def run_pass(ctx, m, J, times, data, pv, mp_, theta0, x0, mut, noises_off=0, flags=0, seed=6, stream=2):
    f = smc.Filter(ctx, m, J, times, data, pv, x0[:, 0], noises_off=noises_off, flags=smc.FLAG_MIF | flags, mif=mp_,
                   like_min=LIKE_MIN)
    f.mif_start(theta0, x0, mut, seed=seed, stream=stream)
    f.run()
    g = f.read(mif=True)
    f.close()
    return g


@pytest.mark.parametrize("J,noises_off,walk", [(300, 3, False), (2000, 0, True), (257, 0, False), (1, 0, True)])
def test_prior_pass_equals_oracle(ctx, J, noises_off, walk):
    N = 12
    m, times, data, pv, mp_, theta0, x0 = prior_setup(N, J, 4)
    if walk:
        mp_ = dict(mp_, sd=[0.0, 0.0, 0.03, 0.02, 0.05, 0.02])
    mut = np.linspace(0.5, 1.0, N)
    print("mif prior", J, noises_off, walk, flush=True)
    g = run_pass(ctx, m, J, times, data, pv, mp_, theta0, x0, mut, noises_off=noises_off)
    o = oracle.mif_pass(m.id, J, times, data, pv, mp_, theta0, x0, mut, noises_off=noises_off, seed=6, stream=2,
                        nthreads=8, like_min=LIKE_MIN)
    check_equal(g, o)


def sirs4_prior(N, J, L):
    """Every component of the SIRS4 map gets a prior (kinds in turn, bounds around its
    natural starting value), exponents 1/N, the first two components until L; row 3 far
    from the model."""
    m, times, data, pv, mp_, theta0, x0, mut = mif_setup("sirs4", N, J, 5)
    data = data.copy()
    data[3] = 1e9
    nth = len(mp_["tr"])
    mn, mx = list(mp_["min"]), list(mp_["max"])
    for o in range(nth):
        if mp_["tr"][o] != 3:
            c = oracle.mif_natural(mp_["tr"][o], 0, 0, float(theta0[o].mean()), 0.0, 0.0, 1.0)
            mn[o], mx[o] = c * 0.9 - 0.01, c * 1.1 + 0.01
    mp_ = dict(mp_, min=mn, max=mx, prior=[(o % 4) - 1 for o in range(nth)],
               prior_pow=[1.0 / N] * nth, prior_until=[L, L] + [N] * (nth - 2))
    return m, times, data, pv, mp_, theta0, x0, mut


def test_prior_pass_sirs4_equals_oracle(ctx):
    N, J = 20, 1500
    m, times, data, pv, mp_, theta0, x0, mut = sirs4_prior(N, J, 6)
    print("mif prior sirs4", flush=True)
    g = run_pass(ctx, m, J, times, data, pv, mp_, theta0, x0, mut)
    o = oracle.mif_pass(m.id, J, times, data, pv, mp_, theta0, x0, mut, seed=6, stream=2, nthreads=8,
                        like_min=LIKE_MIN)
    check_equal(g, o)


def test_prior_pass_ode_equals_oracle(ctx):
    N, J = 12, 400
    m, times, data, pv, mp_, theta0, x0 = prior_setup(N, J, 4, seed=9)
    mut = np.full(N, 0.7)
    print("mif prior ode", flush=True)
    g = run_pass(ctx, m, J, times, data, pv, mp_, theta0, x0, mut, flags=smc.FLAG_ODE)
    o = oracle.mif_pass(m.id, J, times, data, pv, mp_, theta0, x0, mut, seed=6, stream=2, nthreads=8, ode=True,
                        like_min=LIKE_MIN)
    check_equal(g, o)


@pytest.mark.parametrize("lanes", [1, 2])
def test_prior_batched_with_passes_without(ctx, lanes):
    N = 12
    cases = []
    m, times, data, pv, mp_, theta0, x0 = prior_setup(N, 700, 4, seed=1)
    cases.append((m, times, data, pv, dict(mp_, sd=[0.0, 0.0, 0.03, 0.02, 0.05, 0.02]), theta0, x0))
    plain = {k: v for k, v in cases[0][4].items() if not k.startswith("prior")}
    cases.append((m, times, data, pv, plain, theta0, x0))
    m4, t4, d4, pv4, mp4, th4, x04, _ = sirs4_prior(N, 900, 5)
    cases.append((m4, t4, d4, pv4, mp4, th4, x04))
    fs, alone = [], []
    mut = np.linspace(0.5, 1.0, N)
    for k, (mm, t, d, p, mpk, th, x) in enumerate(cases):
        print("mif prior batch", mm.name, k, flush=True)
        f = smc.Filter(ctx, mm, th.shape[1], t, d, p, x[:, 0], flags=smc.FLAG_MIF, mif=mpk, like_min=LIKE_MIN)
        fs.append(f)
        alone.append(oracle.mif_pass(mm.id, th.shape[1], t, d, p, mpk, th, x, mut, seed=3, stream=k, nthreads=8,
                                     like_min=LIKE_MIN))
    b = smc.Batch(ctx, fs)
    b.configure(lanes)
    for rep in range(2):
        for k, (mm, t, d, p, mpk, th, x) in enumerate(cases):
            fs[k].mif_start(th, x, mut, seed=3, stream=k)
        b.launch(use_graph=True)
        for f, o in zip(fs, alone):
            check_equal(f.read(mif=True), o)
    b.close()
    for f in fs:
        f.close()


def test_prior_refusals(ctx):
    N = 12
    m, times, data, pv, mp_, theta0, x0 = prior_setup(N, 10, 2)
    bad = [dict(mp_, prior_pow=None), dict(mp_, prior=[5] + [0] * 5), dict(mp_, min=[1.0] * 6, max=[1.0] * 6),
           dict(mp_, prior_pow=[-1.0] * 6)]
    for mpb in bad:
        if mpb.get("prior_pow") is None:
            with pytest.raises((smc.PlomError, TypeError, ValueError)):
                smc.Filter(ctx, m, 10, times, data, pv, x0[:, 0], flags=smc.FLAG_MIF, mif=mpb)
            continue
        with pytest.raises(smc.PlomError):
            smc.Filter(ctx, m, 10, times, data, pv, x0[:, 0], flags=smc.FLAG_MIF, mif=mpb)
