"""School-term forcing on the GPU: the example model school_terms (terms_forcing(e) on the
transmission rate, a holiday calendar per population; pinned on the CPU in
test_school_terms.py) compiled into its own model library, against its CPU oracle.  FP64
sde, psr and ode: bit-exact log-likelihoods, ESS, ancestors and final cloud; with an empty
calendar (forcing exactly 1) the GPU filter equals the GPU filter of noise_no_remainder,
in FP64 and in FP32 (the single-precision forcing)."""
import numpy as np
import pytest

from iip_amd import smc
from oracle import oracle
import user_model as U

pytestmark = pytest.mark.gpu

NAME = "school_terms"


@pytest.fixture(scope="module")
def model():
    return U.load(name=NAME)


def inputs(m, N, psr=False):
    times, data = m.dataset(N)
    x0 = m.x0(U.X0_PROPS)
    return times, data, m.pack(U.TRUTHS[NAME]), np.round(x0) if psr else x0, m.covariates()


@pytest.mark.parametrize("J,N,flags", [(3000, 60, 0), (1000, 40, smc.FLAG_PSR), (300, 20, smc.FLAG_ODE)])
def test_filter_equals_oracle(ctx, model, J, N, flags):
    psr, ode = bool(flags & smc.FLAG_PSR), bool(flags & smc.FLAG_ODE)
    times, data, pv, x0, cov = inputs(model, N, psr)
    print("school terms", J, N, flags, flush=True)
    g = smc.run(ctx, model, J, times, data, pv, x0, ancestry=True, final_state=True, flags=flags, seed=4, stream=8,
                cov=cov)
    o = oracle.smc(NAME, J, times, data, pv, x0, seed=4, stream=8, ancestry=True, final_state=True, nthreads=8, cov=cov,
                   psr=psr, ode=ode)
    assert np.isfinite(g.loglik)
    assert g.loglik == o["loglik"]
    np.testing.assert_array_equal(g.loglik_n, o["loglik_n"])
    np.testing.assert_array_equal(g.ess_n, o["ess_n"])
    np.testing.assert_array_equal(g.ancestry, o["ancestry"])
    np.testing.assert_array_equal(g.final_state, o["final_state"])


def test_empty_calendar_equals_the_model_without_forcing(ctx, model):
    base = U.load()
    times, data, pv, x0, cov = inputs(model, 50)
    lay = model.desc.layout
    empty = np.zeros(18)
    empty[1] = 1.0            # no periods: prop_school 1
    pv[lay["school"]:] = np.concatenate([empty, empty])
    print("school terms, empty calendar", flush=True)
    a = smc.run(ctx, model, 2000, times, data, pv, x0, seed=7, stream=3, cov=cov)
    b = smc.run(ctx, base, 2000, times, data, base.pack(U.TRUTH), x0, seed=7, stream=3, cov=cov)
    assert a.loglik == b.loglik
    np.testing.assert_array_equal(a.loglik_n, b.loglik_n)
    c = smc.run(ctx, model, 2000, times, data, model.pack(U.TRUTHS[NAME]), x0, seed=7, stream=3, cov=cov)
    assert c.loglik != a.loglik


def test_fp32_forcing(ctx, model):
    """The single-precision forcing: with an empty calendar it is exactly 1 (1 / (1 + e 0)),
    so the FP32 filter equals the FP32 filter of noise_no_remainder bit for bit; the
    calendar changes it."""
    base = U.load()
    times, data, pv, x0, cov = inputs(model, 50)
    lay = model.desc.layout
    empty = np.zeros(18)
    empty[1] = 1.0
    pe = pv.copy()
    pe[lay["school"]:] = np.concatenate([empty, empty])
    print("school terms fp32", flush=True)
    a = smc.run(ctx, model, 4000, times, data, pe, x0, precision="fp32", seed=5, stream=1, cov=cov)
    b = smc.run(ctx, base, 4000, times, data, base.pack(U.TRUTH), x0, precision="fp32", seed=5, stream=1, cov=cov)
    c = smc.run(ctx, model, 4000, times, data, pv, x0, precision="fp32", seed=5, stream=1, cov=cov)
    assert np.isfinite(a.loglik) and np.isfinite(c.loglik)
    assert a.loglik == b.loglik
    np.testing.assert_array_equal(a.loglik_n, b.loglik_n)
    assert c.loglik != a.loglik
