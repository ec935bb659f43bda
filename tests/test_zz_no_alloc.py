"""plom_smc_run_noalloc (the filter laid out in the library's fixed device scratch, no device
allocation) against plom_smc_run on the same inputs, through the C ABI: bit-identical
log-likelihood, per-interval log-likelihoods and ESS, ancestry, final cloud and final
log-weights, on calls in a row whose later inputs are smaller (so they run over the earlier
calls' leftovers in the scratch), and the size limit."""
import numpy as np
import pytest

from iip_amd import configs, smc
from iip_amd.smc import Model

pytestmark = pytest.mark.gpu


def inputs(name, N):
    m = Model.get(name)
    times, data = m.dataset(N)
    pv = m.pack(configs.TRUTH[name])
    x0 = m.x0(configs.X0_PROPS[name])
    return m, times, data, pv, x0


def same(a, b):
    assert a.loglik == b.loglik and a.n_all_fail == b.n_all_fail
    np.testing.assert_array_equal(a.loglik_n, b.loglik_n)
    np.testing.assert_array_equal(a.ess_n, b.ess_n)
    np.testing.assert_array_equal(a.final_logw, b.final_logw)
    if a.ancestry is not None:
        np.testing.assert_array_equal(a.ancestry, b.ancestry)
    if a.final_state is not None:
        np.testing.assert_array_equal(a.final_state, b.final_state)


CASES = [
    # (model, J, N, flags, precision): each call smaller than the one before
    ("sirs4", 20000, 104, smc.FLAG_STORE_ANCESTRY | smc.FLAG_FINAL_STATE, "fp64"),
    ("sirs4", 3000, 52, smc.FLAG_PSR | smc.FLAG_FINAL_STATE, "fp64"),
    ("sir", 700, 20, smc.FLAG_STORE_ANCESTRY | smc.FLAG_FINAL_STATE, "fp32"),
]


def test_noalloc_equals_smc_run_on_calls_in_a_row(ctx):
    prev = None
    for name, J, N, flags, prec in CASES:
        m, times, data, pv, x0 = inputs(name, N)
        if flags & smc.FLAG_PSR:
            x0 = np.round(x0)
        desc = smc._desc(m, J, N, prec, 0, flags, 5, 1, 0.25, 1e-17)
        nbytes = smc.lib().plom_filter_device_bytes(smc.C.byref(desc))
        assert 0 < nbytes <= smc.NOALLOC_MAX_BYTES
        if prev is not None:
            assert nbytes < prev
        prev = nbytes
        print("noalloc case", name, J, N, flags, prec, nbytes, flush=True)
        a = smc.smc_run(ctx, m, J, times, data, pv, x0, noalloc=True, precision=prec, flags=flags, seed=5, stream=1)
        b = smc.smc_run(ctx, m, J, times, data, pv, x0, noalloc=False, precision=prec, flags=flags, seed=5, stream=1)
        assert np.isfinite(a.loglik)
        same(a, b)


def test_noalloc_refuses_a_filter_larger_than_the_scratch(ctx):
    m, times, data, pv, x0 = inputs("sirs4", 10)
    J = 400000
    desc = smc._desc(m, J, 10, "fp64", 0, 0, 1, 0, 0.25, 1e-17)
    assert smc.lib().plom_filter_device_bytes(smc.C.byref(desc)) > smc.NOALLOC_MAX_BYTES
    print("noalloc refusal", J, flush=True)
    with pytest.raises(smc.PlomError) as e:
        smc.smc_run(ctx, m, J, times, data, pv, x0, noalloc=True)
    assert e.value.code == smc.ERR_ARG and "fixed scratch" in str(e.value)
    # the next call that fits runs as before
    a = smc.smc_run(ctx, m, 1000, times, data, pv, x0, noalloc=True, seed=2)
    b = smc.smc_run(ctx, m, 1000, times, data, pv, x0, seed=2)
    same(a, b)
