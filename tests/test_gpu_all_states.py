"""PLOM_FLAG_ALL_STATES (smc -r: every particle after every interval, print_X,
C/core/print.c:360-392) against the CPU oracle: the stored clouds as resampling leaves them
(row n = the propagated cloud of interval n-1 gathered through the ancestry; row 0 = x0)
bit for bit, and every particle's observation mean and drawn observation equal to the
oracle's plom_obs_draw on that state (slot j + 1), with constant and time-varying
covariates; the smc -r file carries the same values."""
import numpy as np
import pytest

from oracle import oracle
from iip_amd import configs, smc
from iip_amd.smc import Model

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,J,N,cov", [("sir", 300, 20, False), ("noise", 257, 15, True), ("sirs4", 64, 10, False)])
def test_all_states_equal_oracle(ctx, name, J, N, cov):
    m = Model.get(name)
    times, data = m.dataset(N)
    data = data.copy()
    data[3] = np.nan                                   # an interval without resampling
    pv, x0 = m.pack(configs.TRUTH[name]), m.x0(configs.X0_PROPS[name])
    cv = m.covariates() if cov else None
    print("all states", name, J, N, cov, flush=True)
    f = smc.Filter(ctx, m, J, times, data, pv, x0, flags=smc.FLAG_ALL_STATES, seed=6, stream=3, cov=cv)
    f.run()
    g = f.read(all_states=True, ancestry=True)
    f.close()
    o = oracle.smc(m.id, J, times, data, pv, x0, seed=6, stream=3, all_states=True, ancestry=True, nthreads=8, cov=cv)
    np.testing.assert_array_equal(g.ancestry, o["ancestry"])
    ref = np.empty_like(o["all_states"])
    ref[0] = o["all_states"][0]
    for n in range(1, N + 1):
        ref[n] = o["all_states"][n][o["ancestry"][n - 1]]
    np.testing.assert_array_equal(g.all_states, ref)
    ctab = None
    if cov:
        ctab, kmax = oracle.cov_table(m.id, times, 0.25, cv)
    rng = np.random.default_rng(1)
    for n in range(N + 1):
        cvd = ctab[N * kmax + n] if cov else None
        for j in rng.choice(J, 5, replace=False):
            mean, draw = oracle.obs_draw(m.id, ref[n, j], pv, times[n], 6, 3, n, j + 1, cvd)
            np.testing.assert_array_equal(g.all_obs[n, j, :, 0], mean)
            np.testing.assert_array_equal(g.all_obs[n, j, :, 1], draw)
    assert np.all(g.all_obs[:, :, :, 1] >= 0.0)
