"""Drifting parameters (diffusions: compute_drift, C/core/drift.c:27-42, and the drifted
rates of step_sde_*, C/templates/core_template.c:251-271) in the CPU oracle, pinned without
the GPU.  Model: the reference's drift example (r0 diffuses on the log scale with volatility
vol).  With the demographic and environmental noises off, the only normals a particle draws
are the drift's, so

* the drift state after every substep is its previous value plus (vol_g sqrt(dt)) z_g, z_g the
  particle's Box-Muller normals in group order (one pair per two groups, per substep):
  checked bit for bit against the stream itself; populations of one group move together;
* the states follow the Euler scheme of the model with r0 = exp(drift state) at the start
  of every substep (numpy brute force with scipy covariate splines, 1e-10);
* hat reports the drifting parameter on its natural scale (f_inv, hat.c:141-168).
"""
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import oracle
from iip_amd import configs
from iip_amd.smc import Model


def setup(N, groups=None, vol=None):
    """vol: the natural volatility of every group (written into the pv slot the map names)."""
    m = Model.get("drift")
    times, data = m.dataset(N)
    pv = m.pack(configs.TRUTH["drift"])
    x0, drift = configs.drift_setup(m, m.x0(configs.X0_PROPS["drift"]), configs.TRUTH["drift"], groups=groups)
    if vol is not None:
        for g, v in enumerate(vol):
            if drift["vol_slot"][g] >= 0:
                pv[drift["vol_slot"][g]] = v
    return m, times, data, pv, x0, drift


def substeps(t0, t1, dt=0.25):
    k, t = 0, t0
    while t < t1:
        t += dt
        k += 1
    return k


@pytest.mark.parametrize("groups,vol", [([[0, 1]], [0.05, 0.08]), ([[0, 0]], [0.07, 0.0]), ([[1, 0]], [0.05, 0.08])])
def test_drift_is_the_random_walk_of_the_stream(groups, vol):
    m, times, data, pv, x0, drift = setup(12, groups, vol)
    seed, stream, j = 9, 4, 3
    got = oracle.trajectory(m.id, times, pv, x0, noises_off=3, seed=seed, stream=stream, j=j, drift=drift)
    off = m.desc.drift_offset
    sqdt = math.sqrt(0.25)
    ng = max(groups[0]) + 1
    x = [x0[off], x0[off + 1]]
    for n in range(12):
        K = substeps(times[n], times[n + 1])
        per = 2 * ((ng + 1) // 2)                      # normals drawn per substep (whole pairs)
        z = oracle.normals(seed, stream, n, j, per * K)
        for k in range(K):
            for c in (0, 1):
                g = groups[0][c]
                x[c] = x[c] + (vol[g] * sqdt) * z[per * k + g]
        assert got[n + 1, off] == x[0] and got[n + 1, off + 1] == x[1], n
    if groups[0][0] == groups[0][1]:
        np.testing.assert_array_equal(got[:, off], got[:, off + 1])


def test_states_follow_the_drifted_rates():
    m, times, data, pv, x0, drift = setup(20, vol=[0.05, 0.08])
    cov = m.covariates()
    got = oracle.trajectory(m.id, times, pv, x0, noises_off=3, seed=2, stream=1, j=0, drift=drift, cov=cov)
    lay, fo = m.layout, m.layout["fix_off"]
    sp = [CubicSpline(x, y, bc_type="natural") if len(x) >= 4 else
          (lambda t, x=x, y=y: np.interp(t, x, y)) if len(x) > 1 else (lambda t, y=y: y[0]) for x, y in cov]
    c = lambda name, cac, t: float(sp[lay["fixed"][name] + cac - fo](t))
    v = pv[lay["proc"]["v"]]
    off = m.desc.drift_offset
    S, I = [x0[0], x0[1]], [x0[2], x0[3]]
    for n in range(20):
        t = times[n]
        K = substeps(times[n], times[n + 1])
        # the drift path of this interval comes from the oracle's own (exactly pinned) drift states:
        # rebuild it from the stream as above
        z = oracle.normals(2, 1, n, 0, 2 * K)
        d = [got[n, off], got[n, off + 1]]
        inc_out = inc_in = 0.0
        for k in range(K):
            for cac in (0, 1):
                N, mb, md = c("N", cac, t), c("mu_b", cac, t), c("mu_d", cac, t)
                r0 = math.exp(d[cac])
                inf = r0 / N * v * I[cac] * S[cac]
                dS = mb * N - inf - md * S[cac]
                dI = inf - v * I[cac] - md * I[cac]
                inc_out += (v * I[cac] + md * I[cac]) * 0.25
                if cac == 1:
                    inc_in += inf * 0.25
                S[cac] = max(S[cac] + dS * 0.25, 0.0)
                I[cac] = max(I[cac] + dI * 0.25, 0.0)
            for cac in (0, 1):
                d[cac] = d[cac] + (pv[drift["vol_slot"][cac]] * 0.5) * z[2 * k + cac]
            t += 0.25
        ref = np.array([S[0], S[1], I[0], I[1], inc_out, inc_in])
        np.testing.assert_allclose(got[n + 1, [0, 1, 2, 3, 6, 7]], ref, rtol=1e-10, atol=1e-8)


def test_hat_reports_the_natural_parameter_and_no_drift_keeps_pv():
    m, times, data, pv, x0, drift = setup(15, vol=[0.05, 0.08])
    o = oracle.smc(m.id, 1, times, data, pv, x0, noises_off=3, drift=drift, hat=True, seed=2, stream=1)
    path = oracle.trajectory(m.id, times, pv, x0, noises_off=3, seed=2, stream=1, j=0, drift=drift)
    off, nh = m.desc.drift_offset, m.n_hat
    for n in range(16):
        for cac in (0, 1):
            assert o["hat"][n, nh - 2 + cac, 1] == pytest.approx(math.exp(path[n, off + cac]), rel=1e-15)
    # without the drift map (--no_diff): r0 keeps its pv value, the drift states do not move,
    # hat reports NaN for them
    a = oracle.smc(m.id, 1, times, data, pv, x0, noises_off=3, hat=True, seed=2, stream=1)
    assert np.isnan(a["hat"][:, nh - 2:, :]).all()
    p2 = oracle.trajectory(m.id, times, pv, x0, noises_off=3, seed=2, stream=1, j=0)
    np.testing.assert_array_equal(p2[:, off:off + 2], np.repeat(x0[None, off:off + 2], 16, axis=0))
    assert a["loglik"] != o["loglik"]


# iterated filtering with drift: theta = r0 of each city (the drift states' initial values,
# fitted by fixed-lag smoothing: no walk), v, vol of each city (the particle's own
# volatility), rep, phi
DRIFT_MAP = {"pv_src": [-1, -1, 2, 2, 3, 4, 5, 5, 5, 5, 6, 6, 6, 6], "tr": [1, 1, 1, 1, 1, 2, 1], "scale": [0] * 7,
             "duration": [0] * 7, "min": [0.0] * 7, "max": [0.0] * 7, "multiplier": [1.0] * 7,
             "sd": [0.0, 0.0, 0.02, 0.05, 0.05, 0.02, 0.02]}


def mif_drift_setup(N, J, seed, spread=0.0):
    m, times, data, pv, x0, drift = setup(N)
    t = configs.TRUTH["drift"]
    th = np.array([math.log(t["r0"]), math.log(t["r0"]) + 0.1, math.log(t["v"]), math.log(t["vol"]),
                   math.log(t["vol"]) + 0.3, math.log(t["rep"] / (1 - t["rep"])), math.log(t["phi"])])
    rng = np.random.default_rng(seed)
    theta0 = th[:, None] + spread * rng.normal(0.0, 1.0, (7, J))
    X0 = np.repeat(x0[:, None], J, axis=1)
    off = m.desc.drift_offset
    X0[off:off + 2] = theta0[0:2]
    return m, times, data, pv, drift, theta0, X0


def natural_pv(pv, theta):
    out = pv.copy()
    for s, o in enumerate(DRIFT_MAP["pv_src"]):
        if o >= 0:
            out[s] = oracle.mif_natural(DRIFT_MAP["tr"][o], 0, 0, theta[o], 0.0, 0.0, 1.0)
    return out


def test_mif_pass_with_drift_and_equal_parameters_is_the_filter():
    """Equal theta in every particle, no walk: the pass is the drift filter of those
    parameters, bit for bit (the drift states start at theta's r0)."""
    N, J = 30, 600
    m, times, data, pv, drift, theta0, X0 = mif_drift_setup(N, J, 1)
    pv_eq = natural_pv(pv, theta0[:, 0])
    r = oracle.mif_pass(m.id, J, times, data, pv_eq, dict(DRIFT_MAP, sd=[0.0] * 7), theta0, X0, np.full(N, 0.7),
                        seed=5, stream=3, nthreads=8, drift=drift, cov=m.covariates())
    o = oracle.smc(m.id, J, times, data, pv_eq, X0[:, 0], seed=5, stream=3, nthreads=8, drift=drift, cov=m.covariates())
    assert r["loglik"] == o["loglik"]
    np.testing.assert_array_equal(r["loglik_n"], o["loglik_n"])
    np.testing.assert_array_equal(r["ess_n"], o["ess_n"])


def test_mif_drift_uses_each_particles_volatility():
    """Different theta per particle, only the drift's noise: at the first interval every
    particle's log-weight is the likelihood of its own path (its volatility, its initial
    drift state, its stream)."""
    N, J = 6, 40
    m, times, data, pv, drift, theta0, X0 = mif_drift_setup(N, J, 2, spread=0.3)
    cov = m.covariates()
    r = oracle.mif_pass(m.id, J, times, data, pv, DRIFT_MAP, theta0, X0, np.full(N, 0.7), noises_off=3, seed=7,
                        stream=2, drift=drift, cov=cov, logw=True)
    tab, kmax = oracle.cov_table(m.id, times, 0.25, cov)
    for j in range(J):
        pvj = natural_pv(pv, theta0[:, j])
        path = oracle.trajectory(m.id, times[:2], pvj, X0[:, j], noises_off=3, seed=7, stream=2, j=j, drift=drift,
                                 cov=cov)
        pvo = pvj.copy()
        pvo[m.fix_off:] = tab[N * kmax + 1]
        obs, ll, nf = oracle.obs_loglik(m.id, path[1], data[0], pvo, times[1])
        assert r["logw"][0][j] == (-math.inf if nf == m.n_ts else ll), j   # every series floored: failure
    # the volatility matters: the drift states of two particles with different vol differ
    assert len(set(np.round(r["logw"][0], 12))) > J // 2
