"""A user's model outside the library (CPU part): its model library registers through
plom_model_load with the dimensions the code generator gives it, bad libraries are
refused, and its CPU oracle is pinned on the no-remainder rewrite (N -> S + I + R,
Cmodel.py:110-153): a deterministic Euler scheme written here from the reaction list, the
population balance that the rewrite implies, and the one-particle likelihood against
scipy."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.stats import norm

from iip_amd import smc
from oracle import oracle
import user_model as U


@pytest.fixture(scope="module")
def model():
    return U.load(build_if_missing=True)


def test_model_library_registers(model):
    d = model.desc
    # after the 5 built-in models (other user models loaded by earlier tests may follow it)
    assert 5 <= model.id < smc.lib().plom_model_count()
    assert d.remainder is None and d.par_sv == ["S", "I", "R"]
    mi = smc.ModelInfo()
    smc.check(smc.lib().plom_model_info_get(model.id, C.byref(mi)))
    assert (mi.size_proj, mi.n_ts, mi.n_cac, mi.n_par_sv, mi.n_hat) == (8, 4, 2, 3, 10)
    assert mi.pv_size == d.layout["size"]
    buf = C.create_string_buffer(64)
    smc.check(smc.lib().plom_model_name(model.id, buf, 64))
    assert buf.value.decode() == U.NAME
    assert smc.Model.load(U.DIR, U.LIBRARY).id == model.id          # registered once
    i = C.c_int32()
    assert smc.lib().plom_model_load(b"/nonexistent/libplom_model_x.so", C.byref(i)) == smc.ERR_MODEL
    # a library that is not a model library
    assert smc.lib().plom_model_load(smc.lib()._name.encode(), C.byref(i)) == smc.ERR_MODEL
    assert b"plom_model_plugin" in smc.lib().plom_last_error()


def euler_no_remainder(m, pv, x0, times, dt, cov):
    """Deterministic Euler of the model with the covariates (scipy splines) at every substep
    start: births mu_b (S+I+R), infection r0/(S+I+R) v I S, recovery v I, deaths mu_d S,
    mu_d I, mu_d R; inc_out = recoveries + deaths of I over both cities, inc_in =
    infections in city 2."""
    lay = m.layout
    fo = lay["fix_off"]
    sp = [CubicSpline(x, y, bc_type="natural") if len(x) >= 4 else
          (lambda t, x=x, y=y: np.interp(t, x, y)) if len(x) > 1 else (lambda t, y=y: y[0]) for x, y in cov]
    c = lambda name, cac, t: float(sp[lay["fixed"][name] + cac - fo](t))
    r0, v = pv[lay["proc"]["r0"]], pv[lay["proc"]["v"]]
    S, I, R = [x0[0], x0[1]], [x0[2], x0[3]], [x0[4], x0[5]]
    rows = [np.concatenate([S, I, R, [0.0, 0.0]])]
    for n in range(len(times) - 1):
        inc_out, inc_in = 0.0, 0.0
        t = times[n]
        while t < times[n + 1]:
            for cac in (0, 1):
                P = S[cac] + I[cac] + R[cac]
                mb, md = c("mu_b", cac, t), c("mu_d", cac, t)
                inf = r0 / P * v * I[cac] * S[cac]
                dS = mb * P - inf - md * S[cac]
                dI = inf - v * I[cac] - md * I[cac]
                dR = v * I[cac] - md * R[cac]
                inc_out += (v * I[cac] + md * I[cac]) * dt
                if cac == 1:
                    inc_in += inf * dt
                S[cac] = max(S[cac] + dS * dt, 0.0)
                I[cac] = max(I[cac] + dI * dt, 0.0)
                R[cac] = max(R[cac] + dR * dt, 0.0)
            t += dt
        rows.append(np.concatenate([S, I, R, [inc_out, inc_in]]))
    return np.array(rows), sp


def test_no_remainder_path_and_likelihood(model):
    m = model
    N = 30
    times, data = m.dataset(N)
    cov = m.covariates()
    pv, x0 = m.pack(U.TRUTH), m.x0(U.X0_PROPS)
    ref, sp = euler_no_remainder(m, pv, x0, times, 0.25, cov)
    got = oracle.trajectory(U.NAME, times, pv, x0, noises_off=3, cov=cov)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-6)
    # the rewrite makes the population its own: S + I + R grows by (mu_b - mu_d) of itself,
    # and leaves the covariate N (which the model with a remainder follows) behind
    lay, fo = m.layout, m.layout["fix_off"]
    for cac in (0, 1):
        tot = got[:, cac] + got[:, 2 + cac] + got[:, 4 + cac]
        t, P = times[0], tot[0]
        for n in range(N):
            while t < times[n + 1]:
                P = P + (float(sp[lay["fixed"]["mu_b"] + cac - fo](t)) - float(sp[lay["fixed"]["mu_d"] + cac - fo](t))) \
                    * P * 0.25
                t += 0.25
            assert tot[n + 1] == pytest.approx(P, rel=1e-9)
        Nt = np.array([float(sp[lay["fixed"]["N"] + cac - fo](x)) for x in times])
        assert np.max(np.abs(tot - Nt) / Nt) > 1e-3
    # one particle, no noise: loglik_n is the log-likelihood of its path
    o = oracle.smc(U.NAME, 1, times, data, pv, x0, noises_off=3, cov=cov)
    rep, phi = pv[lay["obs"]["rep"]], pv[lay["obs"]["phi"]]
    for n in range(N):
        x = ref[n + 1]
        obs = [x[6], x[6], x[7], x[2]]               # inc_out (CDC, google), inc_in, prevalence city 1
        ll = 0.0
        for ts in range(4):
            y = data[n, ts]
            if math.isnan(y):
                continue
            prop = float(sp[lay["fixed_obs"]["prop"] + ts - fo](times[n + 1]))
            mu = rep * prop * obs[ts]
            sd = math.sqrt(rep * (1 - rep) * prop * obs[ts] + (rep * phi * prop * obs[ts]) ** 2)
            like = norm.cdf(y + 0.5, mu, sd) - (norm.cdf(y - 0.5, mu, sd) if y > 0 else 0.0)
            ll += math.log(max(like, 1e-17))
        assert o["loglik_n"][n] == pytest.approx(ll, rel=1e-9, abs=1e-9), n
