"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so). TEST INFRASTRUCTURE:
imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as
the checker of the HIP filter, never as the thing measured or shipped."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
CANONICAL, REFERENCE = 0, 1
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_extra = {}


def extra_lib(name, gen_dir):
    """The oracle with a user's model (generated header plom_<name>.h in gen_dir) compiled
    in after the built-in models (make extra); its model name then resolves in smc(),
    mif_pass(), trajectory() ... like a built-in id."""
    if name not in _extra:
        path = os.path.join(HERE, "_build", "liboracle_%s.so" % name)
        if not os.path.exists(path) or os.path.exists(os.path.join(HERE, "Makefile")):
            # make rebuilds it when the oracle's sources or the model's header are newer
            subprocess.check_call(["make", "-s", "-C", HERE, "extra", "NAME=" + name, "GEN=" + os.path.abspath(gen_dir)])
        _extra[name] = _setup(C.CDLL(path))
    return _extra[name]


def _resolve(model):
    """(library, model index) of a built-in model id or name, or of a registered extra model's name."""
    if not isinstance(model, str):
        return lib(), model
    L = lib()
    for i in range(L.oracle_model_count()):
        if L.oracle_model_name(i).decode() == model:
            return L, i
    if model in _extra:
        X = _extra[model]
        return X, X.oracle_model_count() - 1
    raise KeyError("unknown oracle model %r (register a user's model with extra_lib)" % model)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _setup(C.CDLL(LIB))
    return _lib


def _setup(L):
    """argument and result types of the oracle's entry points"""
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    L.oracle_smc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                             C.c_double, C.c_double, dp, dp, dp, dp, C.c_int, C.c_int,
                             dp, ip, dp, dp, ip, dp, dp, dp, ip, dp, dp, dp, dp, dp, ip, dp, dp, C.c_double, C.c_double,
                             ip, ip, ip, ip, dp, dp, dp, ip]
    L.oracle_trajectory.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_double, dp, dp, dp, dp, ip, dp, dp, C.c_double, C.c_double,
                                    ip, ip, ip, ip, dp, dp, dp, ip]
    L.oracle_ode_test.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, ip]
    L.oracle_cov_table.argtypes = [C.c_int, C.c_int, dp, C.c_double, ip, dp, dp, dp, ip]
    L.oracle_spline.argtypes = [C.c_int, dp, dp, C.c_int, dp, dp]
    L.oracle_obs_draw.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, dp]
    L.oracle_obs_loglik.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_double, dp, dp, ip]
    L.oracle_math.argtypes = [C.c_int, dp, dp]
    L.oracle_normals.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, dp]
    L.oracle_philox.argtypes = [C.POINTER(C.c_uint32)] * 3
    L.oracle_bm_pairs.argtypes = [C.c_int, C.POINTER(C.c_uint32), dp]
    L.oracle_icdf.argtypes = [C.c_int, C.POINTER(C.c_uint32), dp]
    L.oracle_log_tab.argtypes = [C.c_int, dp, dp]
    L.oracle_terms_forcing.argtypes = [C.c_int, C.c_double, dp, dp, dp]
    L.oracle_xoshiro.argtypes = [C.POINTER(C.c_uint32), C.c_int, C.POINTER(C.c_uint32)]
    L.oracle_model_name.restype = C.c_char_p
    L.oracle_psr_draws.argtypes = [C.c_int, C.c_double, C.c_double, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.c_int, dp, dp]
    L.oracle_mif_natural.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    L.oracle_mif_natural.restype = C.c_double
    i32p = C.POINTER(C.c_int32)
    L.oracle_mif_pass.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double,
                                  C.c_double, dp, dp, dp, C.c_int, i32p, i32p, i32p, i32p, dp, dp, dp, dp, dp,
                                  dp, dp, C.c_int, dp, i32p, dp, dp, dp, dp, dp, dp, i32p, i32p, dp, dp, C.c_double,
                                  C.c_double, dp, i32p, dp, i32p, ip, ip, ip, ip, dp, dp, dp, ip]
    L.oracle_mif_prior_pdf.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
    L.oracle_mif_prior_pdf.restype = C.c_double
    L.oracle_mif_prior_term.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double]
    L.oracle_mif_prior_term.restype = C.c_double
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None


IMPL_PSR = 0x100   # the psr implementation (PLOM_IMPL_PSR of csrc/plom_model_common.h)
IMPL_ODE = 0x200   # the ode implementation (PLOM_IMPL_ODE)


def _impl(noises_off, psr, ode):
    return noises_off | (IMPL_PSR if psr else 0) | (IMPL_ODE if ode else 0)


_keep_drift = []


def _drift_args(drift):
    """drift: dict of group, tr, scale, duration [n_drift * n_cac] ints and min, max,
    multiplier doubles and vol_slot ints (plom_drift_map), or None."""
    if drift is None:
        return (None,) * 8
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    arr = [i32(drift["group"]), i32(drift["tr"]), i32(drift["scale"]), i32(drift["duration"]), f64(drift["min"]),
           f64(drift["max"]), f64(drift["multiplier"]), i32(drift["vol_slot"])]
    _keep_drift[:] = arr
    return tuple(_ip(a) for a in arr[:4]) + tuple(_dp(a) for a in arr[4:7]) + (_ip(arr[7]),)


def _eps(eps_abs, eps_rel):
    """Both 0: the reference's defaults (1e-6, 1e-3), as the library takes them."""
    return (eps_abs, eps_rel) if (eps_abs > 0 or eps_rel > 0) else (1e-6, 1e-3)


def ode_test(kind, p, y, t0, t1, h, eps_abs, eps_rel):
    """The RKF45 integrator of csrc/plom_ode_impl.h on an analytic system (see oracle_ode_test):
    (y at t1, right-hand side evaluations, status)."""
    pp = np.zeros(5)
    pp[:len(p)] = p
    yy = np.zeros(2)
    yy[:len(y)] = y
    n = C.c_int32()
    rc = lib().oracle_ode_test(kind, _dp(pp), _dp(yy), t0, t1, h, eps_abs, eps_rel, C.byref(n))
    return yy[:len(y)], n.value, rc


def _cov(cov):
    """cov: one (x, y) knot pair per covariate slot of pv, or None -> (n, x, y) arrays."""
    if cov is None:
        return None, None, None
    n = np.ascontiguousarray([len(x) for x, _ in cov], dtype=np.int32)
    xs = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x, _ in cov]), dtype=np.float64)
    ys = np.ascontiguousarray(np.concatenate([np.asarray(y, dtype=np.float64) for _, y in cov]), dtype=np.float64)
    return n, xs, ys


_keep = []


def _cov_args(cov):
    n, xs, ys = _cov(cov)
    _keep[:] = [n, xs, ys]
    return _ip(n), _dp(xs), _dp(ys)


def cov_table(model_id, times, dt, cov):
    """The covariate table of a filter (csrc/plom_cov.h): (table [rows][n_fix], kmax)."""
    L, model_id = _resolve(model_id)
    times = np.ascontiguousarray(times, dtype=np.float64)
    N = len(times) - 1
    nfix = L.oracle_model_n_fix(model_id)
    kmax_bound = int(np.max(np.diff(times)) / dt) + 2 if N else 0
    out = np.zeros(((N * kmax_bound + N + 1), max(nfix, 1)))
    km = C.c_int32()
    n, xs, ys = _cov(cov)
    rc = L.oracle_cov_table(model_id, N, _dp(times), dt, _ip(n), _dp(xs), _dp(ys), _dp(out), C.byref(km))
    if rc != 0:
        raise ValueError("covariate table: %d" % rc)
    rows = N * km.value + N + 1
    return out.reshape(-1)[:rows * nfix].reshape(rows, nfix), km.value


def spline(x, y, t):
    """The interpolating spline of csrc/plom_cov.h through (x, y), at the times t."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    out = np.zeros(len(t))
    if lib().oracle_spline(len(x), _dp(x), _dp(y), len(t), _dp(t), _dp(out)) != 0:
        raise RuntimeError("oracle_spline failed")
    return out


def smc(model_id, J, times, data, pv, x0, noises_off=0, flags=0, seed=1, stream=0, dt=0.25,
        like_min=1e-17, mode=CANONICAL, nthreads=1, ancestry=False, final_state=False, logw=False,
        lineage=False, all_states=False, hat=False, psr=False, cov=None, ode=False, eps_abs=0.0, eps_rel=0.0,
        drift=None):
    """The filter restated on the CPU.  lineage: the sampled path (slots [N], states
    [N+1][S]); all_states: every particle's state after every interval, before resampling
    ([N+1][J][S], for tests on small runs); psr: the Euler-multinomial implementation;
    cov: time-varying covariates, one (x, y) knot pair per covariate slot of pv; ode: the
    adaptive RKF45 implementation under eps_abs / eps_rel."""
    L, model_id = _resolve(model_id)
    noises_off = _impl(noises_off, psr, ode)
    times = np.ascontiguousarray(times, dtype=np.float64)
    N = len(times) - 1
    data = np.ascontiguousarray(data, dtype=np.float64)
    pv = np.ascontiguousarray(pv, dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    S = L.oracle_model_size_proj(model_id)
    ll = C.c_double()
    naf = C.c_int32()
    ll_n = np.zeros(N)
    ess = np.zeros(N)
    anc = np.zeros((N, J), dtype=np.int32) if ancestry else None
    fs = np.zeros((S, J)) if final_state else None
    lw = np.zeros((N, J)) if logw else None
    lsl = np.zeros(N, dtype=np.int32) if lineage else None
    lst = np.zeros((N + 1, S)) if lineage else None
    alls = np.zeros((N + 1, J, S)) if all_states else None
    NH = L.oracle_model_n_hat(model_id)
    n_ts = L.oracle_model_n_ts(model_id)
    ht = np.zeros((N + 1, NH, 3)) if hat else None
    lob = np.zeros((N + 1, n_ts, 2)) if lineage else None
    pr = np.zeros((N, n_ts, 2)) if hat else None
    el = C.c_double()
    rc = L.oracle_smc(model_id, J, N, noises_off, flags, seed, stream, dt, like_min, _dp(times), _dp(data),
                          _dp(pv), _dp(x0), mode, nthreads, C.byref(ll), C.byref(naf), _dp(ll_n), _dp(ess),
                          _ip(anc), _dp(fs), _dp(lw), C.byref(el), _ip(lsl), _dp(lst), _dp(alls), _dp(ht), _dp(pr), _dp(lob),
                          *_cov_args(cov), *_eps(eps_abs, eps_rel), *_drift_args(drift))
    if rc != 0:
        raise RuntimeError("oracle_smc failed: %d" % rc)
    return {"loglik": ll.value, "n_all_fail": naf.value, "loglik_n": ll_n, "ess_n": ess,
            "ancestry": anc, "final_state": fs, "logw": lw, "elapsed_s": el.value,
            "lineage_slots": lsl, "lineage": lst, "lineage_obs": lob, "all_states": alls, "hat": ht, "pred_res": pr}


def mif_pass(model_id, J, times, data, pv, mif, theta0, x0, mut_fac, noises_off=0, seed=1, stream=0, dt=0.25,
             like_min=1e-17, nthreads=1, theta_out=False, logw=False, ancestry=False, psr=False, cov=None, ode=False,
             eps_abs=0.0, eps_rel=0.0, drift=None):
    """One pass of iterated filtering restated on the CPU (oracle_mif_pass).  mif: dict of
    pv_src [n_pp], tr, scale, duration, min, max, multiplier, sd [n_theta], optionally chol
    [n_theta, n_theta] (a correlated walk) and optionally prior, prior_pow, prior_until
    [n_theta] (the prior patch of the weights); theta0 [n_theta, J]; x0 [S, J]; mut_fac [N];
    drift: the drift map of a model with diffusions (the volatilities are read from each
    particle's own parameters)."""
    L, model_id = _resolve(model_id)
    noises_off = _impl(noises_off, psr, ode)
    times = np.ascontiguousarray(times, dtype=np.float64)
    N = len(times) - 1
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    m = {k: (i32 if k in ("pv_src", "tr", "scale", "duration") else f64)(mif[k])
         for k in ("pv_src", "tr", "scale", "duration", "min", "max", "multiplier", "sd")}
    nth = len(m["tr"])
    chol = f64(mif["chol"]).reshape(-1) if mif.get("chol") is not None else None
    pr = [i32(mif["prior"]), f64(mif["prior_pow"]), i32(mif["prior_until"])] if mif.get("prior") is not None \
        else [None, None, None]
    theta0, x0, mut_fac = f64(theta0), f64(x0), f64(mut_fac)
    data, pv = f64(data), f64(pv)
    ll, naf = C.c_double(), C.c_int32()
    ll_n, ess = np.zeros(N), np.zeros(N)
    mm, mv = np.zeros((N + 1, nth)), np.zeros((N + 1, nth))
    tho = np.zeros((N + 1, nth, J)) if theta_out else None
    lw = np.zeros((N, J)) if logw else None
    anc = np.zeros((N, J), dtype=np.int32) if ancestry else None
    rc = L.oracle_mif_pass(model_id, J, N, noises_off, seed, stream, dt, like_min, _dp(times), _dp(data), _dp(pv),
                               nth, _ip(m["pv_src"]), _ip(m["tr"]), _ip(m["scale"]), _ip(m["duration"]), _dp(m["min"]),
                               _dp(m["max"]), _dp(m["multiplier"]), _dp(m["sd"]), _dp(mut_fac), _dp(theta0), _dp(x0),
                               nthreads, C.byref(ll), C.byref(naf), _dp(ll_n), _dp(ess), _dp(mm), _dp(mv), _dp(tho),
                               _dp(lw), _ip(anc), *_cov_args(cov), *_eps(eps_abs, eps_rel), _dp(chol), _ip(pr[0]),
                               _dp(pr[1]), _ip(pr[2]), *_drift_args(drift))
    if rc != 0:
        raise RuntimeError("oracle_mif_pass failed: %d" % rc)
    return {"loglik": ll.value, "n_all_fail": naf.value, "loglik_n": ll_n, "ess_n": ess, "mif_mean": mm,
            "mif_var": mv, "theta": tho, "logw": lw, "ancestry": anc}


PSR_UNIFORM, PSR_GAMMA, PSR_POISSON, PSR_BINOMIAL, PSR_LFACT, PSR_MULTINOMIAL, PSR_UNIFORM32, PSR_LFACT2 = range(8)


def psr_draws(kind, a, count, b=0.0, seed=1, stream=0, n=0, j=0, p=None):
    """count draws of one PSR sampler (csrc/plom_psr.h) from the stream (seed, stream, n, j);
    for PSR_MULTINOMIAL, the counts of N = a over the probabilities p."""
    pp = np.ascontiguousarray(p, dtype=np.float64) if p is not None else None
    out = np.zeros(len(pp) if pp is not None else count)
    lib().oracle_psr_draws(kind, float(a), float(b), seed, stream, n, j, len(out), _dp(pp), _dp(out))
    return out


def mif_natural(tr, sc, dur, x, a, b, mult):
    """plom_mif_natural (csrc/plom_mif.h) through the oracle build."""
    return lib().oracle_mif_natural(int(tr), int(sc), int(dur), float(x), float(a), float(b), float(mult))


PRIOR_NONE, PRIOR_UNIFORM, PRIOR_NORMAL, PRIOR_PSEUDO_UNIFORM = -1, 0, 1, 2


def mif_prior_pdf(kind, x, a, b):
    """plom_mif_prior_pdf (csrc/plom_mif.h) through the oracle build."""
    return lib().oracle_mif_prior_pdf(int(kind), float(x), float(a), float(b))


def mif_prior_term(kind, e, tr, theta, a, b, like_min=1e-17):
    """plom_mif_prior_term: e log(max(prior(f_inv(theta)), like_min))."""
    return lib().oracle_mif_prior_term(int(kind), float(e), int(tr), float(theta), float(a), float(b), float(like_min))


def trajectory(model_id, times, pv, x0, noises_off=0, seed=1, stream=0, j=0, dt=0.25, psr=False, cov=None, ode=False,
               eps_abs=0.0, eps_rel=0.0, drift=None):
    L, model_id = _resolve(model_id)
    noises_off = _impl(noises_off, psr, ode)
    times = np.ascontiguousarray(times, dtype=np.float64)
    N = len(times) - 1
    S = L.oracle_model_size_proj(model_id)
    out = np.zeros((N + 1, S))
    rc = L.oracle_trajectory(model_id, N, noises_off, seed, stream, j, dt, _dp(times),
                                 _dp(np.ascontiguousarray(pv, dtype=np.float64)),
                                 _dp(np.ascontiguousarray(x0, dtype=np.float64)), _dp(out), *_cov_args(cov),
                                 *_eps(eps_abs, eps_rel), *_drift_args(drift))
    if rc != 0:
        raise RuntimeError("oracle_trajectory failed")
    return out


def obs_draw(model_id, x, pv, t, seed, stream, n, slot, cvd=None):
    """(observation means [n_ts], drawn observations [n_ts]) of state x (plom_obs_draw)."""
    L, model_id = _resolve(model_id)
    n_ts = L.oracle_model_n_ts(model_id)
    out = np.zeros(2 * n_ts)
    L.oracle_obs_draw(model_id, _dp(np.ascontiguousarray(x, dtype=np.float64)),
                          _dp(np.ascontiguousarray(pv, dtype=np.float64)),
                          _dp(np.ascontiguousarray(cvd, dtype=np.float64)) if cvd is not None else None, t, seed, stream,
                          n, slot, _dp(out))
    return out[0::2], out[1::2]


def obs_loglik(model_id, x, y, pv, t=0.0, like_min=1e-17):
    L, model_id = _resolve(model_id)
    n_ts = L.oracle_model_n_ts(model_id)
    obs = np.zeros(max(n_ts, 1))
    ll = C.c_double()
    nf = C.c_int32()
    L.oracle_obs_loglik(model_id, _dp(np.ascontiguousarray(x, dtype=np.float64)),
                            _dp(np.ascontiguousarray(y, dtype=np.float64)),
                            _dp(np.ascontiguousarray(pv, dtype=np.float64)), t, like_min, _dp(obs),
                            C.byref(ll), C.byref(nf))
    return obs, ll.value, nf.value


def math(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((x.size, 8))
    lib().oracle_math(x.size, _dp(x), _dp(out))
    return out


def normals(seed, stream, n0, j, n):
    out = np.zeros(n)
    lib().oracle_normals(seed, stream, n0, j, n, _dp(out))
    return out


def bm_pairs(xy):
    """Box-Muller pairs of the 32-bit word pairs xy[:, 0], xy[:, 1] (shape (n, 2))."""
    xy = np.ascontiguousarray(xy, dtype=np.uint32).reshape(-1, 2)
    out = np.zeros(xy.shape)
    lib().oracle_bm_pairs(len(xy), xy.ctypes.data_as(C.POINTER(C.c_uint32)), _dp(out))
    return out


def div_rc(x, c):
    """x / c as the generated steps divide by a covariate (plom_div_rc with the divisor's
    reciprocal)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    c = np.ascontiguousarray(np.broadcast_to(c, x.shape), dtype=np.float64).ravel()
    out = np.zeros(x.size)
    lib().oracle_div_rc(x.size, _dp(x), _dp(c), _dp(out))
    return out


def icdf(words):
    """The stream's Gaussians of the 32-bit words (plom_icdf_normal)."""
    w = np.ascontiguousarray(words, dtype=np.uint32).ravel()
    out = np.zeros(w.size)
    lib().oracle_icdf(w.size, w.ctypes.data_as(C.POINTER(C.c_uint32)), _dp(out))
    return out


def terms_forcing(amp, t, rec):
    """terms_forcing(amp) at the times t (days) for one school record (plom_terms_forcing)."""
    t = np.ascontiguousarray(t, dtype=np.float64).ravel()
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    out = np.zeros(t.size)
    lib().oracle_terms_forcing(t.size, float(amp), _dp(t), _dp(rec), _dp(out))
    return out


def log_tab(x):
    """The psr samplers' table logarithm of positive normal doubles (plom_log_tab)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.zeros(x.size)
    lib().oracle_log_tab(x.size, _dp(x), _dp(out))
    return out


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox(c, k, o)
    return list(o)


def xoshiro(state, n):
    s = (C.c_uint32 * 4)(*state)
    o = (C.c_uint32 * n)()
    lib().oracle_xoshiro(s, n, o)
    return list(o), list(s)


def resample(logw, n=0, n_ts=1, like_min=1e-17, seed=1, stream=0, mode=CANONICAL):
    logw = np.ascontiguousarray(logw, dtype=np.float64)
    J = logw.size
    sel = np.zeros(J, dtype=np.int32)
    ln = C.c_double()
    es = C.c_double()
    L = lib()
    L.oracle_resample.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int, C.c_double, C.c_uint32,
                                  C.c_uint32, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
    r = L.oracle_resample(J, _dp(logw), n, n_ts, like_min, seed, stream, mode, _ip(sel), C.byref(ln), C.byref(es))
    return sel, ln.value, es.value, r


def resample_num(seed, stream, n):
    L = lib()
    L.oracle_resample_num.restype = C.c_uint64
    L.oracle_resample_num.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    return L.oracle_resample_num(seed, stream, n)
