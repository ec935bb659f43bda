"""CPU restatement of the pMCMC driver (particle marginal Metropolis-Hastings).
TEST INFRASTRUCTURE: imported only by tests/, as the checker of the `pmcmc` command,
never as the thing measured or shipped.

It restates, in plain Python, the reference's chain (model_builder/C/pmcmc/pmcmc.c:40-266):

* increment_iteration_counters (core/mcmc_util.c:163-182);
* get_var_and_sd_fac (mcmc_util.c:88-147): epsilon adaptation
  eps <- eps * exp(a^(m-1) (ar - 0.234)) while m > E and m ar < S, else 1, capped at
  eps_max; sd_fac = eps 2.38 / sqrt(n); the empirical covariance once m ar >= S;
  sequential updates shuffle the estimated components at the start of every cycle;
* proposals (proposal.c:40-84, mvn.c:105-150, mcmc_util.c:186-197), redrawn until the
  initial-condition constraint holds;
* metropolis_hastings (metropolis_hastings_prior.c:24-50) with log_prob_proposal
  (proposal.c:120-188: the transformed-scale density divided by the derivative of the
  inverse transform) and log_prob_prior (metropolis_hastings_prior.c:60-92);
* compute_acceptance_rates (mcmc_util.c:156-166) and eval_var_emp (mvn.c:196-233).

The random numbers are the `pmcmc` command's: Philox-4x32-10 keyed (seed, 0xffffffff),
counter (m, 1, attempt, block) for the proposal normals (Box-Muller of word pairs, four
normals per block), (m, 2, 0, 0) for the acceptance uniform (w + 1/2) 2^-32 and
(m, 3, block, 0) for the Fisher-Yates shuffle (j = floor(w (i + 1) / 2^32)).  The
likelihood is a callable, so the chain logic is pinned against closed-form targets
(tests/test_pmcmc_oracle.py) independently of the particle filter.
"""
import math

import numpy as np

from . import oracle

CHAIN_STREAM = 0xFFFFFFFF
PURPOSE_PROPOSAL, PURPOSE_ACCEPT, PURPOSE_SHUFFLE = 1, 2, 3


# ---------------------------------------------------------------- transforms (transform.c)
def f_inv(tr, x, a, b):
    if tr == "log":
        return math.exp(x)
    if tr == "logit":
        return 1.0 / (1.0 + math.exp(-x)) if x > 0 else math.exp(x) / (1.0 + math.exp(x))
    if tr == "logit_ab":
        if a == b:
            return x
        return (b * math.exp(x) + a) / (1.0 + math.exp(x)) if x < 0 else (b + a * math.exp(-x)) / (1.0 + math.exp(-x))
    return x


def f_inv_der(tr, x, a, b):
    if tr == "log":
        return math.exp(x)
    if tr == "logit":
        e = math.exp(-x if x > 0 else x)
        return e / ((1.0 + e) * (1.0 + e))
    if tr == "logit_ab":
        if a == b:
            return x
        e = math.exp(-x if x > 0 else x)
        return (b - a) * e / ((e + 1.0) * (e + 1.0))
    return x


def gauss_pdf(u, sd):
    return math.exp(-u * u / (2.0 * sd * sd)) / (math.sqrt(2.0 * math.pi) * sd)


def prior_pdf(kind, x, a, b):
    if kind == "normal":
        return gauss_pdf(x - 0.5 * (b + a), (b - a) / 4.0)
    if kind == "pseudo_uniform":
        d = 0.01 * (b - a)
        h = 1.0 / (b - a - 2.0 / 3.0 * d)
        c2 = -h / (d * d)
        c1 = -2.0 * (a + d) * c2
        c0 = h + c2 * (a + d) * (a + d)
        if x <= a or x >= b:
            return 0.0
        if x < a + d:
            return (c2 * x + c1) * x + c0
        if x < b - d:
            return h
        y = x - (b - a - 2.0 * d)
        return (c2 * y + c1) * y + c0
    return 1.0 / (b - a) if a <= x < b else 0.0


# ---------------------------------------------------------------- random numbers
def words(seed, c0, c1, c2, c3):
    return oracle.philox([c0 & 0xFFFFFFFF, c1, c2, c3], [seed & 0xFFFFFFFF, CHAIN_STREAM])


def normals(seed, m, attempt, n):
    z = []
    b = 0
    while len(z) < n:
        w = words(seed, m, PURPOSE_PROPOSAL, attempt, b)
        pairs = oracle.bm_pairs(np.array([[w[0], w[1]], [w[2], w[3]]], dtype=np.uint32))
        z += [pairs[0, 0], pairs[0, 1], pairs[1, 0], pairs[1, 1]]
        b += 1
    return z[:n]


# ---------------------------------------------------------------- the chain
def run(problem, loglik, M, seed, full=False, m_eps=50, a=0.999, eps_max=50.0, alpha=0.02, m_switch=-1, smooth=False,
        like_min=1e-17, jacobian=True):
    """One chain of M iterations.

    problem: {"theta": [...] transformed initial values, "var": n*n initial covariance,
    "estimated": [offsets], "components": [{"tr", "min", "max", "prior", "remainder"}],
    "constrain": callable(theta) -> number of violated initial-condition constraints
    (assigns the theta remainder in place), "follow": callable(theta) or None}.
    loglik(theta, m) -> the filter's log-likelihood at iteration m.
    jacobian=False drops the change-of-variable factor (only for tests that show the
    restatement can tell the difference).
    Returns the trace rows (index, natural values of the non-remainder components,
    log-likelihood), the empirical covariance, the acceptance count and, per iteration,
    the proposal, the likelihood and whether it was accepted."""
    comps = problem["components"]
    n = len(problem["theta"])
    var0 = np.array(problem["var"], dtype=float).reshape(n, n)
    tbe0 = list(problem["estimated"])
    nt = len(tbe0)
    est = [False] * n
    for o in tbe0:
        est[o] = True
    constrain, follow = problem["constrain"], problem.get("follow")
    if m_switch < 0:
        m_switch = 5 * nt * nt
    mean = list(problem["theta"])
    proposed = list(mean)
    mean_s = [0.0] * n
    var_s = np.zeros((n, n))
    tbe = list(tbe0)
    ar = [1.0] * n
    global_ar = sm_global_ar = 0.0
    eps = 1.0
    has_cycled, m_full, cycle_id = 1, 0, nt - 1
    keep = [o for o in range(n) if not comps[o].get("remainder")]

    def natural(th):
        return [f_inv(comps[o]["tr"], th[o], comps[o]["min"], comps[o]["max"]) for o in keep]

    def log_prior(th):
        err, lp = False, 0.0
        for o in range(n):
            if not est[o]:
                continue
            c = comps[o]
            p = prior_pdf(c["prior"], f_inv(c["tr"], th[o], c["min"], c["max"]), c["min"], c["max"])
            if math.isnan(p) or math.isinf(p) or p < 0.0:
                err, p = True, like_min
            elif p <= like_min:
                p = like_min
            lp += math.log(p)
        return err, lp

    def cholesky(var, s):
        """lower factor of s var[tbe0, tbe0] by the Cholesky-Banachiewicz recursion, or None"""
        L = [[0.0] * nt for _ in range(nt)]
        for i in range(nt):
            for j in range(i + 1):
                v = s * var[tbe0[i], tbe0[j]]
                for k in range(j):
                    v -= L[i][k] * L[j][k]
                if i == j:
                    if not v > 0.0:
                        return None
                    L[i][i] = math.sqrt(v)
                else:
                    L[i][j] = v / L[j][j]
        return L

    def log_q(x, y, var, sd_fac, L):
        """log density of proposing x from y, on the natural scale"""
        lp = 0.0
        if full:
            # log N(x; y, L L^T) by forward substitution
            yv, q, ld = [0.0] * nt, 0.0, 0.0
            for i in range(nt):
                v = x[tbe0[i]] - y[tbe0[i]]
                for k in range(i):
                    v -= L[i][k] * yv[k]
                yv[i] = v / L[i][i]
                q += yv[i] * yv[i]
                ld += math.log(L[i][i])
            lp = -0.5 * nt * math.log(2.0 * math.pi) - ld - 0.5 * q
        for o in range(n):
            if not est[o]:
                continue
            c = comps[o]
            der = f_inv_der(c["tr"], x[o], c["min"], c["max"]) if jacobian else 1.0
            if not full:
                p = gauss_pdf(x[o] - y[o], sd_fac * math.sqrt(var[o, o])) / der
                if math.isnan(p) or math.isinf(p) or p < 0.0:
                    return True, 0.0
                lp += math.log(p)
            else:
                if math.isnan(der) or math.isinf(der) or der <= 0.0:
                    return True, 0.0
                lp -= math.log(der)
        if math.isnan(lp) or math.isinf(lp):
            return True, 0.0
        return False, lp

    out = {"trace": [], "means": [], "proposals": [], "updated": [], "loglik": [], "accepted": [], "alpha": [],
           "sd_fac": []}
    L_new = loglik(list(mean), 0)
    L_prev = L_new
    out["trace"].append((0, natural(mean), L_new))
    out["means"].append(list(mean))
    out["loglik"].append(L_new)
    accept = 0
    for m in range(1, M):
        if full:
            has_cycled, m_full, cycle_id = 1, m_full + 1, nt
        elif cycle_id >= nt - 1:
            has_cycled, m_full, cycle_id = 1, m_full + 1, 0
        else:
            has_cycled, cycle_id = 0, cycle_id + 1
        var, sd_fac, L = var0, 1.0, None
        if full:
            if m > m_eps and m * global_ar < m_switch:
                eps *= math.exp(a ** (m - 1) * ((sm_global_ar if smooth else global_ar) - 0.234))
            else:
                eps = 1.0
            eps = min(eps, eps_max)
            sd_fac = eps * 2.38 / math.sqrt(nt)
            if m * global_ar >= m_switch:
                var = var_s
            L = cholesky(var, sd_fac * sd_fac)
            if L is None:
                var = var0
                L = cholesky(var, sd_fac * sd_fac)
        elif nt > 0 and has_cycled:
            for i in range(nt - 1, 0, -1):
                w = words(seed, m, PURPOSE_SHUFFLE, i // 4, 0)
                j = (w[i % 4] * (i + 1)) >> 32
                tbe[i], tbe[j] = tbe[j], tbe[i]
        attempt = 0
        while True:
            if full:
                z = normals(seed, m, attempt, nt)
                proposed = list(mean)
                for i, o in enumerate(tbe0):
                    v = 0.0
                    for q in range(i + 1):
                        v += L[i][q] * z[q]
                    proposed[o] = mean[o] + v
            elif nt > 0:
                z = normals(seed, m, attempt, 1)
                o = tbe[cycle_id]
                proposed[o] = mean[o] + sd_fac * math.sqrt(var[o, o]) * z[0]
                if attempt == 0:
                    out["updated"].append(o)
            if constrain(proposed) == 0:
                break
            attempt += 1
        if follow:
            follow(proposed)
        L_new = loglik(list(proposed), m)
        out["proposals"].append(list(proposed))
        out["sd_fac"].append(sd_fac)
        out["loglik"].append(L_new)
        e1, lqn = log_q(proposed, mean, var, sd_fac, L)
        e2, lqp = log_q(mean, proposed, var, sd_fac, L)
        e3, lpn = log_prior(proposed)
        e4, lpp = log_prior(mean)
        acc, al = False, 0.0
        if not (e1 or e2 or e3 or e4):
            al = math.exp(L_new - L_prev + lqp - lqn + lpn - lpp)
            u = (words(seed, m, PURPOSE_ACCEPT, 0, 0)[0] + 0.5) * 2.0 ** -32
            acc = u < al
        out["accepted"].append(acc)
        out["alpha"].append(al)
        if acc:
            L_prev = L_new
            mean = list(proposed)
            accept += 1
        elif not full and nt > 0:
            o = tbe[cycle_id]
            proposed[o] = mean[o]
        x = 1.0 if acc else 0.0
        if not full and nt > 0:
            o = tbe[cycle_id]
            ar[o] += (x - ar[o]) / m_full
        sm_global_ar = (1.0 - alpha) * sm_global_ar + alpha * x
        global_ar += (x - global_ar) / m
        if cycle_id >= nt - 1:
            mf = float(m_full)
            for i in range(n):
                for k in range(i + 1):
                    val = ((mf - 1.0) / mf) * (mean[i] - mean_s[i]) * (mean[k] - mean_s[k])
                    var_s[i, k] = ((mf - 1.0) * var_s[i, k] + val) / mf
            for i in range(n):
                for k in range(i):
                    var_s[k, i] = var_s[i, k]
            for i in range(n):
                mean_s[i] += (mean[i] - mean_s[i]) / mf
            out["trace"].append((m_full, natural(mean), L_prev))
            out["means"].append(list(mean))
    out["var_sampling"] = var_s
    out["accept"] = accept
    out["mean"] = mean
    out["global_ar"] = global_ar
    return out
