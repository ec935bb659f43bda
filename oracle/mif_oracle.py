"""CPU restatement of the MIF driver (iterated filtering, Ionides et al. 2006).
TEST INFRASTRUCTURE: imported only by tests/, as the checker of the `mif` command, never
as the thing measured or shipped.

It restates, in plain Python, the iteration loop of the reference
(model_builder/C/mif/mif.c:21-247, mif/methods.c):

* fill_theta_bart_and_Vt_mif (methods.c:114-138): every row of the filtered means starts
  at the current theta, Vt[0] = (b FREEZE)^2 var_kk, FREEZE = a^(m-1);
* the particles' starting parameters (ran_proposal, proposal.c:71-84): mean +
  (b FREEZE) sqrt(var_kk) z for every estimated component, or with a covariance in theta
  (ran_proposal_chol, proposal.c:40-69) mean + (b FREEZE) Lv z, Lv the Cholesky factor of
  the estimated components' covariance; redrawn until the initial proportions fit
  (proposal.c:87-114), then followers copy their leaders;
* one pass with per-particle parameters, given as a callable (the CPU filter oracle,
  oracle.mif_pass, in the tests);
* per observation with data: the filtered mean, and the variance plus delta_t FREEZE^2
  var_kk (mean_var_theta_theoretical_mif, methods.c:146-205), for the estimated process /
  observation components (all estimated components when the mif file is written);
* the fixed-lag value of the estimated initial conditions: their filtered mean at the
  last observation up to L;
* update_theta_best_stable_mif (methods.c:395-418) for m <= SWITCH, else
  update_theta_best_king_mif (methods.c:426-461);
* trace rows (index, f_inv(theta), log-likelihood; row 0 has no log-likelihood).

Random numbers are the `mif` command's: the starting parameters from Philox-4x32-10 keyed
(seed, 0xfffffffe) with counter (j, m, attempt, block), four Box-Muller normals per
block, in the order of the estimated components.
"""
import math

import numpy as np

from . import oracle

START_STREAM = 0xFFFFFFFE
TR = {0: "identity", 1: "log", 2: "logit", 3: "logit_ab"}


def f_inv(tr, x, a, b):
    if tr == 1:
        return math.exp(x)
    if tr == 2:
        return 1.0 / (1.0 + math.exp(-x)) if x > 0 else math.exp(x) / (1.0 + math.exp(x))
    if tr == 3:
        if a == b:
            return x
        return (b * math.exp(x) + a) / (1.0 + math.exp(x)) if x < 0 else (b + a * math.exp(-x)) / (1.0 + math.exp(-x))
    return x


def start_normals(seed, j, m, attempt, nblocks):
    z = []
    for blk in range(nblocks):
        w = oracle.philox([j, m, attempt, blk], [seed & 0xFFFFFFFF, START_STREAM])
        p = oracle.bm_pairs(np.array([[w[0], w[1]], [w[2], w[3]]], dtype=np.uint32))
        z += [p[0, 0], p[0, 1], p[1, 0], p[1, 1]]
    return z


def run(dry, pass_fn, x0_of, constrain, M, J, seed, a=0.975, b=2.0, switch=5, print_traj=False, follow=None,
        keep=None):
    """dry: the `mif --dry_run` inputs (theta, var, estimated, sv, tr, min, max, delta, N,
    lag_row); pass_fn(theta [n, J], x0 [S, J], mut [N], seed, stream) -> dict with loglik,
    mif_mean, mif_var ([(N+1), n]); x0_of(theta) -> initial state [S]; constrain(theta) ->
    number of violated initial-condition constraints (assigns a theta remainder in place);
    keep: the components written in the trace (default all)."""
    n = len(dry["theta"])
    N = dry["N"]
    var = np.array(dry["var"], dtype=float).reshape(n, n)
    est, sv = dry["estimated"], dry["sv"]
    tr, mn, mx = dry["tr"], dry["min"], dry["max"]
    delta = dry["delta"]
    lag_row = dry["lag_row"]
    keep = list(range(n)) if keep is None else keep
    mean = list(dry["theta"])
    est_list = [o for o in range(n) if est[o]]
    nblocks = (len(est_list) + 3) // 4
    ne = len(est_list)
    Lv = np.array(dry.get("chol_est") or [], dtype=float)
    Lv = Lv.reshape(ne, ne) if Lv.size else None
    rows = [r for r in range(N) if not math.isnan(delta[r]) and r in dry["data_rows"]]
    trace = [(0, [f_inv(tr[o], mean[o], mn[o], mx[o]) for o in keep], float("nan"))]
    for m in range(1, M + 1):
        FREEZE = a ** (m - 1)
        mut = np.array([FREEZE * math.sqrt(delta[r]) for r in range(N)])
        sd_fac = b * FREEZE
        bart = np.array([list(mean)] * (N + 1))
        vt = np.zeros((N + 1, n))
        for o in range(n):
            vt[0, o] = (sd_fac * sd_fac) * var[o, o]
        theta = np.zeros((n, J))
        x0 = None
        for j in range(J):
            attempt = 0
            while True:
                th = list(mean)
                z = start_normals(seed, j, m, attempt, nblocks)
                for q, o in enumerate(est_list):
                    if Lv is None:
                        th[o] = mean[o] + sd_fac * math.sqrt(var[o, o]) * z[q]
                    else:
                        acc = 0.0
                        for p_ in range(q + 1):
                            acc = acc + Lv[q, p_] * z[p_]
                        th[o] = mean[o] + sd_fac * acc
                if constrain(th) == 0:
                    break
                attempt += 1
            if follow:
                follow(th)
            theta[:, j] = th
            xs = x0_of(th)
            if x0 is None:
                x0 = np.zeros((len(xs), J))
            x0[:, j] = xs
        r = pass_fn(theta, x0, mut, seed, m)
        for row in rows:
            var_fac = delta[row] * (FREEZE * FREEZE)
            for o in range(n):
                if not est[o] or (sv[o] and not print_traj):
                    continue
                v = r["mif_var"][row + 1, o]
                if v < 0.0 or math.isinf(v) or math.isnan(v):
                    v = 0.0
                bart[row + 1, o] = r["mif_mean"][row + 1, o]
                vt[row + 1, o] = v + var_fac * var[o, o]
        if lag_row >= 0:
            for o in range(n):
                if est[o] and sv[o]:
                    mean[o] = r["mif_mean"][lag_row + 1, o]
        for o in range(n):
            if not est[o] or sv[o]:
                continue
            if m <= switch:
                tmp = 0.0
                for row in rows:
                    tmp += bart[row + 1, o]
                mean[o] = tmp / len(rows)
            else:
                nnp1 = rows[0]
                tmp = (bart[nnp1 + 1, o] - bart[0, o]) / vt[nnp1 + 1, o]
                for q in range(1, len(rows)):
                    nn, nnp1 = rows[q - 1], rows[q]
                    tmp += (bart[nnp1 + 1, o] - bart[nn + 1, o]) / vt[nnp1 + 1, o]
                mean[o] = mean[o] + ((dry["times0"] + b * b) * FREEZE * FREEZE * var[o, o] * tmp)
        trace.append((m, [f_inv(tr[o], mean[o], mn[o], mx[o]) for o in keep], r["loglik"]))
    return {"trace": trace, "mean": mean}
