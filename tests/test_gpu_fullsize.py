"""Oracle parity at the workloads' full sizes (and at the particle counts that exercise
every shape of the chunked tile trees).

The filter loop is smc.c:379-467 (weight :27-83, systematic_sampling :90-112, resample_X
:142-161); the CPU oracle restates it and is pinned in test_oracle.py.  Bar: FP64 results
bit-identical to the oracle (log-likelihood, per-interval log-likelihoods, ESS, every
ancestor, final cloud, lineage); FP32 propagation: mean log-likelihood over replicates
within 3 Monte Carlo standard errors of FP64.

Tile trees: a filter of J particles has T = ceil(J / 256) tiles; the reductions over tiles
(finalize_weights, tile_tree, k_mif_reduce, k_hat_select) run in chunks of 256 tiles,
then a pairwise tree over the chunks zero padded to a power of two.  J = 150000 gives 3
chunks (padded to 4), J = 300000 gives 5 (padded to 8), J = 1e6 gives 16 (3907 tiles).
"""
import os

import numpy as np
import pytest

from oracle import oracle
from iip_amd import configs, smc
from iip_amd.smc import Model

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def inputs(name, N, params=None):
    m = Model.get(name)
    times, data = m.dataset(N)
    return m, times, data, m.pack(params or configs.TRUTH[name]), m.x0(configs.X0_PROPS[name])


def assert_same(g, o, ancestry=True, final_state=True):
    assert g.loglik == o["loglik"]
    assert g.n_all_fail == o["n_all_fail"]
    np.testing.assert_array_equal(g.loglik_n, o["loglik_n"])
    np.testing.assert_array_equal(g.ess_n, o["ess_n"])
    if ancestry:
        np.testing.assert_array_equal(g.ancestry, o["ancestry"])
    if final_state:
        np.testing.assert_array_equal(g.final_state, o["final_state"])


def test_sirs4_j100000_n520_fp64(ctx):
    """The single large filter: SIRS 4 x 5e5, J = 1e5, N = 520, FP64; every ancestor of
    every interval and the final cloud."""
    J, N = 100000, 520
    m, times, data, pv, x0 = inputs("sirs4", N)
    print("sirs4 J", J, "N", N, flush=True)
    g = smc.run(ctx, m, J, times, data, pv, x0, ancestry=True, final_state=True, seed=configs.FILTER_SEED, stream=0)
    o = oracle.smc(m.id, J, times, data, pv, x0, seed=configs.FILTER_SEED, stream=0, ancestry=True,
                   final_state=True, nthreads=THREADS)
    assert_same(g, o)


def pmcmc_sample(k=8):
    """k filters of the 256-filter pMCMC batch at their own J and N: evenly spaced in the
    order of J * N, from the smallest to the largest."""
    draws = configs.pmcmc_batch(256)
    order = sorted(range(len(draws)), key=lambda i: draws[i]["J"] * draws[i]["N"])
    pick = [order[int(round(q * (len(order) - 1) / (k - 1)))] for q in range(k)]
    return [draws[i] for i in pick]


def test_pmcmc_batch_draws_batched_graph_two_lanes(ctx):
    """8 filters of the pMCMC batch at their real sizes (J up to 2e4, N up to 520), run as
    one batch on 2 lanes from a captured hipGraph, replayed: each equals the oracle."""
    sample = pmcmc_sample()
    fs, ins = [], []
    for d in sample:
        m, times, data, pv, x0 = inputs(d["model"], d["N"], d["params"])
        ins.append((m, times, data, pv, x0))
        fs.append(smc.Filter(ctx, m, d["J"], times, data, pv, x0, flags=smc.FLAG_STORE_ANCESTRY | smc.FLAG_FINAL_STATE,
                             seed=configs.FILTER_SEED, stream=d["stream"]))
    b = smc.Batch(ctx, fs)
    b.configure(2)
    print("pmcmc batch sample", [(d["J"], d["N"]) for d in sample], flush=True)
    b.launch(use_graph=True)       # capture and first replay
    b.launch(use_graph=True)       # second replay
    for f, d, (m, times, data, pv, x0) in zip(fs, sample, ins):
        g = f.read(ancestry=True, final_state=True)
        o = oracle.smc(m.id, d["J"], times, data, pv, x0, seed=configs.FILTER_SEED, stream=d["stream"],
                       ancestry=True, final_state=True, nthreads=THREADS)
        assert_same(g, o)
    b.close()
    for f in fs:
        f.close()


def test_seir10_j1000000_ancestry_lineage(ctx):
    """The memory-heavy model at its particle count: SEIR 10 x 1e6 (S = 40), J = 1e6
    (3907 tiles, 16 chunks), 8 intervals, ancestry and lineage kept."""
    J, N = 1000000, 8
    m, times, data, pv, x0 = inputs("seir10", N)
    print("seir10 J", J, "N", N, flush=True)
    g = smc.run(ctx, m, J, times, data, pv, x0, ancestry=True, final_state=True, lineage=True,
                seed=configs.FILTER_SEED, stream=999)
    o = oracle.smc(m.id, J, times, data, pv, x0, seed=configs.FILTER_SEED, stream=999, ancestry=True,
                   final_state=True, lineage=True, nthreads=THREADS)
    assert_same(g, o)
    np.testing.assert_array_equal(g.lineage_slots, o["lineage_slots"])
    np.testing.assert_array_equal(g.lineage, o["lineage"])


def test_fp32_mean_loglik_at_the_large_filter_shape(ctx):
    """FP32 propagation against FP64 at J = 1e5, N = 520 (SIRS 4 x 5e5): 30 replicates each
    (their own random streams); the means agree within 3 combined standard errors."""
    R, J, N = 30, 100000, 520
    m, times, data, pv, x0 = inputs("sirs4", N)
    ll = {}
    for prec in ("fp64", "fp32"):
        print("replicates", prec, flush=True)
        fs = [smc.Filter(ctx, m, J, times, data, pv, x0, precision=prec, seed=11, stream=200 + r) for r in range(R)]
        smc.batch_run(ctx, fs, use_graph=True)
        ll[prec] = np.array([f.read().loglik for f in fs])
        for f in fs:
            f.close()
    a, b = ll["fp64"], ll["fp32"]
    se = np.sqrt(a.var(ddof=1) / R + b.var(ddof=1) / R)
    print("fp64 mean %.4f sd %.4f | fp32 mean %.4f sd %.4f | se %.4f" % (a.mean(), a.std(ddof=1), b.mean(),
                                                                          b.std(ddof=1), se))
    assert abs(a.mean() - b.mean()) < 3 * se


@pytest.mark.parametrize("J", [150000, 300000, 1000000])
def test_tile_tree_chunks_plain(ctx, J):
    """3, 5 and 16 chunks of tiles in the weight reduction and the tile prefix."""
    N = 4
    m, times, data, pv, x0 = inputs("sir", N)
    print("plain sir J", J, flush=True)
    g = smc.run(ctx, m, J, times, data, pv, x0, ancestry=True, final_state=True, seed=3, stream=J)
    o = oracle.smc(m.id, J, times, data, pv, x0, seed=3, stream=J, ancestry=True, final_state=True, nthreads=THREADS)
    assert_same(g, o)


@pytest.mark.parametrize("J", [150000, 1000000])
def test_tile_tree_chunks_hat(ctx, J):
    """The hat means (tree over the tiles' value x weight sums) and bounds, and the
    prediction residuals, over 3 and 16 chunks of tiles."""
    N = 3
    m, times, data, pv, x0 = inputs("sir", N)
    print("hat sir J", J, flush=True)
    f = smc.Filter(ctx, m, J, times, data, pv, x0, flags=smc.FLAG_HAT, seed=3, stream=4)
    f.run()
    g = f.read(hat=True)
    f.close()
    o = oracle.smc(m.id, J, times, data, pv, x0, seed=3, stream=4, hat=True, nthreads=THREADS)
    assert g.loglik == o["loglik"]
    np.testing.assert_array_equal(g.hat, o["hat"])
    np.testing.assert_array_equal(g.pred_res, o["pred_res"])


@pytest.mark.parametrize("J", [150000, 1000000])
def test_tile_tree_chunks_mif(ctx, J):
    """The theta moments of an iterated-filtering pass (the weight sums and the
    (count, mean, M2) merges over the tiles) over 3 and 16 chunks of tiles."""
    from test_gpu_mif import mif_setup, check_equal
    N = 3
    m, times, data, pv, mp, theta0, x0, mut = mif_setup("sir", N, J, 13)
    print("mif sir J", J, flush=True)
    f = smc.Filter(ctx, m, J, times, data, pv, x0[:, 0], flags=smc.FLAG_MIF, mif=mp, seed=1, stream=0)
    f.mif_start(theta0, x0, mut, seed=6, stream=5)
    f.run()
    g = f.read(mif=True)
    f.close()
    o = oracle.mif_pass(m.id, J, times, data, pv, mp, theta0, x0, mut, seed=6, stream=5, nthreads=THREADS)
    check_equal(g, o)
