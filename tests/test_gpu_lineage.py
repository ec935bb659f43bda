"""Lineage sampling (pmcmc --full) on the GPU against the CPU oracle.

The device draws the final particle from the last interval's cumulative weights, walks
its ancestry back and regenerates its path with the counter-based noise; the result must
equal the oracle's (which is pinned against brute-force stored states) bit for bit in
FP64, alone, in a batch (any number of lanes) and served."""
import numpy as np
import pytest

from oracle import oracle
from iip_amd import configs, smc
from iip_amd.smc import Model

pytestmark = pytest.mark.gpu


def setup(name, N):
    m = Model.get(name)
    times, data = m.dataset(N)
    return m, times, data, m.pack(configs.TRUTH[name]), m.x0(configs.X0_PROPS[name])


@pytest.mark.parametrize("name,J,N,flags", [("sir", 1000, 52, 0), ("sirs4", 3000, 40, 0), ("seir10", 700, 20, 0),
                                            ("sir", 257, 30, smc.FLAG_NO_FILTER), ("sirs4", 1, 10, 0)])
def test_lineage_equals_oracle(ctx, name, J, N, flags):
    m, times, data, pv, x0 = setup(name, N)
    print("lineage", name, J, N, flags, flush=True)
    g = smc.run(ctx, m, J, times, data, pv, x0, lineage=True, flags=flags, seed=3, stream=11)
    o = oracle.smc(m.id, J, times, data, pv, x0, flags=flags, seed=3, stream=11, lineage=True, nthreads=8)
    np.testing.assert_array_equal(g.lineage_slots, o["lineage_slots"])
    np.testing.assert_array_equal(g.lineage, o["lineage"])
    np.testing.assert_array_equal(g.lineage_obs, o["lineage_obs"])
    assert g.loglik == o["loglik"]


def test_lineage_with_missing_last_row(ctx):
    """No data at the last interval: the final particle is drawn with equal weights."""
    m, times, data, pv, x0 = setup("sirs4", 25)
    data = data.copy()
    data[-1, :] = np.nan
    g = smc.run(ctx, m, 2000, times, data, pv, x0, lineage=True, seed=2, stream=5)
    o = oracle.smc(m.id, 2000, times, data, pv, x0, seed=2, stream=5, lineage=True, nthreads=8)
    np.testing.assert_array_equal(g.lineage_slots, o["lineage_slots"])
    np.testing.assert_array_equal(g.lineage, o["lineage"])


@pytest.mark.parametrize("lanes", [1, 3])
def test_lineage_batched_equals_alone(ctx, lanes):
    cases = [("sir", 900, 30, 1, "fp64"), ("sirs4", 1500, 22, 2, "fp64"), ("sirs4", 800, 35, 3, "fp32"),
             ("seir10", 500, 12, 4, "fp64")]
    fs, alone = [], []
    for name, J, N, stream, prec in cases:
        m, times, data, pv, x0 = setup(name, N)
        fs.append(smc.Filter(ctx, m, J, times, data, pv, x0, precision=prec, flags=smc.FLAG_LINEAGE, seed=9,
                             stream=stream))
        alone.append(smc.run(ctx, m, J, times, data, pv, x0, precision=prec, lineage=True, seed=9, stream=stream))
    b = smc.Batch(ctx, fs)
    b.configure(lanes)
    b.launch(use_graph=True)
    for f, a in zip(fs, alone):
        r = f.read(lineage=True)
        np.testing.assert_array_equal(r.lineage_slots, a.lineage_slots)
        np.testing.assert_array_equal(r.lineage, a.lineage)
        np.testing.assert_array_equal(r.lineage_obs, a.lineage_obs)
    b.close()
    for f in fs:
        f.close()


def test_lineage_served_equals_alone(ctx):
    srv = smc.Server(ctx)
    reqs = [("sirs4", 1200, 30, 31, True), ("sir", 3000, 45, 32, False), ("sir", 600, 20, 33, True)]
    ids = []
    for name, J, N, stream, lin in reqs:
        m, times, data, pv, x0 = setup(name, N)
        ids.append(srv.submit(m, J, times, data, pv, x0, seed=9, stream=stream, lineage=lin))
    got = {}
    for _ in range(100000):
        srv.step()
        for rid in srv.done():
            if rid not in got:
                got[rid] = srv.result(rid)
        if len(got) == len(ids):
            break
    for rid, (name, J, N, stream, lin) in zip(ids, reqs):
        m, times, data, pv, x0 = setup(name, N)
        a = smc.run(ctx, m, J, times, data, pv, x0, lineage=lin, seed=9, stream=stream)
        assert got[rid].loglik == a.loglik
        if lin:
            np.testing.assert_array_equal(got[rid].lineage, a.lineage)
            np.testing.assert_array_equal(got[rid].lineage_slots, a.lineage_slots)
        else:
            assert got[rid].lineage is None
    srv.close()
