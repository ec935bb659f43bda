"""Tiles per propagate block (1, 2 or 4 tiles of 256 particles in one block of 256, 512 or
1024 threads) never change a result: every tile keeps its own weight statistics, so the
summation tree (and every FP64 result) is that of the CPU oracle.  Covered: single filters
whose last block is partly empty, psr, batches tuned by plom_batch_tune or set through the
context, and served requests."""
import numpy as np
import pytest

from oracle import oracle
from iip_amd import configs, smc
from iip_amd.smc import Model

pytestmark = pytest.mark.gpu


def inputs(name, N):
    m = Model.get(name)
    times, data = m.dataset(N)
    return m, times, data, m.pack(configs.TRUTH[name]), m.x0(configs.X0_PROPS[name])


@pytest.fixture
def tctx(ctx):
    yield ctx
    ctx.set_tiles_per_block(1)


@pytest.mark.parametrize("tpb", [1, 2, 4])
@pytest.mark.parametrize("name,J,N,flags", [("sirs4", 1300, 25, 0), ("sir", 257, 20, 0), ("seir10", 2100, 8, 0),
                                            ("sirs4", 900, 12, smc.FLAG_PSR)])
def test_single_filter_equals_oracle(tctx, tpb, name, J, N, flags):
    tctx.set_tiles_per_block(tpb)
    m, times, data, pv, x0 = inputs(name, N)
    if flags & smc.FLAG_PSR:
        x0 = np.round(x0)
    print("tiles per block", tpb, name, J, N, flags, flush=True)
    g = smc.run(tctx, m, J, times, data, pv, x0, ancestry=True, final_state=True, flags=flags, seed=2, stream=tpb)
    o = oracle.smc(m.id, J, times, data, pv, x0, flags=flags, seed=2, stream=tpb, ancestry=True, final_state=True,
                   nthreads=8, psr=bool(flags & smc.FLAG_PSR))
    assert g.loglik == o["loglik"]
    np.testing.assert_array_equal(g.loglik_n, o["loglik_n"])
    np.testing.assert_array_equal(g.ess_n, o["ess_n"])
    np.testing.assert_array_equal(g.ancestry, o["ancestry"])
    np.testing.assert_array_equal(g.final_state, o["final_state"])


def test_fp32_equals_across_tiles_per_block(tctx):
    m, times, data, pv, x0 = inputs("sirs4", 20)
    ll = []
    for tpb in (1, 2, 4):
        tctx.set_tiles_per_block(tpb)
        ll.append(smc.run(tctx, m, 1500, times, data, pv, x0, precision="fp32", seed=2, stream=5).loglik)
    assert ll[0] == ll[1] == ll[2]


CASES = [("sirs4", 1300, 25, 1), ("sir", 700, 30, 2), ("sirs4", 3000, 10, 3), ("seir10", 600, 12, 4),
         ("sir", 4099, 18, 5)]


@pytest.mark.parametrize("how", ["tuned", 2, 4])
def test_batch_equals_alone_and_oracle(tctx, how):
    fs, ins = [], []
    for name, J, N, stream in CASES:
        m, times, data, pv, x0 = inputs(name, N)
        ins.append((m, times, data, pv, x0))
        fs.append(smc.Filter(tctx, m, J, times, data, pv, x0, flags=smc.FLAG_STORE_ANCESTRY, seed=9, stream=stream))
    if how != "tuned":
        tctx.set_tiles_per_block(how)
    b = smc.Batch(tctx, fs)
    if how == "tuned":
        t = b.tune(rounds=4)
        print("tuning", t, flush=True)
        assert {g["model"] for g in t["groups"]} == {"sir", "sirs4", "seir10"}
    else:
        b.configure(2)
    b.launch(use_graph=True)
    b.launch(use_graph=True)
    for f, (name, J, N, stream), (m, times, data, pv, x0) in zip(fs, CASES, ins):
        r = f.read(ancestry=True)
        o = oracle.smc(m.id, J, times, data, pv, x0, seed=9, stream=stream, ancestry=True, nthreads=8)
        assert r.loglik == o["loglik"], (name, J, N)
        np.testing.assert_array_equal(r.ancestry, o["ancestry"])
    b.close()
    for f in fs:
        f.close()


def test_served_with_four_tiles_per_block(tctx):
    tctx.set_tiles_per_block(4)
    srv = smc.Server(tctx)
    ids = []
    for name, J, N, stream in CASES:
        m, times, data, pv, x0 = inputs(name, N)
        ids.append(srv.submit(m, J, times, data, pv, x0, seed=9, stream=stream))
    got = {}
    for _ in range(5000):
        srv.step()
        for rid in srv.done():
            got[rid] = srv.result(rid)
        if len(got) == len(ids):
            break
    srv.close()
    tctx.set_tiles_per_block(1)
    for rid, (name, J, N, stream) in zip(ids, CASES):
        m, times, data, pv, x0 = inputs(name, N)
        assert got[rid].loglik == smc.run(tctx, m, J, times, data, pv, x0, seed=9, stream=stream).loglik
