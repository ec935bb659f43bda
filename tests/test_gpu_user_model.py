"""A user's model outside the library on the GPU: the no-remainder example compiled into a
model library of its own and registered with plom_model_load, against its CPU oracle
(pinned in test_user_model.py).  FP64: bit-exact log-likelihoods, ancestors, final cloud,
hat and lineage, for the sde and psr implementations; batched beside a built-in model on
two lanes from a captured graph, and served beside built-in requests, each equal to its
run alone."""
import numpy as np
import pytest

from iip_amd import configs, smc
from iip_amd.smc import Model
from oracle import oracle
import user_model as U

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model():
    return U.load()


def inputs(m, N):
    times, data = m.dataset(N)
    return times, data, m.pack(U.TRUTH), m.x0(U.X0_PROPS), m.covariates()


@pytest.mark.parametrize("J,N,flags", [(3000, 60, 0), (1000, 30, smc.FLAG_PSR), (257, 20, 0)])
def test_filter_equals_oracle(ctx, model, J, N, flags):
    m = model
    times, data, pv, x0, cov = inputs(m, N)
    psr = bool(flags & smc.FLAG_PSR)
    if psr:
        x0 = np.round(x0)
    print("user model", J, N, flags, flush=True)
    g = smc.run(ctx, m, J, times, data, pv, x0, ancestry=True, final_state=True, flags=flags, seed=4, stream=8,
                cov=cov)
    o = oracle.smc(U.NAME, J, times, data, pv, x0, flags=flags, seed=4, stream=8, ancestry=True, final_state=True,
                   nthreads=8, cov=cov, psr=psr)
    assert g.loglik == o["loglik"]
    np.testing.assert_array_equal(g.loglik_n, o["loglik_n"])
    np.testing.assert_array_equal(g.ess_n, o["ess_n"])
    np.testing.assert_array_equal(g.ancestry, o["ancestry"])
    np.testing.assert_array_equal(g.final_state, o["final_state"])


def test_hat_and_lineage_equal_oracle(ctx, model):
    m = model
    times, data, pv, x0, cov = inputs(m, 40)
    print("user model hat + lineage", flush=True)
    f = smc.Filter(ctx, m, 1500, times, data, pv, x0, flags=smc.FLAG_HAT | smc.FLAG_LINEAGE, seed=6, stream=1, cov=cov)
    f.run()
    g = f.read(hat=True, lineage=True)
    f.close()
    o = oracle.smc(U.NAME, 1500, times, data, pv, x0, seed=6, stream=1, hat=True, lineage=True, nthreads=8, cov=cov)
    np.testing.assert_array_equal(g.hat, o["hat"])
    np.testing.assert_array_equal(g.pred_res, o["pred_res"])
    np.testing.assert_array_equal(g.lineage, o["lineage"])
    np.testing.assert_array_equal(g.lineage_slots, o["lineage_slots"])


def test_batched_and_served_beside_builtin_models(ctx, model):
    m = model
    times, data, pv, x0, cov = inputs(m, 50)
    nm = Model.get("noise")
    nt, nd = nm.dataset(40)
    npv, nx0, ncov = nm.pack(configs.TRUTH["noise"]), nm.x0(configs.X0_PROPS["noise"]), nm.covariates()
    fs = [smc.Filter(ctx, m, 2000, times, data, pv, x0, seed=3, stream=1, cov=cov),
          smc.Filter(ctx, nm, 1200, nt, nd, npv, nx0, seed=3, stream=2, cov=ncov),
          smc.Filter(ctx, m, 700, times[:31], data[:30], pv, x0, seed=3, stream=3, cov=cov)]
    b = smc.Batch(ctx, fs)
    b.configure(2)
    print("user model batched", flush=True)
    b.launch(use_graph=True)
    b.launch(use_graph=True)
    got = [f.read().loglik for f in fs]
    b.close()
    for f in fs:
        f.close()
    alone = [smc.run(ctx, m, 2000, times, data, pv, x0, seed=3, stream=1, cov=cov).loglik,
             smc.run(ctx, nm, 1200, nt, nd, npv, nx0, seed=3, stream=2, cov=ncov).loglik,
             smc.run(ctx, m, 700, times[:31], data[:30], pv, x0, seed=3, stream=3, cov=cov).loglik]
    assert got == alone
    assert alone[0] == oracle.smc(U.NAME, 2000, times, data, pv, x0, seed=3, stream=1, nthreads=8, cov=cov)["loglik"]
    print("user model served", flush=True)
    srv = smc.Server(ctx)
    sm = Model.get("sir")
    st, sd = sm.dataset(30)
    ids = [srv.submit(m, 2000, times, data, pv, x0, seed=3, stream=1, cov=cov),
           srv.submit(sm, 900, st, sd, sm.pack(configs.TRUTH["sir"]), sm.x0(configs.X0_PROPS["sir"]), seed=3, stream=4),
           srv.submit(m, 700, times[:31], data[:30], pv, x0, seed=3, stream=3, cov=cov)]
    res = {}
    for _ in range(5000):
        srv.step()
        for rid in srv.done():
            res[rid] = srv.result(rid)
        if len(res) == len(ids):
            break
    srv.close()
    assert res[ids[0]].loglik == alone[0]
    assert res[ids[2]].loglik == alone[2]
