"""The served stream: continuous batching with admission by free HBM.

Every request's result equals the same request run alone (bit for bit in FP64),
whatever arrives before, beside or after it; admission respects the memory budget."""
import numpy as np
import pytest

from iip_amd import configs, smc
from iip_amd.smc import Model

pytestmark = pytest.mark.gpu

REQS = [("sir", 700, 52, 11), ("sirs4", 2000, 30, 12), ("sir", 5000, 12, 13), ("sirs4", 300, 45, 14),
        ("seir10", 800, 20, 15), ("sir", 257, 52, 16), ("sirs4", 9000, 8, 17), ("sir", 1500, 25, 18)]


def inputs(name, N, scale=1.0):
    m = Model.get(name)
    times, data = m.dataset(N)
    p = dict(configs.TRUTH[name])
    p["beta"] = p["beta"] * scale
    return m, times, data, m.pack(p), m.x0(configs.X0_PROPS[name])


def alone(ctx, name, J, N, stream, precision="fp64"):
    m, times, data, pv, x0 = inputs(name, N)
    return smc.run(ctx, m, J, times, data, pv, x0, precision=precision, seed=5, stream=stream)


def drain(srv, ids, max_steps=100000):
    for _ in range(max_steps):
        srv.step()
        if all(srv.status(i) == smc.SRV_DONE for i in ids):
            return
    raise AssertionError("server did not finish")


def test_served_equals_alone(ctx):
    srv = smc.Server(ctx)
    ids = []
    for name, J, N, stream in REQS[:5]:
        m, times, data, pv, x0 = inputs(name, N)
        ids.append(srv.submit(m, J, times, data, pv, x0, seed=5, stream=stream))
    # a few intervals in, more requests arrive and join the running batch
    for _ in range(3):
        srv.step()
    for name, J, N, stream in REQS[5:]:
        m, times, data, pv, x0 = inputs(name, N)
        ids.append(srv.submit(m, J, times, data, pv, x0, seed=5, stream=stream))
    drain(srv, ids)
    for rid, (name, J, N, stream) in zip(ids, REQS):
        print("request", name, J, N, flush=True)
        got = srv.result(rid)
        ref = alone(ctx, name, J, N, stream)
        assert got.loglik == ref.loglik
        np.testing.assert_array_equal(got.loglik_n, ref.loglik_n)
        np.testing.assert_array_equal(got.ess_n, ref.ess_n)
    srv.close()


def test_admission_by_memory(ctx):
    m, times, data, pv, x0 = inputs("sirs4", 30)
    one = smc.Filter(ctx, m, 4000, times, data, pv, x0).device_bytes()
    srv = smc.Server(ctx, hbm_budget_bytes=int(2.5 * one))
    ids = [srv.submit(m, 4000, times, data, pv, x0, seed=5, stream=100 + k) for k in range(6)]
    peak = 0
    results = {}
    for _ in range(10000):
        a, q, d = srv.step()
        peak = max(peak, a)
        for i in ids:
            if i not in results and srv.status(i) == smc.SRV_DONE:
                results[i] = srv.result(i)   # frees its memory: the next one can be admitted
        if len(results) == len(ids):
            break
    assert peak == 2 and len(results) == 6
    ref = smc.run(ctx, m, 4000, times, data, pv, x0, seed=5, stream=103)
    assert results[ids[3]].loglik == ref.loglik
    with pytest.raises(smc.PlomError):
        srv.submit(m, 40000, times, data, pv, x0)
    srv.close()


def test_served_fp32_mix(ctx):
    """FP32 requests served beside FP64 ones give the FP32 result alone."""
    srv = smc.Server(ctx)
    m, times, data, pv, x0 = inputs("sirs4", 20)
    a = srv.submit(m, 3000, times, data, pv, x0, precision="fp32", seed=5, stream=7)
    b = srv.submit(m, 3000, times, data, pv, x0, precision="fp64", seed=5, stream=7)
    drain(srv, [a, b])
    ra, rb = srv.result(a), srv.result(b)
    assert ra.loglik == alone(ctx, "sirs4", 3000, 20, 7, precision="fp32").loglik
    assert rb.loglik == alone(ctx, "sirs4", 3000, 20, 7).loglik
    srv.close()


def test_done_list_and_inflight_bound(ctx):
    """With one step in flight, requests admitted at different times still give their
    results alone; plom_server_done lists them in completion order (shorter N first)."""
    srv = smc.Server(ctx, max_inflight_steps=1)
    order = [("sir", 1200, 40, 21), ("sir", 900, 10, 22), ("sirs4", 600, 25, 23)]
    ids = []
    for k, (name, J, N, stream) in enumerate(order):
        m, times, data, pv, x0 = inputs(name, N)
        ids.append(srv.submit(m, J, times, data, pv, x0, seed=5, stream=stream))
        srv.step()
    finished = []
    for _ in range(1000):
        srv.step()
        for rid in srv.done():
            if rid not in finished:
                finished.append(rid)
        if len(finished) == len(ids):
            break
    assert finished == [ids[1], ids[2], ids[0]]
    for rid, (name, J, N, stream) in zip(ids, order):
        assert srv.result(rid).loglik == alone(ctx, name, J, N, stream).loglik
    assert srv.done() == []
    srv.close()


@pytest.mark.parametrize("heavy_blocks", [0, 3, -1])
def test_heavy_lane_equals_alone(ctx, heavy_blocks):
    """A request at or above the heavy threshold steps in its own lane beside the small
    ones, its propagate launches limited to heavy_blocks blocks while they run (0: one per
    CU; 3: a 6000-particle request's 24 tiles in 8 launches per interval; -1: no limit);
    every result still equals its run alone."""
    print("heavy lane, heavy_blocks", heavy_blocks, flush=True)
    srv = smc.Server(ctx, heavy_threshold=30000, heavy_blocks=heavy_blocks)   # sirs4 (S = 12) with J >= 2500
    reqs = [("sirs4", 6000, 40, 41), ("sir", 800, 30, 42), ("sir", 1500, 12, 43), ("sirs4", 700, 20, 44),
            ("sirs4", 3000, 10, 45)]
    ids = []
    for name, J, N, stream in reqs:
        m, times, data, pv, x0 = inputs(name, N)
        ids.append(srv.submit(m, J, times, data, pv, x0, seed=5, stream=stream))
    got = {}
    for _ in range(20000):
        srv.step()
        for rid in srv.done():
            got[rid] = srv.result(rid)
        if len(got) == len(ids):
            break
    assert len(got) == len(ids)
    for rid, (name, J, N, stream) in zip(ids, reqs):
        ref = alone(ctx, name, J, N, stream)
        assert got[rid].loglik == ref.loglik
        np.testing.assert_array_equal(got[rid].loglik_n, ref.loglik_n)
    srv.close()


def test_result_straight_after_step_frees_the_slot(ctx):
    """plom_server_result called right after the step that enqueued a request's last
    interval (no status or done call in between, so the request is still on the server's
    finished list): its slot goes back to the free list there.  With one slot, every later
    request is still admitted, and each result equals its run alone."""
    srv = smc.Server(ctx, max_active=1)
    reqs = [("sir", 500, 6, 61), ("sir", 700, 5, 62), ("sirs4", 300, 4, 63), ("sir", 900, 3, 64)]
    ids = []
    for name, J, N, stream in reqs:
        m, times, data, pv, x0 = inputs(name, N)
        ids.append(srv.submit(m, J, times, data, pv, x0, seed=5, stream=stream))
    got, early = {}, 0
    for _ in range(5000):
        srv.step()
        for rid in ids:
            if rid in got:
                continue
            try:
                got[rid] = srv.result(rid)
                early += 1
            except smc.PlomError as e:
                assert e.code == smc.ERR_STATE
        if len(got) == len(ids):
            break
    assert len(got) == len(ids), "a slot was lost: later requests were never admitted"
    assert early == len(ids)
    for rid, (name, J, N, stream) in zip(ids, reqs):
        ref = alone(ctx, name, J, N, stream)
        assert got[rid].loglik == ref.loglik
        np.testing.assert_array_equal(got[rid].loglik_n, ref.loglik_n)
    srv.close()


def test_heavy_slot_reused_by_the_light_lane(ctx):
    """Regression test of the cross-lane slot reuse: two slots; a heavy request with a
    lineage steps on lane 1 beside a long light request on lane 0, and light requests wait
    in the queue.  When the heavy request finishes, the next light request is admitted on
    lane 0 into the slot it frees, only once lane 1's queued steps, lineage launch and
    copy-back of that slot have completed.  Every result (the heavy lineage included)
    equals its run alone and the CPU oracle."""
    from oracle import oracle
    srv = smc.Server(ctx, heavy_threshold=30000, max_active=2)   # sirs4 (S = 12) with J >= 2500 is heavy
    heavy = ("sirs4", 6000, 12, 71)
    light = [("sir", 800, 52, 72), ("sir", 1500, 9, 73), ("sirs4", 700, 7, 74), ("sir", 2500, 5, 75)]
    m, times, data, pv, x0 = inputs(heavy[0], heavy[2])
    hid = srv.submit(m, heavy[1], times, data, pv, x0, seed=5, stream=heavy[3], lineage=True)
    ids = []
    for name, J, N, stream in light:
        m, times, data, pv, x0 = inputs(name, N)
        ids.append(srv.submit(m, J, times, data, pv, x0, seed=5, stream=stream))
    got, order = {}, []
    for _ in range(20000):
        srv.step()
        for rid in srv.done():
            got[rid] = srv.result(rid)
            order.append(rid)
        if len(got) == len(ids) + 1:
            break
    assert len(got) == len(ids) + 1
    # two slots: when the heavy request finishes, one light request holds the other slot,
    # so any further light request that finishes after it was admitted after the heavy
    # slot was freed, into that slot unless the other one had just been freed too
    assert len(order) - 1 - order.index(hid) >= 2
    name, J, N, stream = heavy
    m, times, data, pv, x0 = inputs(name, N)
    ref = smc.run(ctx, m, J, times, data, pv, x0, seed=5, stream=stream, lineage=True)
    o = oracle.smc(m.id, J, times, data, pv, x0, seed=5, stream=stream, lineage=True, nthreads=8)
    assert got[hid].loglik == ref.loglik == o["loglik"]
    np.testing.assert_array_equal(got[hid].lineage, ref.lineage)
    np.testing.assert_array_equal(got[hid].lineage, o["lineage"])
    np.testing.assert_array_equal(got[hid].lineage_slots, o["lineage_slots"])
    for rid, (name, J, N, stream) in zip(ids, light):
        m, times, data, pv, x0 = inputs(name, N)
        o = oracle.smc(m.id, J, times, data, pv, x0, seed=5, stream=stream, nthreads=8)
        assert got[rid].loglik == alone(ctx, name, J, N, stream).loglik == o["loglik"]
        np.testing.assert_array_equal(got[rid].loglik_n, o["loglik_n"])
    srv.close()
