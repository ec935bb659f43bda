"""Drifting parameters on the GPU (the drift states move after every substep, the rates read
their natural values) against the CPU oracle, which is pinned in test_drift.py: FP64
bit-exact for the SDE (one group per population, and one shared group) and psr
implementations, hat and lineages, batched with models without drift; PLOM_NO_DRIFT_BIT
keeps the pv values; the refusals (MIF, a missing map)."""
import numpy as np
import pytest

from oracle import oracle
from iip_amd import configs, smc
from iip_amd.smc import Model

pytestmark = pytest.mark.gpu


def setup(N, groups=None):
    m = Model.get("drift")
    times, data = m.dataset(N)
    pv = m.pack(configs.TRUTH["drift"])
    x0, drift = configs.drift_setup(m, m.x0(configs.X0_PROPS["drift"]), configs.TRUTH["drift"], groups=groups)
    return m, times, data, pv, x0, drift, m.covariates()


@pytest.mark.parametrize("J,N,groups,noises_off,flags", [(2000, 104, None, 0, 0), (700, 50, [[0, 0]], 0, 0),
                                                         (1000, 60, None, 2, smc.FLAG_PSR), (257, 30, [[1, 0]], 1, 0)])
def test_drift_filter_equals_oracle(ctx, J, N, groups, noises_off, flags):
    m, times, data, pv, x0, drift, cov = setup(N, groups)
    if flags & smc.FLAG_PSR:
        x0[:m.desc.drift_offset] = np.round(x0[:m.desc.drift_offset])
    print("drift", J, N, groups, noises_off, flags, flush=True)
    f = smc.Filter(ctx, m, J, times, data, pv, x0, noises_off=noises_off, flags=flags | smc.FLAG_HAT | smc.FLAG_LINEAGE,
                   seed=5, stream=2, cov=cov, drift=drift)
    f.run()
    g = f.read(hat=True, lineage=True)
    f.close()
    o = oracle.smc(m.id, J, times, data, pv, x0, noises_off=noises_off, seed=5, stream=2, hat=True, lineage=True,
                   nthreads=8, cov=cov, drift=drift, psr=bool(flags & smc.FLAG_PSR))
    np.testing.assert_array_equal(g.loglik_n, o["loglik_n"])
    np.testing.assert_array_equal(g.ess_n, o["ess_n"])
    np.testing.assert_array_equal(g.hat, o["hat"])
    np.testing.assert_array_equal(g.lineage, o["lineage"])
    assert g.loglik == o["loglik"]


def test_no_drift_bit_keeps_pv(ctx):
    m, times, data, pv, x0, drift, cov = setup(40)
    print("drift off", flush=True)
    g = smc.run(ctx, m, 1500, times, data, pv, x0, noises_off=smc.NO_DRIFT, seed=3, cov=cov, drift=drift)
    o = oracle.smc(m.id, 1500, times, data, pv, x0, seed=3, cov=cov, nthreads=8)
    np.testing.assert_array_equal(g.loglik_n, o["loglik_n"])
    # without the bit the drift moves r0: another likelihood
    d = smc.run(ctx, m, 1500, times, data, pv, x0, seed=3, cov=cov, drift=drift)
    assert d.loglik != g.loglik


@pytest.mark.parametrize("lanes", [1, 2])
def test_drift_batched_with_other_models(ctx, lanes):
    m, times, data, pv, x0, drift, cov = setup(50)
    s = Model.get("sirs4")
    st, sd = s.dataset(30)
    cases = [(m, times, data, pv, x0, cov, drift, 1200), (s, st, sd, s.pack(configs.TRUTH["sirs4"]),
                                                          s.x0(configs.X0_PROPS["sirs4"]), None, None, 900),
             (m, times[:21], data[:20], pv, x0, None, drift, 3000)]
    fs, alone = [], []
    for k, (mm, t, d, p, x, c, dr, J) in enumerate(cases):
        print("drift batch", mm.name, J, flush=True)
        fs.append(smc.Filter(ctx, mm, J, t, d, p, x, seed=8, stream=k, cov=c, drift=dr))
        alone.append(smc.run(ctx, mm, J, t, d, p, x, seed=8, stream=k, cov=c, drift=dr))
    b = smc.Batch(ctx, fs)
    b.configure(lanes)
    for rep in range(2):
        b.launch(use_graph=True)
        for f, a in zip(fs, alone):
            r = f.read()
            np.testing.assert_array_equal(r.loglik_n, a.loglik_n)
    b.close()
    for f in fs:
        f.close()


def test_drift_fp32_close_to_fp64_without_noise(ctx):
    """Only the drift's noise, equal random streams: the FP32 path stays within 1e-3 of FP64
    over 30 weeks (the same normals rounded to FP32)."""
    m, times, data, pv, x0, drift, cov = setup(30)
    a = smc.run(ctx, m, 1, times, data, pv, x0, noises_off=3, seed=4, cov=cov, drift=drift, lineage=True)
    b = smc.run(ctx, m, 1, times, data, pv, x0, noises_off=3, seed=4, cov=cov, drift=drift, lineage=True,
                precision="fp32")
    np.testing.assert_allclose(b.lineage[:, :4], a.lineage[:, :4], rtol=1e-3)


def test_drift_refusals(ctx):
    m, times, data, pv, x0, drift, cov = setup(10)
    with pytest.raises(smc.PlomError):
        smc.Filter(ctx, m, 10, times, data, pv, x0)                       # no map, drift on
    srv = smc.Server(ctx)
    try:
        with pytest.raises(smc.PlomError):
            srv.submit(m, 100, times, data, pv, x0)
        rid = srv.submit(m, 100, times, data, pv, x0, noises_off=smc.NO_DRIFT)
        while srv.status(rid) != smc.SRV_DONE:
            srv.step()
        r = srv.result(rid)
        o = oracle.smc(m.id, 100, times, data, pv, x0)
        assert r.loglik == o["loglik"]
    finally:
        srv.close()


@pytest.mark.parametrize("J,N,noises_off,flags", [(1500, 40, 0, 0), (300, 20, 3, 0), (500, 20, 0, smc.FLAG_PSR)])
def test_mif_pass_with_drift_equals_oracle(ctx, J, N, noises_off, flags):
    """Iterated filtering with drifting parameters: the drift states start at the particles'
    theta and move with each particle's own volatility (a walked component)."""
    from test_drift import DRIFT_MAP, mif_drift_setup
    from test_gpu_mif import check_equal
    m, times, data, pv, drift, theta0, X0 = mif_drift_setup(N, J, 3, spread=0.05)
    if flags & smc.FLAG_PSR:
        off = m.desc.drift_offset
        X0[:off] = np.round(X0[:off])
    mut = np.linspace(0.5, 1.0, N)
    cov = m.covariates()
    print("mif drift", J, N, noises_off, flags, flush=True)
    f = smc.Filter(ctx, m, J, times, data, pv, X0[:, 0], noises_off=noises_off, flags=smc.FLAG_MIF | flags,
                   mif=DRIFT_MAP, cov=cov, drift=drift)
    f.mif_start(theta0, X0, mut, seed=6, stream=2)
    f.run()
    g = f.read(mif=True)
    f.close()
    o = oracle.mif_pass(m.id, J, times, data, pv, DRIFT_MAP, theta0, X0, mut, noises_off=noises_off, seed=6, stream=2,
                        nthreads=8, drift=drift, cov=cov, psr=bool(flags & smc.FLAG_PSR))
    check_equal(g, o)


def test_mif_command_with_drift(tmp_path):
    """`mif` on the drift example: r0's groups are fitted by fixed-lag smoothing (drift
    states from theta), vol by the random walk; the command's trace equals the restatement's
    (oracle/mif_oracle.py driving the CPU pass with the dry run's map)."""
    import json
    import os
    import subprocess
    from oracle import mif_oracle
    from test_cli import PKG, model_dir
    from test_gpu_mif_cli import MIF, read_csv
    d = model_dir(tmp_path, "drift")
    theta = os.path.join(PKG, "examples", "drift", "theta.json")
    J, M = 300, 3
    args = ["sde", "-J", str(J), "-M", str(M), "-S", "1", "--trace_17g", "-i", "0"]

    def run(a):
        with open(theta) as fh:
            return subprocess.run([MIF] + a, cwd=str(d), stdin=fh, capture_output=True, text=True, timeout=600)
    dry = json.loads(run(args + ["--dry_run"]).stdout)
    assert dry["drift"] is not None and dry["sv"][4] == 1 and dry["sd"][4] == 0.0
    print("mif command drift", flush=True)
    r = run(args)
    assert r.returncode == 0, r.stderr
    m = Model.get("drift")
    times, data = m.dataset(dry["N"])
    mp = {k: dry[k] for k in ("pv_src", "tr", "scale", "duration", "min", "max", "multiplier", "sd")}
    pv, pop0, tr, mn, mx = np.array(dry["pv"]), dry["pop0"], dry["tr"], dry["min"], dry["max"]
    cov = m.covariates()

    def pass_fn(th, x0, mut, seed, stream):
        return oracle.mif_pass(m.id, th.shape[1], times, data, pv, mp, th, x0, mut, seed=seed, stream=stream,
                               nthreads=8, drift=dry["drift"], cov=cov)

    def x0_of(t):
        return ([mif_oracle.f_inv(tr[o], t[o], mn[o], mx[o]) * pop0[o % 2] for o in range(4)] + [t[4], t[5]] +
                [0.0] * (m.S - 6))

    def constrain(t):
        return sum(int(mif_oracle.f_inv(tr[c], t[c], mn[c], mx[c]) + mif_oracle.f_inv(tr[2 + c], t[2 + c], mn[2 + c],
                                                                                      mx[2 + c]) > 1.0) for c in (0, 1))

    np.testing.assert_array_equal(x0_of(dry["theta"]), dry["x0"])
    o = mif_oracle.run(dry, pass_fn, x0_of, constrain, M=M, J=J, seed=2, switch=1)
    head, rows = read_csv(tmp_path / "drift" / "trace_0.csv")
    assert len(rows) == M + 1
    for (idx, vals, ll), row in zip(o["trace"], rows):
        assert row[0] == idx
        np.testing.assert_array_equal(row[1:-1], vals)
        if idx:
            assert row[-1] == ll
