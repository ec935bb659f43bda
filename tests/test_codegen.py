"""Model builder against the reference's own codegen tests (golden fixtures transcribed
from model_builder/test_Cmodel.py and test_Ccoder.py) and against an independent
Euler integration of the model equations."""
import json
import math
import os

import numpy as np
import pytest

from iip_amd.codegen import PlomModel, expr as E, load_model, emit_header, generate_all
from oracle import oracle
from iip_amd import configs
from iip_amd.smc import Model

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture():
    return json.load(open(os.path.join(GOLD, "cmodel_fixture.json")))


def strip(pm):
    return [{k: v for k, v in r.items() if k not in ("ind_cache", "ind_dem_sto", "reactions")} for r in pm]


def strip_obs(defs):
    return [[({k: v for k, v in d.items() if k != "reactions"} if isinstance(d, dict) else d) for d in lst] for lst in defs]


@pytest.mark.parametrize("which", ["m", "m2", "m3"])
def test_cmodel_rewriting(which):
    fx = fixture()
    m = PlomModel(fx["context"], fx["process"][which], fx["link"])
    exp = fx["expected"][which]
    assert m.par_sv == exp["par_sv"]
    assert m.remainder == exp["remainder"]
    assert strip(m.proc_model) == exp["proc_model"]
    assert strip_obs(m.obs_var_def) == exp["obs_var_def"]


def test_cmodel_lists():
    fx = fixture()
    m = PlomModel(fx["context"], fx["process"]["m"], fx["link"])
    c = fx["expected"]["common"]
    assert set(m.par_proc) == set(c["par_proc"])
    assert set(m.par_obs) == set(c["par_obs"])
    assert set(m.par_fixed) == set(c["par_fixed"]) and m.par_fixed[0] == "N"
    assert set(m.par_fixed_obs) == set(c["par_fixed_obs"])
    assert m.drift_par_proc == c["drift_par_proc"]
    assert m.vol_par_proc == c["vol_par_proc"]
    assert m.drift_var == c["drift_var"]
    assert m.obs_var == c["obs_var"]
    assert E.change_user_input(c["change_user_input"]["in"]) == c["change_user_input"]["out"]


def test_parser_against_reference_human_forms():
    """The reference's SymPy-simplified forms must evaluate to the same numbers as the
    inputs under this build's parser (precedence, associativity, unary minus, **)."""
    pairs = json.load(open(os.path.join(GOLD, "ccoder_terms.json")))["pairs"]
    funcs = {"sin": math.sin, "cos": math.cos, "pow": math.pow}
    rng = np.random.default_rng(1)
    for p in pairs:
        a, b = E.parse(p["x"]), E.parse(p["h"])
        for _ in range(50):
            env = {s: float(rng.uniform(0.1, 3.0)) for s in E.symbols(a) | E.symbols(b)}
            va, vb = E.evaluate(a, env, funcs), E.evaluate(b, env, funcs)
            assert va == pytest.approx(vb, rel=1e-13, abs=1e-13)


def test_parser_precedence():
    f = {}
    assert E.evaluate(E.parse("-2**2"), {}, f) == -4.0
    assert E.evaluate(E.parse("2**3**2"), {}, f) == 512.0
    assert E.evaluate(E.parse("8/4/2"), {}, f) == 1.0
    assert E.evaluate(E.parse("1-2-3"), {}, f) == -4.0
    assert E.evaluate(E.parse("2*-3"), {}, f) == -6.0
    assert E.evaluate(E.parse("1e-3*2"), {}, f) == 0.002


def test_generated_headers_up_to_date():
    """The committed gen/ headers are what the generator produces from examples/."""
    gen = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "iterated-inference-pipeline_amd", "gen")
    for name in ("sir", "sirs4", "seir10", "noise", "drift"):
        text = emit_header(load_model(name))
        assert open(os.path.join(gen, "plom_%s.h" % name)).read() == text


def test_obs2ts_and_layout():
    m = load_model("sirs4")
    assert m.N_CAC == 4 and m.N_TS == 4 and m.size_proj == 12
    assert [o["n_ts_unique"] for o in m.obs2ts] == [4]
    assert m.obs2ts[0]["cac"] == [[[0, 0]], [[1, 0]], [[2, 0]], [[3, 0]]]
    s = load_model("seir10")
    assert s.size_proj == 40 and s.N_CAC == 10


def python_euler(model, x0, pv, times, dt, noise=None):
    """Euler(-Maruyama) of the model written from the reaction list (Ccoder.py:366-505,
    core_template.c:258-336): for every reaction with term T = rate*X_from (rate alone
    from U or the remainder), the increment T dt + sqrt(T) dW_r + T sd dW_env (the last
    only for reactions carrying white noise) leaves `from`, enters `to` and is added to
    the observed incidences of that reaction; states and incidences are clipped at 0
    after each step.  noise = (seed, stream, j, noises_off) draws dW = sqrt(dt) z from
    oracle.normals in the order population-major, demographic noises (one per reaction)
    then environmental ones, pairs drawn per substep (an open pair's second value is
    dropped at the end of the substep); noise = None is the deterministic skeleton."""
    m = model.desc
    lay = m.layout
    trees = [E.parse(r["rate"]) for r in m.proc_model]
    C = np.asarray(m.coupling) if m.coupling is not None else None
    R, W = len(m.proc_model), len(m.white_noise)
    dem_on = noise is not None and not (noise[3] & 1)
    env_on = noise is not None and not (noise[3] & 2) and W > 0
    K = m.N_CAC * ((R if dem_on else 0) + (W if env_on else 0))
    per_sub = 2 * ((K + 1) // 2)
    sq = math.sqrt(dt)
    out = [x0.copy()]
    x = x0.copy()
    for n in range(len(times) - 1):
        for u in range(m.N_TS_INC_UNIQUE):
            x[m.inc_offset + u] = 0.0
        nsub, tt = 0, times[n]
        while tt < times[n + 1]:
            nsub, tt = nsub + 1, tt + dt
        z = oracle.normals(noise[0], noise[1], n, noise[2], nsub * per_sub) if K else None
        t = times[n]
        ksub = 0
        while t < times[n + 1]:
            new = x.copy()
            zs = z[ksub * per_sub:ksub * per_sub + K] if K else None
            ksub += 1
            i = 0
            for c in range(m.N_CAC):
                dem = [0.0] * R
                env = [0.0] * W
                if dem_on:
                    dem = [sq * zs[i + r] for r in range(R)]
                    i += R
                if env_on:
                    env = [sq * zs[i + w] for w in range(W)]
                    i += W
                env_v = {s: x[q * m.N_CAC + c] for q, s in enumerate(m.par_sv)}
                env_v.update({p: pv[off + c] for p, off in lay["proc"].items()})
                env_v.update({p: pv[off + c] for p, off in lay["fixed"].items()})
                env_v.update({"t": t, "M_PI": math.pi, "ONE_YEAR": 365.0})
                funcs = {"sin": math.sin, "cos": math.cos, "correct_rate": lambda v: v,
                         "mix": None}

                def mix_factory(c=c):
                    return lambda _v: None
                for ri, (r, tree) in enumerate(zip(m.proc_model, trees)):
                    if "mix(" in r["rate"]:
                        name = r["rate"].split("mix(")[1].split(")")[0]
                        k = m.par_sv.index(name)
                        mixed = sum(C[c, kk] * x[k * m.N_CAC + kk] for kk in range(m.N_CAC))
                        funcs["mix"] = lambda _v, mixed=mixed: mixed
                    rate = E.evaluate(tree, env_v, funcs)
                    flow = rate * (x[m.par_sv.index(r["from"]) * m.N_CAC + c] if r["from"] in m.par_sv else 1.0)
                    inc = flow * dt + math.sqrt(flow) * dem[ri]
                    if "white_noise" in r:
                        w = [wn["name"] for wn in m.white_noise].index(r["white_noise"]["name"])
                        inc += flow * pv[lay["proc"][r["white_noise"]["sd"]] + c] * env[w]
                    if r["from"] in m.par_sv:
                        new[m.par_sv.index(r["from"]) * m.N_CAC + c] -= inc
                    if r["to"] in m.par_sv:
                        new[m.par_sv.index(r["to"]) * m.N_CAC + c] += inc
                    # observed incidences of this reaction
                    u = 0
                    for o, lst in enumerate(m.obs_var_def):
                        if not m.is_inc[o]:
                            continue
                        for tsu in range(m.obs2ts[o]["n_ts_unique"]):
                            cacs = [ci * m.N_AC + ai for ci, ai in m.obs2ts[o]["cac"][tsu]]
                            if c in cacs and any(d["from"] == r["from"] and d["to"] == r["to"] for d in lst):
                                new[m.inc_offset + u] += inc
                            u += 1
            x = np.maximum(new, 0.0)
            t += dt
        out.append(x.copy())
    return np.array(out)


@pytest.mark.parametrize("name", ["sir", "sirs4", "seir10"])
def test_generated_euler_matches_independent_euler(name):
    model = Model.get(name)
    times = np.arange(9, dtype=np.float64) * 7.0
    pv = model.pack(configs.TRUTH[name])
    x0 = model.x0(configs.X0_PROPS[name])
    got = oracle.trajectory(model.id, times, pv, x0, noises_off=3, dt=0.25)
    ref = python_euler(model, x0, pv, times, 0.25)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-7)


@pytest.mark.parametrize("name,noises_off", [("sir", 0), ("sirs4", 0), ("seir10", 0), ("sirs4", 1), ("sirs4", 2)])
def test_generated_sde_matches_independent_euler_maruyama(name, noises_off):
    """The generated Euler-Maruyama step (demographic and environmental noise) against the
    restatement above fed with the same Gaussian stream."""
    model = Model.get(name)
    times = np.arange(7, dtype=np.float64) * 7.0
    pv = model.pack(configs.TRUTH[name])
    x0 = model.x0(configs.X0_PROPS[name])
    got = oracle.trajectory(model.id, times, pv, x0, noises_off=noises_off, seed=3, stream=8, j=5, dt=0.25)
    ref = python_euler(model, x0, pv, times, 0.25, noise=(3, 8, 5, noises_off))
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-6)
    # the noise matters: the deterministic skeleton is far from it
    det = python_euler(model, x0, pv, times, 0.25)
    assert np.max(np.abs(got - det)) > 1.0
