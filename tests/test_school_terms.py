"""School-term forcing, terms_forcing(amplitude) in a rate (C/core/special_functions.c:21-47,
the school terms loaded by C/core/build.c:834-868), on the CPU:

* plom_terms_forcing against its closed form on a grid of times over three years: the
  fraction of the year y = t/365 - floor(t/365) in exact rational arithmetic decides
  holiday (y in some [begin, end], bounds included) or school; the value is
  amplitude / (1 + amplitude H) in holidays and 1 / (1 + amplitude H) in school, H the
  holiday share of the year (1 - prop_school) summed from the calendar; the yearly mean
  is (amplitude H + 1 - H) / (1 + amplitude H);
* the model's pv record (n_terms, prop_school, periods), the settings (IS_SCHOOL_TERMS,
  data.school_terms) and the `smc` command's packing of the same record;
* the filter of the example model school_terms with an empty calendar (forcing exactly 1)
  equals, bit for bit, the filter of noise_no_remainder, the same model without the factor.
"""
import json
import math
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle
import user_model as U

CAL = [0.0, 0.0137, 0.1315, 0.1589, 0.25, 0.5, 0.8082, 0.8356, 0.9699, 1.0]


def record(cal):
    rec = np.zeros(18)
    rec[0] = len(cal) // 2
    rec[1] = 1.0 - sum(cal[k + 1] - cal[k] for k in range(0, len(cal), 2))
    rec[2:2 + len(cal)] = cal
    return rec


def in_holiday(t, cal):
    """exact: the fraction of the year of t (a double, days) and the periods as doubles"""
    ty = Fraction(t) / 365
    y = ty - math.floor(ty)
    return any(Fraction(cal[k]) <= y <= Fraction(cal[k + 1]) for k in range(0, len(cal), 2))


def near_boundary(t, cal, eps=1e-9):
    ty = t / 365.0
    y = ty - math.floor(ty)
    return any(abs(y - b) < eps for b in cal) or abs(y - 1.0) < eps


@pytest.mark.parametrize("amp", [0.3, 1.0, 1.6, 4.0])
def test_terms_forcing_closed_form(amp):
    rec = record(CAL)
    H = sum(CAL[k + 1] - CAL[k] for k in range(0, len(CAL), 2))
    hol, sch = amp / (1.0 + amp * H), 1.0 / (1.0 + amp * H)
    t = np.arange(0.0, 3 * 365.0, 0.0625)
    got = oracle.terms_forcing(amp, t, rec)
    n_hol = 0
    for ti, g in zip(t, got):
        if near_boundary(ti, CAL):
            continue
        want = hol if in_holiday(ti, CAL) else sch
        assert g == pytest.approx(want, rel=1e-15), ti
        n_hol += in_holiday(ti, CAL)
    assert 0 < n_hol < len(t)
    # holiday / school = amplitude; the yearly mean (one year of the uniform grid)
    assert hol / sch == pytest.approx(amp, rel=1e-15)
    year = got[t < 365.0]
    assert float(np.mean(year)) == pytest.approx((amp * H + 1.0 - H) / (1.0 + amp * H), abs=2e-3)


def test_terms_forcing_bounds_are_included():
    cal = [0.25, 0.5]
    rec = record(cal)
    amp = 2.0
    hol, sch = amp / (1.0 + amp * 0.25), 1.0 / (1.0 + amp * 0.25)
    # 91.25 / 365 = 0.25 and 182.5 / 365 = 0.5 exactly in double, also a year later
    for t in (91.25, 182.5, 365.0 + 91.25, 730.0 + 182.5):
        ty = t / 365.0
        assert ty - math.floor(ty) in (0.25, 0.5)
        assert oracle.terms_forcing(amp, [t], rec)[0] == hol
    for t in (np.nextafter(91.25, 0.0), np.nextafter(182.5, 1e9), 0.0, 300.0):
        assert oracle.terms_forcing(amp, [t], rec)[0] == sch
    # no periods: prop_school 1, the forcing is 1 for every amplitude
    assert np.all(oracle.terms_forcing(3.7, np.linspace(0, 800, 97), record([])) == 1.0)


@pytest.fixture(scope="module")
def model():
    return U.load(build_if_missing=True, name="school_terms")


def test_pv_record_and_settings(model):
    d = model.desc
    lay = d.layout
    assert d.is_school_terms and lay["school"] is not None and lay["size"] == lay["school"] + 2 * 18
    pv = model.pack(U.TRUTHS["school_terms"])
    cal = json.load(open(os.path.join(U.paths("school_terms")[0], "context.json")))["data"]["school_terms"]
    for cac in range(2):
        np.testing.assert_array_equal(pv[lay["school"] + 18 * cac:lay["school"] + 18 * (cac + 1)], record(cal[cac]))
    s = d.settings()
    assert s["cst"]["IS_SCHOOL_TERMS"] == 1 and s["data"]["school_terms"] == cal
    # the other example has no school block
    assert U.load(build_if_missing=True).desc.settings()["cst"]["IS_SCHOOL_TERMS"] == 0


def test_cli_packs_the_school_record(model, tmp_path):
    from iip_amd import PKG_DIR
    d = U.paths("school_terms")[0]
    subprocess.check_call(["python", os.path.join(PKG_DIR, "plom_build.py"), d, "-o", str(tmp_path),
                           "--model_library", U.paths("school_terms")[2]], stdout=subprocess.DEVNULL)
    theta = os.path.join(d, "theta.json")
    r = subprocess.run([os.path.join(PKG_DIR, "cli", "smc"), "sde", "-J", "10", "--dry_run"], stdin=open(theta), cwd=str(tmp_path),
                       capture_output=True, text=True, check=True)
    pv = np.array(json.loads(r.stdout)["pv"])
    lay = model.desc.layout
    np.testing.assert_array_equal(pv[lay["school"]:], model.pack(U.TRUTHS["school_terms"])[lay["school"]:])


@pytest.mark.parametrize("psr", [False, True])
def test_empty_calendar_equals_the_model_without_forcing(model, psr):
    base = U.load(build_if_missing=True)
    N = 40
    times, data = model.dataset(N)
    cov = model.covariates()
    x0 = model.x0(U.X0_PROPS)
    if psr:
        x0 = np.round(x0)
    pv = model.pack(U.TRUTHS["school_terms"])
    lay = model.desc.layout
    pv[lay["school"]:] = np.concatenate([record([]), record([])])
    a = oracle.smc("school_terms", 500, times, data, pv, x0, seed=3, stream=2, cov=cov, psr=psr, nthreads=8)
    b = oracle.smc(U.NAME, 500, times, data, base.pack(U.TRUTH), x0, seed=3, stream=2, cov=cov, psr=psr, nthreads=8)
    assert a["loglik"] == b["loglik"]
    np.testing.assert_array_equal(a["loglik_n"], b["loglik_n"])
    # the calendar changes the filter
    c = oracle.smc("school_terms", 500, times, data, model.pack(U.TRUTHS["school_terms"]), x0, seed=3, stream=2, cov=cov,
                   psr=psr, nthreads=8)
    assert c["loglik"] != a["loglik"]
