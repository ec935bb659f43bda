"""The C ABI library loads and exports every entry point include/plom_smc.h declares
(no GPU needed: only host-side queries are called here)."""
import ctypes as C
import os
import re

import pytest

from iip_amd import smc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(plom_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = smc.lib()
    names = declared("plom_smc.h")
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(smc.EXPORTS) <= set(names)


def test_host_side_queries():
    L = smc.lib()
    assert L.plom_model_count() == 5
    for i, name in enumerate(["sir", "sirs4", "seir10", "noise", "drift"]):
        assert L.plom_model_find(name.encode()) == i
        info = smc.ModelInfo()
        assert L.plom_model_info_get(i, C.byref(info)) == 0
        m = smc.Model.get(name)
        assert info.size_proj == m.S and info.n_ts == m.n_ts and info.pv_size == m.layout["size"]
        assert info.n_hat == m.n_hat == {"sir": 4, "sirs4": 16, "seir10": 50, "noise": 10, "drift": 12}[name]
        assert info.fix_off == m.fix_off and info.n_fix == m.n_fix == {"sir": 1, "sirs4": 4, "seir10": 10, "noise": 12,
                                                                       "drift": 12}[name]
        assert info.n_drift == m.n_drift == (1 if name == "drift" else 0)
    assert L.plom_model_find(b"nope") < 0
    assert b"unknown model" in L.plom_last_error()


def test_device_bytes_accounting():
    d = smc._desc(smc.Model.get("seir10"), 1000000, 1040, "fp64", 0, smc.FLAG_STORE_ANCESTRY, 1, 0, 0.25, 1e-17)
    b = smc.lib().plom_filter_device_bytes(C.byref(d))
    # two FP64 state slices of 40 x 1e6, ancestry int32 1040 x 1e6, plus per-particle scratch
    assert 2 * 40 * 8 * 10 ** 6 + 4 * 1040 * 10 ** 6 < b < 2 * 40 * 8 * 10 ** 6 + 4 * 1040 * 10 ** 6 + 64 * 10 ** 6
    bad = smc._desc(smc.Model.get("sir"), 0, 10, "fp64", 0, 0, 1, 0, 0.25, 1e-17)
    assert smc.lib().plom_filter_device_bytes(C.byref(bad)) < 0
