"""The user-model fixtures of the tests: models compiled outside the library, each into its
own model library (build.py --model) and into an extra CPU oracle (oracle/Makefile `extra`).

noise_no_remainder  the reference's example model without a remainder state
                    (model_builder/example/noise_no_remainder: S, I and R all explicit, so N
                    in the rates is the sum S + I + R, Cmodel.py:110-153), with the context
                    and observation model of examples/noise;
school_terms        the same model with school-term forcing of the transmission rate,
                    r0/N*v*terms_forcing(e)*I (C/core/special_functions.c:21-47), and a
                    holiday calendar per population (context data.school_terms)."""
import os

from iip_amd import PKG_DIR
from iip_amd.smc import Model
from oracle import oracle

NAME = "noise_no_remainder"
TRUTHS = {
    "noise_no_remainder": {"r0": 1.8, "v": 0.1, "sto": 0.1, "rep": 0.6, "phi": 0.1},
    "school_terms": {"r0": 1.8, "v": 0.1, "e": 1.6, "sto": 0.1, "rep": 0.6, "phi": 0.1},
}
X0_PROPS = {"S": [0.7, 0.65], "I": [0.001, 0.0005], "R": [0.299, 0.3495]}
TRUTH = TRUTHS[NAME]


def paths(name):
    d = os.path.join(PKG_DIR, "examples", name)
    b = os.path.join(d, "build")
    return d, b, os.path.join(b, "libplom_model_%s.so" % name)


DIR, BUILD, LIBRARY = paths(NAME)


def load(build_if_missing=False, name=NAME):
    """The registered model (Model.load) with its oracle library; the model library is
    built here only when asked (CPU tests), never on the GPU box."""
    d, b, lib = paths(name)
    if not os.path.exists(lib):
        if not build_if_missing:
            raise FileNotFoundError("%s is not built (build.py --model %s)" % (lib, d))
        import importlib.util
        spec = importlib.util.spec_from_file_location("plom_build_lib", os.path.join(PKG_DIR, "build.py"))
        bl = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(bl)
        bl.build_model_lib(d, b)
    m = Model.load(d, lib)
    oracle.extra_lib(name, b)
    return m
