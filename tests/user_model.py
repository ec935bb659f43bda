"""The user-model fixture of the tests: the reference's example model without a remainder
state (model_builder/example/noise_no_remainder: S, I and R all explicit, so N in the
rates is the sum S + I + R, Cmodel.py:110-153), with the context and observation model of
examples/noise, compiled outside the library into its own model library
(build.py --model) and into an extra CPU oracle (oracle/Makefile `extra`)."""
import os

from iip_amd import PKG_DIR
from iip_amd.smc import Model
from oracle import oracle

NAME = "noise_no_remainder"
DIR = os.path.join(PKG_DIR, "examples", NAME)
BUILD = os.path.join(DIR, "build")
LIBRARY = os.path.join(BUILD, "libplom_model_%s.so" % NAME)
TRUTH = {"r0": 1.8, "v": 0.1, "sto": 0.1, "rep": 0.6, "phi": 0.1}
X0_PROPS = {"S": [0.7, 0.65], "I": [0.001, 0.0005], "R": [0.299, 0.3495]}


def load(build_if_missing=False):
    """The registered model (Model.load) with its oracle library; the model library is
    built here only when asked (CPU tests), never on the GPU box."""
    if not os.path.exists(LIBRARY):
        if not build_if_missing:
            raise FileNotFoundError("%s is not built (build.py --model %s)" % (LIBRARY, DIR))
        import importlib.util
        spec = importlib.util.spec_from_file_location("plom_build_lib", os.path.join(PKG_DIR, "build.py"))
        b = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(b)
        b.build_model_lib(DIR, BUILD)
    m = Model.load(DIR, LIBRARY)
    oracle.extra_lib(NAME, BUILD)
    return m
