"""Register the package directory ``iterated-inference-pipeline_amd`` (not a Python
identifier) as the module ``iip_amd``."""
import importlib.util
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "iterated-inference-pipeline_amd")


def load_package():
    if "iip_amd" in sys.modules:
        return sys.modules["iip_amd"]
    spec = importlib.util.spec_from_file_location("iip_amd", os.path.join(PKG, "__init__.py"),
                                                  submodule_search_locations=[PKG])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["iip_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
