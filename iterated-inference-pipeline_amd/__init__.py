"""MI355X-native PLOM SMC particle filter (host-side Python mirror).

The directory name is not a Python identifier; load it with ``load_package()`` from a
checkout (tests, bench.py and __graft_entry__.py do) and it registers as ``iip_amd``.

Layout:
  csrc/      HIP kernels for gfx950 and the C ABI (libplom_smc.so), shared math/RNG
  codegen/   model builder: PLOM model JSON -> generated step/observation code
  gen/       generated model headers compiled into the library
  cli/       the drop-in `smc` command (C) over the C ABI
  examples/  model files (reference grammar) of the benchmark configurations
  smc.py     ctypes binding of the C ABI and the Python filter API
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))


def lib_path():
    return os.path.join(PKG_DIR, "csrc", "libplom_smc.so")


def __getattr__(name):
    # lazy: importing the package must not need the built library
    if name in ("smc",):
        import importlib
        return importlib.import_module(__name__ + "." + name)
    raise AttributeError(name)
