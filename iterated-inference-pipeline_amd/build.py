"""Build libplom_smc.so (HIP, gfx950) and the `smc`, `pmcmc` and `mif` commands from source,
in-tree; build a user's model into a model library of its own (build_model_lib).

The library is one translation unit of model-independent kernels and host code
(csrc/plom_smc.hip) plus one translation unit per built-in model (csrc/plom_model.hip
compiled with that model's generated header), compiled in parallel and linked together.

Compiler flags that matter for results:
  -ffp-contract=off      no implicit FMA contraction, so host (oracle) and device round
                         the same way; the shared math uses explicit fma() where wanted
  -Xclang -target-feature -Xclang -packed-fp32-ops
                         no packed FP32 instructions on the device

usage: python build.py [--force] [--jobs N]
       python build.py --model <model dir> [-o <out dir>]     (a model library)
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
GEN = os.path.join(PKG, "gen")
INC = os.path.join(REPO, "include")
OBJ = os.path.join(CSRC, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PLOM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable"]


def run(cmd, cwd=None):
    print("+ " + " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def generate():
    sys.path.insert(0, PKG)
    from codegen import generate_all  # noqa: E402
    return generate_all()


def builtin_models():
    sys.path.insert(0, PKG)
    from codegen import BUILTIN_MODELS  # noqa: E402
    return list(BUILTIN_MODELS)


def _headers():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")] + \
           [os.path.join(INC, f) for f in os.listdir(INC) if f.endswith(".h")]


def _stale(out, deps):
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    return any(os.path.getmtime(d) > t for d in deps)


def model_defines(name):
    return ["-DPLOM_MODEL_NAME=" + name, "-DPLOM_MODEL_UNAME=PLOM_" + name.upper(),
            '-DPLOM_MODEL_HEADER="plom_%s.h"' % name]


def _compile_jobs(force):
    """(object, command, dependencies) of every translation unit of the library."""
    os.makedirs(OBJ, exist_ok=True)
    hdrs = _headers()
    inc = ["-I" + CSRC, "-I" + GEN, "-I" + INC]
    jobs = []
    core_src = os.path.join(CSRC, "plom_smc.hip")
    core_o = os.path.join(OBJ, "plom_smc.o")
    jobs.append((core_o, [HIPCC] + HIP_FLAGS + inc + ["-c", core_src, "-o", core_o],
                 [core_src, os.path.join(GEN, "plom_model_list.h")] + hdrs))
    msrc = os.path.join(CSRC, "plom_model.hip")
    for name in builtin_models():
        o = os.path.join(OBJ, "model_%s.o" % name)
        jobs.append((o, [HIPCC] + HIP_FLAGS + inc + model_defines(name) + ["-c", msrc, "-o", o],
                     [msrc, os.path.join(GEN, "plom_%s.h" % name)] + hdrs))
    return [j for j in jobs if force or _stale(j[0], j[2])], [j[0] for j in jobs]


def build_lib(force=True, jobs=None):
    generate()
    todo, objs = _compile_jobs(force)
    jobs = jobs or min(8, os.cpu_count() or 1)
    if todo:
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            for r in [ex.submit(run, cmd) for _, cmd, _ in todo]:
                r.result()
    out = os.path.join(CSRC, "libplom_smc.so")
    if todo or not os.path.exists(out):
        run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-ldl", "-o", out + ".tmp"])
        os.replace(out + ".tmp", out)
    return out


def build_cli(force=True):
    cli = os.path.join(PKG, "cli")
    if not os.path.isdir(cli):
        return None
    shared = [os.path.join(cli, f) for f in ("plom_cli.c", "pjson.c", "plom_cli.h", "pjson.h")]
    lib = os.path.join(CSRC, "libplom_smc.so")
    outs = []
    for name in ("smc", "pmcmc", "mif"):
        out = os.path.join(cli, name)
        src = os.path.join(cli, name + "_main.c")
        if force or _stale(out, [src, lib] + shared + _headers()):
            cmd = ["gcc", "-O2", "-std=gnu99", "-Wall", "-ffp-contract=off", "-I" + INC, "-I" + CSRC, "-I" + GEN,
                   src] + [s for s in shared if s.endswith(".c")] + \
                  ["-L" + CSRC, "-lplom_smc", "-Wl,-rpath,$ORIGIN/../csrc", "-lm", "-o", out]
            run(cmd)
        outs.append(out)
    return outs


def build_model_lib(model_dir, out_dir=None):
    """A user's model (context / process / link JSON in model_dir) into out_dir: its
    generated header plom_<name>.h and the model library libplom_model_<name>.so, which
    plom_model_load registers at run time (the library's built-in models stay as they
    are).  Returns (name, library path)."""
    sys.path.insert(0, PKG)
    from codegen import load_model, emit_header  # noqa: E402
    m = load_model(model_dir)
    name = m.name.lower()
    out_dir = out_dir or os.path.join(model_dir, "build")
    os.makedirs(out_dir, exist_ok=True)
    hdr = os.path.join(out_dir, "plom_%s.h" % name)
    text = emit_header(m)
    if not os.path.exists(hdr) or open(hdr).read() != text:
        with open(hdr, "w") as f:
            f.write(text)
    out = os.path.join(out_dir, "libplom_model_%s.so" % name)
    run([HIPCC] + HIP_FLAGS + ["-shared", "-DPLOM_MODEL_PLUGIN", "-I" + out_dir, "-I" + CSRC, "-I" + GEN, "-I" + INC]
        + model_defines(name) + [os.path.join(CSRC, "plom_model.hip"), "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return name, out


def build_all(force=True):
    lib = build_lib(force)
    cli = build_cli(force)
    return lib, cli


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true", help="recompile every translation unit")
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--model", default=None, help="build this model directory into a model library")
    ap.add_argument("-o", "--out", default=None)
    a = ap.parse_args()
    if a.model:
        print(build_model_lib(a.model, a.out)[1])
        return
    build_lib(a.force, a.jobs)
    build_cli(a.force)


if __name__ == "__main__":
    main()
