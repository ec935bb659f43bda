"""Build libplom_smc.so (HIP, gfx950) and the `smc` command from source, in-tree.

Compiler flags that matter for results:
  -ffp-contract=off      no implicit FMA contraction, so host (oracle) and device round
                         the same way; the shared math uses explicit fma() where wanted
  -Xclang -target-feature -Xclang -packed-fp32-ops
                         no packed FP32 instructions on the device
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
GEN = os.path.join(PKG, "gen")
INC = os.path.join(REPO, "include")
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


def build_lib(verbose=True):
    generate()
    out = os.path.join(CSRC, "libplom_smc.so")
    src = os.path.join(CSRC, "plom_smc.hip")
    cmd = [HIPCC] + HIP_FLAGS + ["-shared", "-I" + CSRC, "-I" + GEN, "-I" + INC, src, "-o", out + ".tmp"]
    run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_cli():
    cli = os.path.join(PKG, "cli")
    if not os.path.isdir(cli):
        return None
    shared = [os.path.join(cli, f) for f in ("plom_cli.c", "pjson.c")]
    outs = []
    for name in ("smc", "pmcmc", "mif"):
        out = os.path.join(cli, name)
        cmd = ["gcc", "-O2", "-std=gnu99", "-Wall", "-ffp-contract=off", "-I" + INC, "-I" + CSRC, "-I" + GEN,
               os.path.join(cli, name + "_main.c")] + shared + \
              ["-L" + CSRC, "-lplom_smc", "-Wl,-rpath,$ORIGIN/../csrc", "-lm", "-o", out]
        run(cmd)
        outs.append(out)
    return outs


def build_all():
    lib = build_lib()
    cli = build_cli()
    return lib, cli


if __name__ == "__main__":
    build_all()
