#!/usr/bin/env python3
"""Model builder command: context.json + process.json + link.json -> .settings.json.

The role of the reference's `pmbuilder` (scripts/pmbuilder:50-105 -> PlomModelBuilder,
model_builder/Builder.py:70-351): it writes the settings file the `smc` command reads
(schema of Builder.make_settings_json, Builder.py:173-259).  A model compiled into
libplom_smc (codegen.BUILTIN_MODELS) needs nothing else; any other model is compiled with
--lib into a model library of its own, libplom_model_<name>.so in out_dir (its generated
code: the role of Builder.py:283-351 and C/templates/Makefile:28-29), and the settings
name it ("model_library"), so that the commands register it (plom_model_load) when they
read the settings.

usage: python plom_build.py <model dir with context/process/link.json> [-o out_dir] [--lib]
"""
import argparse
import datetime
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def make_settings(model_dir):
    sys.path.insert(0, HERE)
    from codegen import load_model, context_data, load_json  # noqa: E402
    m = load_model(model_dir)
    ctx = load_json(os.path.join(model_dir, "context.json"))
    dates, times, data, date_0 = context_data(m, ctx)
    # covariates: days since date_0 (Builder.py:194-220)
    pfv = {}
    for meta in ctx.get("metadata", []):
        order = m.ts_id if meta["id"] in m.par_fixed_obs else m.cac_id
        vals = []
        for k in order:
            src = meta["source"].get(k)
            if src is None:
                continue
            xs, ys = [], []
            for d, v in src["value"]:
                if v is not None:
                    xs.append((datetime.datetime.strptime(d, "%Y-%m-%d").date() - date_0).days)
                    ys.append(v)
            obj = {"id": k, "x": xs, "y": ys, "size": len(ys)}
            for extra in ("unit", "type"):
                if extra in meta:
                    obj[extra] = meta[extra]
            vals.append(obj)
        pfv[meta["id"]] = vals
    s = m.settings(data=data, dates=dates, date_0=date_0, times=times, par_fixed_values=pfv)
    s["par_fixed_obs"] = list(m.par_fixed_obs)
    if m.coupling is not None:
        s["coupling"] = {"matrix": m.coupling}
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("model_dir")
    ap.add_argument("-o", "--out", default=None)
    ap.add_argument("--lib", action="store_true",
                    help="compile the model into its own model library in out_dir (models outside the library)")
    ap.add_argument("--model_library", default=None,
                    help="name an already built model library of this model in the settings instead")
    a = ap.parse_args()
    out = a.out or a.model_dir
    s = make_settings(a.model_dir)
    os.makedirs(out, exist_ok=True)
    sys.path.insert(0, HERE)
    from codegen import BUILTIN_MODELS  # noqa: E402
    if a.lib:
        import build  # noqa: E402
        name, path = build.build_model_lib(a.model_dir, out)
        s["model_library"] = os.path.basename(path)
    elif a.model_library:
        s["model_library"] = os.path.abspath(a.model_library)
    elif s["model"] not in BUILTIN_MODELS:
        sys.exit("model %s is not compiled into libplom_smc: build it with --lib" % s["model"])
    with open(os.path.join(out, ".settings.json"), "w") as f:
        json.dump(s, f)
    print("wrote %s (model %s)" % (os.path.join(out, ".settings.json"), s["model"]))


if __name__ == "__main__":
    main()
