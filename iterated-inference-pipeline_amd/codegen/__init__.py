"""Model builder: PLOM model JSON (context / process / link) -> generated step code.

The role of the reference's model_builder (Python 2 + SymPy + Django templates:
model_builder/Cmodel.py, Ccoder.py, Builder.py, C/templates/core_template.c), rewritten
in Python 3 with its own expression parser so that one expression tree is printed both
as host C (for the CPU oracle) and as HIP device code (for gfx950).
"""
import json
import os

from .model import PlomModel, ModelError, context_data  # noqa: F401
from .emit import emit_header, VARIANTS  # noqa: F401

EXAMPLES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "examples")
GEN_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gen")

# models compiled into the library, in model-id order
BUILTIN_MODELS = ["sir", "sirs4", "seir10", "noise", "drift"]


def load_json(path):
    with open(path) as f:
        return json.load(f)


def load_model(name_or_dir):
    d = name_or_dir if os.path.isdir(name_or_dir) else os.path.join(EXAMPLES, name_or_dir)
    c = load_json(os.path.join(d, "context.json"))
    p = load_json(os.path.join(d, "process.json"))
    l = load_json(os.path.join(d, "link.json"))
    return PlomModel(c, p, l, name=p.get("name", os.path.basename(d)))


def _write_if_changed(path, text):
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)


def generate_all(out_dir=GEN_DIR, names=BUILTIN_MODELS):
    """Write gen/plom_<name>.h for every built-in model, gen/plom_models.h that includes
    them (the CPU oracle) and gen/plom_model_list.h, the list the library registers its
    built-in models from (one translation unit per model); returns the written paths."""
    os.makedirs(out_dir, exist_ok=True)
    written = []
    for n in names:
        m = load_model(n)
        text = emit_header(m)
        path = os.path.join(out_dir, "plom_%s.h" % m.name.lower())
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as f:
                f.write(text)
        written.append(path)
    inc = ["/* Generated: the models compiled into libplom_smc. Do not edit. */",
           "#ifndef PLOM_MODELS_H", "#define PLOM_MODELS_H"]
    for n in names:
        inc.append('#include "plom_%s.h"' % n)
    inc.append("#define PLOM_N_MODELS %d" % len(names))
    inc.append("#endif")
    path = os.path.join(out_dir, "plom_models.h")
    _write_if_changed(path, "\n".join(inc) + "\n")
    written.append(path)
    lst = ["/* Generated: the models compiled into libplom_smc, in model-id order. Do not edit. */",
           "#ifndef PLOM_MODEL_LIST_H", "#define PLOM_MODEL_LIST_H",
           "#define PLOM_BUILTIN_MODELS(X) " + " ".join("X(%s)" % n for n in names), "#endif"]
    path = os.path.join(out_dir, "plom_model_list.h")
    _write_if_changed(path, "\n".join(lst) + "\n")
    written.append(path)
    return written
