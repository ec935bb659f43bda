/*
 * plom_model.hip -- the kernels of one generated model, in a translation unit of its own.
 *
 * Compiled once per model with
 *   -DPLOM_MODEL_NAME=<name> -DPLOM_MODEL_UNAME=PLOM_<NAME> -DPLOM_MODEL_HEADER='"plom_<name>.h"'
 * it instantiates the model-dependent kernels of plom_device.h (k_propagate for every
 * precision and implementation variant, k_lineage, k_hat_tiles, k_pred_res, k_hat0) and
 * exports plom_model_entry_<name>(), which the library lists among its built-in models.
 * With -DPLOM_MODEL_PLUGIN it is built into a shared library of its own (a user's model,
 * plom_build.py --lib) that also exports plom_model_plugin(), the symbol plom_model_load
 * looks up: the role of the reference's per-model build, where the generated
 * core_template.c is compiled into libplomtpl.a and linked into each command
 * (model_builder/C/templates/Makefile:28-29).
 */
#include "plom_device.h"
#include PLOM_MODEL_HEADER

PLOM_DEFINE_MODEL(PLOM_MODEL_NAME, PLOM_MODEL_UNAME)

#ifdef PLOM_MODEL_PLUGIN
extern "C" const model_entry *plom_model_plugin(void) { return PLOM_CAT(plom_model_entry_, PLOM_MODEL_NAME)(); }
#endif
