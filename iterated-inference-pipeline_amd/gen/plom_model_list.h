/* Generated: the models compiled into libplom_smc, in model-id order. Do not edit. */
#ifndef PLOM_MODEL_LIST_H
#define PLOM_MODEL_LIST_H
#define PLOM_BUILTIN_MODELS(X) X(sir) X(sirs4) X(seir10) X(noise) X(drift)
#endif
