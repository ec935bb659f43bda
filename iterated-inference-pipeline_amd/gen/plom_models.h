/* Generated: the models compiled into libplom_smc. Do not edit. */
#ifndef PLOM_MODELS_H
#define PLOM_MODELS_H
#include "plom_sir.h"
#include "plom_sirs4.h"
#include "plom_seir10.h"
#include "plom_noise.h"
#include "plom_drift.h"
#define PLOM_N_MODELS 5
#endif
