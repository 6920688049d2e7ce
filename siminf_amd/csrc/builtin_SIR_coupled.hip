/*
 * builtin_SIR_coupled.hip -- the coupled kernels (external transfers)
 * instantiated for the built-in model SIR.
 */
#include "siminf_builtin_models.h"

SIMINF_COUPLED_INSTANCE(, siminf::ModelSIR);
