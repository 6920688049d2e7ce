/*
 * builtin_SIR_mparse_coupled.hip -- the coupled kernels (external transfers)
 * instantiated for the built-in model SIR_mparse.
 */
#include "siminf_builtin_models.h"

SIMINF_COUPLED_INSTANCE(, siminf::ModelSIRmparse);
