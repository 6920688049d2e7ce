/*
 * builtin_SISe3_coupled.hip -- the coupled kernels (external transfers)
 * instantiated for the built-in model SISe3.
 */
#include "siminf_builtin_models.h"

SIMINF_COUPLED_INSTANCE(, siminf::ModelSISe3);
