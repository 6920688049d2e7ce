/*
 * builtin_SISe_sp_coupled.hip -- the coupled kernels instantiated for the
 * built-in spatial model SISe_sp.
 */
#include "siminf_builtin_models.h"

SIMINF_COUPLED_INSTANCE(, siminf::ModelSISe_sp);
