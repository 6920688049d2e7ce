/*
 * builtin_SEIR.hip -- the node kernels instantiated for the built-in
 * model SEIR, and its plugin record (0 gdata values); the coupled
 * kernels are in builtin_SEIR_coupled.hip.
 */
#include "siminf_builtin_models.h"

SIMINF_COUPLED_INSTANCE(extern, siminf::ModelSEIR);

const SimInf_hip_model *siminf_builtin_SEIR()
{
    static const SimInf_hip_model model =
        SIMINF_MODEL_ENTRY("SEIR", siminf::ModelSEIR, 0);
    return &model;
}
