/*
 * builtin_SIR.hip -- the node kernels instantiated for the built-in
 * model SIR, and its plugin record (0 gdata values); the coupled
 * kernels are in builtin_SIR_coupled.hip.
 */
#include "siminf_builtin_models.h"

SIMINF_COUPLED_INSTANCE(extern, siminf::ModelSIR);

const SimInf_hip_model *siminf_builtin_SIR()
{
    static const SimInf_hip_model model =
        SIMINF_MODEL_ENTRY("SIR", siminf::ModelSIR, 0);
    return &model;
}
