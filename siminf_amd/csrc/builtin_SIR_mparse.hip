/*
 * builtin_SIR_mparse.hip -- the node kernels instantiated for the built-in
 * model SIR_mparse, and its plugin record (2 gdata values); the coupled
 * kernels are in builtin_SIR_mparse_coupled.hip.
 */
#include "siminf_builtin_models.h"

SIMINF_COUPLED_INSTANCE(extern, siminf::ModelSIRmparse);

const SimInf_hip_model *siminf_builtin_SIR_mparse()
{
    static const SimInf_hip_model model =
        SIMINF_MODEL_ENTRY("SIR_mparse", siminf::ModelSIRmparse, 2);
    return &model;
}
