/*
 * builtin_SIS.hip -- the node kernels instantiated for the built-in
 * model SIS, and its plugin record (0 gdata values); the coupled
 * kernels are in builtin_SIS_coupled.hip.
 */
#include "siminf_builtin_models.h"

SIMINF_COUPLED_INSTANCE(extern, siminf::ModelSIS);

const SimInf_hip_model *siminf_builtin_SIS()
{
    static const SimInf_hip_model model =
        SIMINF_MODEL_ENTRY("SIS", siminf::ModelSIS, 0);
    return &model;
}
