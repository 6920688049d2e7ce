/*
 * builtin_SISe3.hip -- the node kernels instantiated for the built-in
 * model SISe3, and its plugin record (12 gdata values); the coupled
 * kernels are in builtin_SISe3_coupled.hip.
 */
#include "siminf_builtin_models.h"

SIMINF_COUPLED_INSTANCE(extern, siminf::ModelSISe3);

const SimInf_hip_model *siminf_builtin_SISe3()
{
    static const SimInf_hip_model model =
        SIMINF_MODEL_ENTRY("SISe3", siminf::ModelSISe3, 12);
    return &model;
}
