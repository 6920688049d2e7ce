/*
 * builtin_SISe.hip -- the node kernels instantiated for the built-in
 * model SISe, and its plugin record (8 gdata values); the coupled
 * kernels are in builtin_SISe_coupled.hip.
 */
#include "siminf_builtin_models.h"

SIMINF_COUPLED_INSTANCE(extern, siminf::ModelSISe);

const SimInf_hip_model *siminf_builtin_SISe()
{
    static const SimInf_hip_model model =
        SIMINF_MODEL_ENTRY("SISe", siminf::ModelSISe, 8);
    return &model;
}
