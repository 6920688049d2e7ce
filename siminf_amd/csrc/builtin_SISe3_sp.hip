/*
 * builtin_SISe3_sp.hip -- the plugin record of the built-in spatial model
 * SISe3_sp (12 gdata values).  Its post-time-step function reads the
 * neighbours' state, so every plan of it runs the coupled kernel (in
 * builtin_SISe3_sp_coupled.hip) and the record has no node kernels.
 */
#include "siminf_builtin_models.h"

SIMINF_COUPLED_INSTANCE(extern, siminf::ModelSISe3_sp);

const SimInf_hip_model *siminf_builtin_SISe3_sp()
{
    static const SimInf_hip_model model =
        SIMINF_MODEL_ENTRY("SISe3_sp", siminf::ModelSISe3_sp, 12);
    return &model;
}
