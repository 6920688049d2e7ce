/*
 * builtin_SISe_sp.hip -- the plugin record of the built-in spatial model
 * SISe_sp (8 gdata values).  Its post-time-step function reads the
 * neighbours' state, so every plan of it runs the coupled kernel (in
 * builtin_SISe_sp_coupled.hip) and the record has no node kernels.
 */
#include "siminf_builtin_models.h"

SIMINF_COUPLED_INSTANCE(extern, siminf::ModelSISe_sp);

const SimInf_hip_model *siminf_builtin_SISe_sp()
{
    static const SimInf_hip_model model =
        SIMINF_MODEL_ENTRY("SISe_sp", siminf::ModelSISe_sp, 8);
    return &model;
}
