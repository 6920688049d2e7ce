/*
 * siminf_plugin_api.h -- what model code written for SimInf.h can use
 * inside a generated plugin: the error codes (SimInf.h:66-89), the
 * SIMINF_UNUSED and R_FINITE helpers, and the seasonal decay helper
 * SimInf_forward_euler_linear_decay (SimInf.h:117-147).  Compiles as C
 * (host plugin for the checker) and as HIP (device plugin).
 */
#ifndef SIMINF_PLUGIN_API_H
#define SIMINF_PLUGIN_API_H

#include "siminf_models.h"

#ifndef SIMINF_UNUSED
#  define SIMINF_UNUSED(x) ((void)(x))
#endif
#ifndef R_FINITE
#  define R_FINITE(x) siminf_isfinite(x)
#endif

/* The error codes, unless the library header already declared them. */
#ifndef SIMINF_HIP_H
enum {
    SIMINF_ERR_NEGATIVE_STATE = -1,
    SIMINF_ERR_V_IS_NOT_FINITE = -7,
    SIMINF_ERR_V_IS_NEGATIVE = -10,
    SIMINF_ERR_INVALID_RATE = -11
};
#endif

SIMINF_FN double SimInf_forward_euler_linear_decay(
    double phi, int day, int end_t1, int end_t2, int end_t3, int end_t4,
    double beta_t1, double beta_t2, double beta_t3, double beta_t4)
{
    return siminf_euler_linear_decay(phi, day, end_t1, end_t2, end_t3, end_t4,
                                     beta_t1, beta_t2, beta_t3, beta_t4);
}

#endif /* SIMINF_PLUGIN_API_H */
