/*
 * siminf_models.h -- the built-in model plugins (transition rate functions
 * and post-time-step functions) in the reference's plugin signatures
 *
 *   double TRFun(const int *u, const double *v, const double *ldata,
 *                const double *gdata, double t);          (SimInf.h:92-97)
 *   int PTSFun(double *v_new, const int *u, const double *v,
 *              const double *ldata, const double *gdata,
 *              int node, double t);                      (SimInf.h:100-107)
 *
 * compiled for both gfx950 and the host.  Every expression keeps the
 * reference's operand order so that the double results are identical on
 * the CPU and the GPU (no contraction: -ffp-contract=off).
 *
 * Models: SIR (src/models/SIR.c), SEIR (src/models/SEIR.c), SIS
 * (src/models/SIS.c), SISe (src/models/SISe.c), SISe3 (src/models/SISe3.c),
 * SISe_sp and SISe3_sp (src/models/SISe_sp.c, SISe3_sp.c, with
 * src/misc/SimInf_local_spread.c) and the model that mparse emits for
 *     transitions = c("S -> beta*S*I/N -> I", "I -> gamma*I -> R",
 *                     "N <- S+I+R"), gdata = c(beta, gamma)
 * (R/mparse.R:647-724, R/C-generator.R:170-297), named "SIR_mparse".
 */
#ifndef SIMINF_MODELS_H
#define SIMINF_MODELS_H

#include "siminf_rand.h"

/* Error codes of the plugin contract (SimInf.h:66-89). */
#define SIMINF_MODEL_ERR_V_IS_NOT_FINITE (-7)
#define SIMINF_MODEL_ERR_V_IS_NEGATIVE   (-10)

SIMINF_FN int siminf_isfinite(double x)
{
    return (siminf_bits(x) & 0x7FF0000000000000ull) != 0x7FF0000000000000ull;
}

/* ---------------------------------------------------------------- SIR */
/* u = (S, I, R); ldata = (beta, gamma).  SIR.c:43-87. */
SIMINF_FN double SIR_S_to_I(const int *u, const double *v, const double *ldata,
                            const double *gdata, double t)
{
    const double S_n = u[0];
    const double I_n = u[1];
    const double n = S_n + I_n + u[2];
    (void)v; (void)gdata; (void)t;
    {
        /* The quotient is formed unconditionally and selected, so that
         * the rate is straight-line code; for n > 0 it is the value of
         * the guarded division. */
        const double r = (ldata[0] * S_n * I_n) / n;
        return n > 0.0 ? r : 0.0;
    }
}

SIMINF_FN double SIR_I_to_R(const int *u, const double *v, const double *ldata,
                            const double *gdata, double t)
{
    (void)v; (void)gdata; (void)t;
    return ldata[1] * u[1];
}

/* Post time step of SIR, SEIR, SIS and mparse models without a pts_fun:
 * no continuous state, never requests a rate refresh. */
SIMINF_FN int siminf_pts_noop(double *v_new, const int *u, const double *v,
                              const double *ldata, const double *gdata,
                              int node, double t)
{
    (void)v_new; (void)u; (void)v; (void)ldata; (void)gdata; (void)node;
    (void)t;
    return 0;
}

/* --------------------------------------------------------------- SEIR */
/* u = (S, E, I, R); ldata = (beta, epsilon, gamma).  SEIR.c:43-112. */
SIMINF_FN double SEIR_S_to_E(const int *u, const double *v, const double *ldata,
                             const double *gdata, double t)
{
    const double S_n = u[0];
    const double I_n = u[2];
    const double n = S_n + u[1] + I_n + u[3];
    (void)v; (void)gdata; (void)t;
    {
        /* Formed unconditionally and selected (as SIR_S_to_I). */
        const double r = ldata[0] * S_n * I_n / n;
        return n > 0.0 ? r : 0.0;
    }
}

SIMINF_FN double SEIR_E_to_I(const int *u, const double *v, const double *ldata,
                             const double *gdata, double t)
{
    (void)v; (void)gdata; (void)t;
    return ldata[1] * u[1];
}

SIMINF_FN double SEIR_I_to_R(const int *u, const double *v, const double *ldata,
                             const double *gdata, double t)
{
    (void)v; (void)gdata; (void)t;
    return ldata[2] * u[2];
}

/* ---------------------------------------------------------------- SIS */
/* u = (S, I); ldata = (beta, gamma).  SIS.c:41-84. */
SIMINF_FN double SIS_S_to_I(const int *u, const double *v, const double *ldata,
                            const double *gdata, double t)
{
    const double S_n = u[0];
    const double I_n = u[1];
    const double n = S_n + I_n;
    (void)v; (void)gdata; (void)t;
    {
        /* The quotient is formed unconditionally and selected, so that
         * the rate is straight-line code; for n > 0 it is the value of
         * the guarded division. */
        const double r = (ldata[0] * S_n * I_n) / n;
        return n > 0.0 ? r : 0.0;
    }
}

SIMINF_FN double SIS_I_to_S(const int *u, const double *v, const double *ldata,
                            const double *gdata, double t)
{
    (void)v; (void)gdata; (void)t;
    return ldata[1] * u[1];
}

/* ------------------------------------------- environmental pressure */
/* Seasonal decay phi * (1 - beta_interval(day)) with four intervals of
 * the year; SimInf_forward_euler_linear_decay.c:56-93.  The interval
 * ends arrive as doubles in ldata and are converted to int like the
 * reference's int parameters. */
SIMINF_FN double siminf_euler_linear_decay(
    double phi, int day, int end_t1, int end_t2, int end_t3, int end_t4,
    double beta_t1, double beta_t2, double beta_t3, double beta_t4)
{
    double beta;
    if (day < end_t2) {
        if (day < end_t1) {
            if (end_t1 < end_t4)
                beta = beta_t1;
            else if (day < end_t4)
                beta = (end_t4 < end_t3 || day >= end_t3) ? beta_t4 : beta_t3;
            else
                beta = beta_t1;
        } else {
            beta = beta_t2;
        }
    } else if (end_t3 < end_t1 || day < end_t3) {
        beta = beta_t3;
    } else if (end_t4 < end_t1 || day < end_t4) {
        beta = beta_t4;
    } else {
        beta = beta_t1;
    }
    return phi * (1.0 - beta);
}

/* --------------------------------------------------------------- SISe */
/* u = (S, I); v = (phi); ldata = (end_t1..end_t4);
 * gdata = (upsilon, gamma, alpha, beta_t1..beta_t4, epsilon).
 * SISe.c:43-144. */
SIMINF_FN double SISe_S_to_I(const int *u, const double *v, const double *ldata,
                             const double *gdata, double t)
{
    (void)ldata; (void)t;
    return gdata[0] * v[0] * u[0];
}

SIMINF_FN double SISe_I_to_S(const int *u, const double *v, const double *ldata,
                             const double *gdata, double t)
{
    (void)v; (void)ldata; (void)t;
    return gdata[1] * u[1];
}

SIMINF_FN int SISe_pts(double *v_new, const int *u, const double *v,
                       const double *ldata, const double *gdata,
                       int node, double t)
{
    const int day = (int)t % 365;
    const double I_n = u[1];
    const double n = u[0] + I_n;
    const double phi = v[0];
    (void)node;
    v_new[0] = siminf_euler_linear_decay(phi, day,
                                         (int)ldata[0], (int)ldata[1],
                                         (int)ldata[2], (int)ldata[3],
                                         gdata[3], gdata[4], gdata[5],
                                         gdata[6]);
    if (n > 0.0)
        v_new[0] += gdata[2] * I_n / n + gdata[7];
    else
        v_new[0] += gdata[7];
    if (!siminf_isfinite(v_new[0]))
        return SIMINF_MODEL_ERR_V_IS_NOT_FINITE;
    if (v_new[0] < 0.0)
        return SIMINF_MODEL_ERR_V_IS_NEGATIVE;
    return phi != v_new[0];
}

/* -------------------------------------------------------------- SISe3 */
/* u = (S_1, I_1, S_2, I_2, S_3, I_3); v = (phi);
 * ldata = (end_t1..end_t4);
 * gdata = (upsilon_1..3, gamma_1..3, alpha, beta_t1..beta_t4, epsilon).
 * SISe3.c:51-242; transition order S_1->I_1, I_1->S_1, S_2->I_2, I_2->S_2,
 * S_3->I_3, I_3->S_3 (SISe3.c:255-258). */
SIMINF_FN double SISe3_S_1_to_I_1(const int *u, const double *v,
                                  const double *ldata, const double *gdata,
                                  double t)
{
    (void)ldata; (void)t;
    return gdata[0] * v[0] * u[0];
}

SIMINF_FN double SISe3_I_1_to_S_1(const int *u, const double *v,
                                  const double *ldata, const double *gdata,
                                  double t)
{
    (void)v; (void)ldata; (void)t;
    return gdata[3] * u[1];
}

SIMINF_FN double SISe3_S_2_to_I_2(const int *u, const double *v,
                                  const double *ldata, const double *gdata,
                                  double t)
{
    (void)ldata; (void)t;
    return gdata[1] * v[0] * u[2];
}

SIMINF_FN double SISe3_I_2_to_S_2(const int *u, const double *v,
                                  const double *ldata, const double *gdata,
                                  double t)
{
    (void)v; (void)ldata; (void)t;
    return gdata[4] * u[3];
}

SIMINF_FN double SISe3_S_3_to_I_3(const int *u, const double *v,
                                  const double *ldata, const double *gdata,
                                  double t)
{
    (void)ldata; (void)t;
    return gdata[2] * v[0] * u[4];
}

SIMINF_FN double SISe3_I_3_to_S_3(const int *u, const double *v,
                                  const double *ldata, const double *gdata,
                                  double t)
{
    (void)v; (void)ldata; (void)t;
    return gdata[5] * u[5];
}

SIMINF_FN int SISe3_pts(double *v_new, const int *u, const double *v,
                        const double *ldata, const double *gdata,
                        int node, double t)
{
    const int day = (int)t % 365;
    const double I_n = u[1] + u[3] + u[5];
    const double n = I_n + u[0] + u[2] + u[4];
    const double phi = v[0];
    (void)node;
    v_new[0] = siminf_euler_linear_decay(phi, day,
                                         (int)ldata[0], (int)ldata[1],
                                         (int)ldata[2], (int)ldata[3],
                                         gdata[7], gdata[8], gdata[9],
                                         gdata[10]);
    if (n > 0.0)
        v_new[0] += gdata[6] * I_n / n + gdata[11];
    else
        v_new[0] += gdata[11];
    if (!siminf_isfinite(v_new[0]))
        return SIMINF_MODEL_ERR_V_IS_NOT_FINITE;
    if (v_new[0] < 0.0)
        return SIMINF_MODEL_ERR_V_IS_NEGATIVE;
    return phi != v_new[0];
}

/* ----------------------------------------------- spatial models (_sp) */
/* The neighbours' share of a node's phi (SimInf_local_spread.c:45-75).
 * nb lists (index, distance) pairs of replicate-local node indices and
 * ends at an index below zero; phi0 and u0 point at the first node's
 * continuous and compartment state, so that node j's sit at phi0[j] and
 * u0[j * Nc].  A neighbour without individuals adds nothing. */
SIMINF_FN double siminf_local_spread(const double *nb, const double *phi0,
                                     const int *u0, double N_i, double phi_i,
                                     int Nc, double D)
{
    const double own = phi_i * N_i;
    double ls = 0.0;
    for (int j = (int)nb[0]; j >= 0; nb += 2, j = (int)nb[0]) {
        double N_j = 0;
        for (int k = 0; k < Nc; k++)
            N_j += u0[j * Nc + k];
        if (N_j > 0.0)
            ls += ((phi0[j] * N_j - own) * D) / (N_i * nb[1]);
    }
    return ls;
}

/* SISe_sp: u = (S, I); v = (phi); ldata = (end_t1..end_t4, then the
 * neighbour pairs); gdata = (upsilon, gamma, alpha, beta_t1..beta_t4,
 * coupling).  The rates are SISe's (SISe_sp.c:43-75); the post-time-step
 * function adds the local spread and reads the other nodes' u and v
 * through negative offsets from its own (SISe_sp.c:90-138), so u and v
 * must point into the replicate's node-major arrays. */
SIMINF_FN int SISe_sp_pts(double *v_new, const int *u, const double *v,
                          const double *ldata, const double *gdata,
                          int node, double t)
{
    const int day = (int)t % 365;
    const double I_i = u[1];
    const double N_i = u[0] + I_i;
    const double phi = v[0];
    v_new[0] = siminf_euler_linear_decay(phi, day,
                                         (int)ldata[0], (int)ldata[1],
                                         (int)ldata[2], (int)ldata[3],
                                         gdata[3], gdata[4], gdata[5],
                                         gdata[6]);
    if (N_i > 0.0)
        v_new[0] += gdata[2] * I_i / N_i +
                    siminf_local_spread(ldata + 4, v - node, u - 2 * node,
                                        N_i, phi, 2, gdata[7]);
    if (!siminf_isfinite(v_new[0]))
        return SIMINF_MODEL_ERR_V_IS_NOT_FINITE;
    if (v_new[0] < 0.0)
        return SIMINF_MODEL_ERR_V_IS_NEGATIVE;
    return phi != v_new[0];
}

/* SISe3_sp: SISe3's compartments and rates; gdata = (upsilon_1..3,
 * gamma_1..3, alpha, beta_t1..beta_t4, coupling); ldata as SISe_sp's
 * (SISe3_sp.c:206-255). */
SIMINF_FN int SISe3_sp_pts(double *v_new, const int *u, const double *v,
                           const double *ldata, const double *gdata,
                           int node, double t)
{
    const int day = (int)t % 365;
    const double I_i = u[1] + u[3] + u[5];
    const double N_i = u[0] + u[2] + u[4] + I_i;
    const double phi = v[0];
    v_new[0] = siminf_euler_linear_decay(phi, day,
                                         (int)ldata[0], (int)ldata[1],
                                         (int)ldata[2], (int)ldata[3],
                                         gdata[7], gdata[8], gdata[9],
                                         gdata[10]);
    if (N_i > 0.0)
        v_new[0] += gdata[6] * I_i / N_i +
                    siminf_local_spread(ldata + 4, v - node, u - 6 * node,
                                        N_i, phi, 6, gdata[11]);
    if (!siminf_isfinite(v_new[0]))
        return SIMINF_MODEL_ERR_V_IS_NOT_FINITE;
    if (v_new[0] < 0.0)
        return SIMINF_MODEL_ERR_V_IS_NEGATIVE;
    return phi != v_new[0];
}

/* --------------------------------------------------------- SIR_mparse */
/* What mparse emits for "S -> beta*S*I/N -> I", "I -> gamma*I -> R",
 * "N <- S+I+R" with compartments (S, I, R) and gdata (beta, gamma):
 * the variable N is declared "const double N = u[0]+u[1]+u[2];" and the
 * propensities are "return gdata[0]*u[0]*u[1]/N;" and
 * "return gdata[1]*u[1];" (R/C-generator.R:170-247). */
SIMINF_FN double SIR_mparse_trFun1(const int *u, const double *v,
                                   const double *ldata, const double *gdata,
                                   double t)
{
    const double N = u[0]+u[1]+u[2];
    (void)v; (void)ldata; (void)t;
    return gdata[0]*u[0]*u[1]/N;
}

SIMINF_FN double SIR_mparse_trFun2(const int *u, const double *v,
                                   const double *ldata, const double *gdata,
                                   double t)
{
    (void)v; (void)ldata; (void)t;
    return gdata[1]*u[1];
}

#endif /* SIMINF_MODELS_H */
