/*
 * siminf_abc.h -- the arithmetic of the ABC-SMC weights (Toni et al.
 * 2009; SimInf_abc_weights, SimInf_abc.c:283-406), written once and
 * compiled for gfx950 and for the host (C99), so that the device kernel
 * and the CPU checker compute identical weights:
 *
 *   - siminf_prior_logdens: the log-density of a prior (dgamma, dlnorm,
 *     dnorm, dunif with log = TRUE);
 *   - siminf_cholesky: the lower Cholesky factor of the proposal
 *     covariance and the log normaliser of its Gaussian density;
 *   - siminf_forward_subst / siminf_mvn_logpdf_z: the log-density of the
 *     proposal kernel on points transformed by L^-1.
 */
#ifndef SIMINF_ABC_H
#define SIMINF_ABC_H

#include "siminf_rand.h"

#define SIMINF_ABC_MAX_PARS 16

/* Prior kinds: the first letter of the distribution's name. */
#define SIMINF_PRIOR_GAMMA 'g'      /* p1 shape, p2 rate */
#define SIMINF_PRIOR_LOGNORMAL 'l'  /* p1 meanlog, p2 sdlog */
#define SIMINF_PRIOR_NORMAL 'n'     /* p1 mean, p2 sd */
#define SIMINF_PRIOR_UNIFORM 'u'    /* p1 lower, p2 upper */

/* log(2 pi) / 2 */
#define SIMINF_HALF_LOG_2PI 0.91893853320467274178

/* Log-density of x under a prior.  mean replaces p1 for the normal and
 * lognormal kinds (the weights centre them on a particle of the previous
 * generation); lg is log Gamma(p1) for the gamma kind. */
SIMINF_FN double siminf_prior_logdens(int kind, double x, double mean, double p1,
                                      double p2, double lg)
{
    const double NEG_INF = -1.0 / 0.0;
    double z;
    switch (kind) {
    case SIMINF_PRIOR_GAMMA:
        if (x < 0.0)
            return NEG_INF;
        if (x == 0.0)
            return p1 < 1.0 ? 1.0 / 0.0 : (p1 == 1.0 ? siminf_log(p2) : NEG_INF);
        return (((p1 - 1.0) * siminf_log(x) - p2 * x) + p1 * siminf_log(p2)) - lg;
    case SIMINF_PRIOR_LOGNORMAL:
        if (!(x > 0.0))
            return NEG_INF;
        z = (siminf_log(x) - mean) / p2;
        return ((-0.5 * (z * z) - siminf_log(p2)) - SIMINF_HALF_LOG_2PI) - siminf_log(x);
    case SIMINF_PRIOR_NORMAL:
        z = (x - mean) / p2;
        return (-0.5 * (z * z) - siminf_log(p2)) - SIMINF_HALF_LOG_2PI;
    default:
        return (x >= p1 && x <= p2) ? -siminf_log(p2 - p1) : NEG_INF;
    }
}

/* Lower Cholesky factor L (row-major, d x d) of the symmetric matrix
 * sigma (column-major), and log((2 pi)^(d/2) |L|).  Returns 0, or -1 when
 * sigma is not positive definite. */
SIMINF_FN int siminf_cholesky(const double *sigma, int d, double *L, double *lognorm)
{
    int i, j, k;
    double s = 0.0;
    for (i = 0; i < d * d; i++)
        L[i] = 0.0;
    for (j = 0; j < d; j++) {
        double dj = sigma[j * d + j];
        for (k = 0; k < j; k++)
            dj = dj - L[j * d + k] * L[j * d + k];
        if (!(dj > 0.0))
            return -1;
        L[j * d + j] = siminf_sqrt(dj);
        for (i = j + 1; i < d; i++) {
            double v = sigma[j * d + i];
            for (k = 0; k < j; k++)
                v = v - L[i * d + k] * L[j * d + k];
            L[i * d + j] = v / L[j * d + j];
        }
    }
    for (j = 0; j < d; j++)
        s = s + siminf_log(L[j * d + j]);
    *lognorm = (double)d * SIMINF_HALF_LOG_2PI + s;
    return 0;
}

/* z = L^-1 x by forward substitution (x and z strided).  The proposal
 * kernel's log-density at xx for mean x is -|L^-1 xx - L^-1 x|^2 / 2 -
 * lognorm: both points are transformed once, then every pair costs d
 * subtractions and squares. */
SIMINF_FN void siminf_forward_subst(const double *x, long long sx, const double *L, int d,
                                    double *z, long long sz)
{
    int i, k;
    for (i = 0; i < d; i++) {
        double v = x[i * sx];
        for (k = 0; k < i; k++)
            v = v - L[i * d + k] * z[k * sz];
        z[i * sz] = v / L[i * d + i];
    }
}

/* -|a - b|^2 / 2 - lognorm for transformed points a, b (strided). */
SIMINF_FN double siminf_mvn_logpdf_z(const double *a, long long sa, const double *b,
                                     long long sb, int d, double lognorm)
{
    double q = 0.0;
    int k;
    for (k = 0; k < d; k++) {
        const double t = a[k * sa] - b[k * sb];
        q = q + t * t;
    }
    return -0.5 * q - lognorm;
}

#endif /* SIMINF_ABC_H */
