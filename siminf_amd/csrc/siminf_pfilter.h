/*
 * siminf_pfilter.h -- the arithmetic of the bootstrap particle filter
 * (R/pfilter.R:315-499, SimInf_particle_filter.c:32-66), written once and
 * compiled for gfx950 and for the host (C99), so that the device filter
 * and the CPU checker compute identical weights, log-likelihoods and
 * ancestors:
 *
 *   - siminf_exp: exp with a fixed sequence of IEEE operations (the
 *     log-sum-exp of the weights, pfilter.R:468-470);
 *   - siminf_lfact / siminf_log_dbinom / siminf_log_dpois: log n! and the
 *     log-probabilities of the observation processes (stats::dbinom and
 *     stats::dpois with log = TRUE, pfilter.R:279-295);
 *   - siminf_interval_seed: the seed of interval i's run (the reference
 *     draws a new seed for every run() from R's RNG);
 *   - siminf_systematic_resampling: SimInf_particle_filter.c:32-66.
 */
#ifndef SIMINF_PFILTER_H
#define SIMINF_PFILTER_H

#include "siminf_rand.h"

#define SIMINF_DOMAIN_PF 3u     /* stream of interval i's resampling draw */

/* exp(x) by x = k ln2 + r, |r| <= ln2/2, and a degree-13 Taylor sum of
 * exp(r) (truncation below 2^-56); 2^k is applied through the exponent
 * bits, in two steps below the normal range. */
SIMINF_FN double siminf_exp(double x)
{
    const double LOG2E = 1.4426950408889634;
    const double LN2_HI = 6.93147180369123816490e-01;   /* 32 significant bits */
    const double LN2_LO = 1.90821492927058770002e-10;
    double kd, r, p;
    long long k;

    if (x != x)
        return x;
    if (x > 709.78)
        return 1.0 / 0.0;
    if (x < -745.2)
        return 0.0;
    kd = x * LOG2E;
    k = (long long)(kd >= 0.0 ? kd + 0.5 : kd - 0.5);
    kd = (double)k;
    r = (x - kd * LN2_HI) - kd * LN2_LO;
    p = 1.0 / 6227020800.0;                 /* 1/13! */
    p = 1.0 / 479001600.0 + r * p;          /* 1/12! */
    p = 1.0 / 39916800.0 + r * p;
    p = 1.0 / 3628800.0 + r * p;
    p = 1.0 / 362880.0 + r * p;
    p = 1.0 / 40320.0 + r * p;
    p = 1.0 / 5040.0 + r * p;
    p = 1.0 / 720.0 + r * p;
    p = 1.0 / 120.0 + r * p;
    p = 1.0 / 24.0 + r * p;
    p = 1.0 / 6.0 + r * p;
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    if (k > 1020)
        return (p * siminf_from_bits((uint64_t)(k - 2 + 1023) << 52)) * 4.0;
    if (k >= -1021)
        return p * siminf_from_bits((uint64_t)(k + 1023) << 52);
    return (p * siminf_from_bits((uint64_t)(k + 1023 + 64) << 52)) * 0x1p-64;
}

/* log(n!): exact values below 16, Stirling's series for log Gamma(n+1)
 * above (absolute error below 1e-14). */
SIMINF_FN double siminf_lfact(int n)
{
    const double HALF_LOG_2PI = 0.91893853320467274178;
    double m, im, im2;
    if (n < 16) {
        switch (n) {
        case 0: case 1: return 0.0;
        case 2: return 0.6931471805599453;
        case 3: return 1.791759469228055;
        case 4: return 3.1780538303479458;
        case 5: return 4.787491742782046;
        case 6: return 6.579251212010101;
        case 7: return 8.525161361065415;
        case 8: return 10.60460290274525;
        case 9: return 12.801827480081469;
        case 10: return 15.104412573075516;
        case 11: return 17.502307845873887;
        case 12: return 19.987214495661885;
        case 13: return 22.552163853123425;
        case 14: return 25.19122118273868;
        default: return 27.89927138384089;
        }
    }
    m = (double)n + 1.0;
    im = 1.0 / m;
    im2 = im * im;
    return (m - 0.5) * siminf_log(m) - m + HALF_LOG_2PI +
           im * (1.0 / 12.0 - im2 * (1.0 / 360.0 - im2 * (1.0 / 1260.0 -
                                                         im2 * (1.0 / 1680.0))));
}

/* log P(X = x) for X ~ Binomial(n, p); -inf outside the support. */
SIMINF_FN double siminf_log_dbinom(int x, int n, double p)
{
    const double NEG_INF = -1.0 / 0.0;
    if (x < 0 || n < 0 || x > n)
        return NEG_INF;
    if (p <= 0.0)
        return x == 0 ? 0.0 : NEG_INF;
    if (p >= 1.0)
        return x == n ? 0.0 : NEG_INF;
    return (siminf_lfact(n) - siminf_lfact(x) - siminf_lfact(n - x)) +
           ((double)x * siminf_log(p) + (double)(n - x) * siminf_log(1.0 - p));
}

/* log P(X = x) for X ~ Poisson(lambda); lambda <= 0 puts all mass on 0. */
SIMINF_FN double siminf_log_dpois(int x, double lambda)
{
    const double NEG_INF = -1.0 / 0.0;
    if (x < 0)
        return NEG_INF;
    if (!(lambda > 0.0))
        return x == 0 ? 0.0 : NEG_INF;
    return ((double)x * siminf_log(lambda) - lambda) - siminf_lfact(x);
}

/* Observation processes of the filter: the log-weight of observing x
 * when the particle holds X individuals in the observed compartment. */
#define SIMINF_OBS_BINOMIAL 0   /* x ~ Binomial(size = X, prob = p1) */
#define SIMINF_OBS_POISSON 1    /* x ~ Poisson(lambda = p1 * X + p2) */

SIMINF_FN double siminf_obs_logw(int dist, int x, long long X, double p1, double p2)
{
    if (dist == SIMINF_OBS_POISSON)
        return siminf_log_dpois(x, p1 * (double)X + p2);
    return siminf_log_dbinom(x, (int)X, p1);
}

/* Seed of interval i's run: splitmix64 of seed + i. */
SIMINF_FN uint64_t siminf_interval_seed(uint64_t seed, int i)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Systematic resampling of n particles with weights w (non-negative,
 * positive sum) from one uniform u01 in (0, 1); idx gets zero-based
 * ancestors.  Returns 0, or -1 for an invalid weight, -2 for a
 * non-positive sum (SimInf_particle_filter.c:32-66). */
SIMINF_FN int siminf_systematic_resampling(const double *w, int n, double u01,
                                           int *idx)
{
    double cumsum = 0.0, du, u;
    int i, j;
    for (i = 0; i < n; i++) {
        if (!(w[i] >= 0.0) || w[i] == 1.0 / 0.0)
            return -1;
        cumsum += w[i];
    }
    if (cumsum <= 0.0)
        return -2;
    du = cumsum / (double)n;
    u = du * u01;
    cumsum = w[0];
    for (i = 0, j = 0; i < n; i++) {
        while (u > cumsum && j < n - 1)
            cumsum += w[++j];
        idx[i] = j;
        u += du;
    }
    return 0;
}

#endif /* SIMINF_PFILTER_H */
