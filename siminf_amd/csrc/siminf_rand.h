/*
 * siminf_rand.h -- random streams and the univariate variate generators of
 * the solver core, written once and compiled both for gfx950 (hipcc) and
 * for the host (gcc, C99).
 *
 * The reference draws every random number from GSL (gsl_rng_uniform_pos,
 * gsl_ran_binomial, gsl_ran_hypergeometric: SimInf_solver.c:98, :149, :184,
 * :298; SimInf_solver_ssm.c:155, :165; SimInf_solver_aem.c:74, :79, :139).
 * GSL is replaced here by:
 *
 *   - Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 *     as easy as 1, 2, 3", SC'11), a counter-based generator, so that each
 *     (replicate, node), each external-transfer event and each (node,
 *     transition) channel owns an independent stream that a GPU lane can
 *     start anywhere without a seeding pass;
 *   - siminf_log, a log with a fixed sequence of IEEE-754 double operations
 *     so that host and device produce identical bits (-ffp-contract=off on
 *     both sides; no FMA contraction);
 *   - binomial variates: inversion (Kachitvichyanukul & Schmeiser 1988,
 *     BINV) for n*min(p,1-p) < 30 and transformed rejection with
 *     decomposition (Hoermann 1993, BTRD) otherwise;
 *   - hypergeometric variates by simulating the urn draw by draw (the
 *     classical sequential method), drawing the smaller of the sample and
 *     its complement.
 *
 * Stream definition.  A stream is identified by (seed, domain, id).  Its
 * j-th uniform (j = 0, 1, 2, ...) is taken from Philox block b = j >> 1 with
 * counter (b_lo, b_hi, id_lo, domain << 28 | id_hi) and key (seed_lo,
 * seed_hi): words (0,1) give uniform 2b and words (2,3) give uniform 2b+1.
 * A 64-bit word x maps to ((x >> 12) + 0.5) * 2^-52, which lies strictly in
 * (0, 1) like gsl_rng_uniform_pos.
 */
#ifndef SIMINF_RAND_H
#define SIMINF_RAND_H

#include <stdint.h>

#if defined(__HIPCC__)
#  define SIMINF_FN __host__ __device__ static inline __attribute__((always_inline))
#  define SIMINF_COLD_FN __host__ __device__ static __attribute__((noinline))
#else
#  define SIMINF_FN static inline
#  define SIMINF_COLD_FN static __attribute__((noinline))
#endif

#if defined(__clang__)
#  define SIMINF_UNROLL _Pragma("unroll")
#else
#  define SIMINF_UNROLL
#endif

/* Longest run of uniforms a sampler takes from values computed ahead
 * (callers compute at most this many per stream). */
#define SIMINF_PRE_MAX 12

/* Stream domains. */
#define SIMINF_DOMAIN_NODE  0u  /* (replicate, node): SSA and E1 events */
#define SIMINF_DOMAIN_E2    1u  /* (replicate, event): one per E2 event */
#define SIMINF_DOMAIN_AEM   2u  /* (replicate, node, transition) */

typedef struct siminf_stream {
    uint32_t key0, key1;   /* seed */
    uint32_t id0, id1;     /* stream identity (domain in id1's top bits) */
    uint64_t j;            /* index of the next uniform */
    double cache;          /* uniform 2b+1 when j is odd */
    const double *pre;     /* uniforms 0 .. npre-1, computed beforehand */
    uint64_t npre;         /* even; 0 when nothing was computed ahead */
} siminf_stream;

SIMINF_FN uint32_t siminf_mulhi32(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

/* Philox4x32 with 10 rounds.  The multipliers and Weyl key increments
 * are the published Philox4x32 parameters. */
SIMINF_FN void siminf_philox4x32_10(
    uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
    uint32_t k0, uint32_t k1, uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    /* Unrolled: ten rounds of straight-line code that the compiler can
     * interleave with independent work around the call. */
    SIMINF_UNROLL
    for (int round = 0; round < 10; round++) {
        /* One 32x32->64 product per multiplier gives both halves (a
         * single v_mad_u64_u32 on the device). */
        const uint64_t p0 = (uint64_t)M0 * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)M1 * (uint64_t)c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

SIMINF_FN double siminf_u01(uint32_t lo, uint32_t hi)
{
    const uint64_t x = ((uint64_t)hi << 32) | (uint64_t)lo;
    return ((double)(x >> 12) + 0.5) * 0x1p-52;
}

SIMINF_FN siminf_stream siminf_stream_make(
    uint64_t seed, uint32_t domain, uint64_t id)
{
    siminf_stream s;
    s.key0 = (uint32_t)seed;
    s.key1 = (uint32_t)(seed >> 32);
    s.id0 = (uint32_t)id;
    s.id1 = (domain << 28) | ((uint32_t)(id >> 32) & 0x0FFFFFFFu);
    s.j = 0;
    s.cache = 0.0;
    s.pre = 0;
    s.npre = 0;
    return s;
}

/* Next uniform of the stream, in (0, 1).  The first npre uniforms may
 * have been computed ahead (by other lanes) into pre; they are the same
 * values this function would compute. */
SIMINF_FN double siminf_uniform(siminf_stream *s)
{
    uint32_t w[4];
    if (s->j < s->npre)
        return s->pre[s->j++];
    if (s->j & 1u) {
        s->j++;
        return s->cache;
    }
    siminf_philox4x32_10((uint32_t)(s->j >> 1), (uint32_t)(s->j >> 33),
                         s->id0, s->id1, s->key0, s->key1, w);
    s->cache = siminf_u01(w[2], w[3]);
    s->j++;
    return siminf_u01(w[0], w[1]);
}

/* The next aligned pair of uniforms, 2b and 2b+1 with 2b the first even
 * position at or after the stream's (a stream standing at an odd
 * position drops the uniform there).  The SSA takes one pair per
 * attempt -- the waiting time from the first, the transition from the
 * second, which is dropped when the waiting time ends the day -- so that
 * every attempt costs one Philox block and lanes of a wavefront never
 * disagree about computing one. */
SIMINF_FN void siminf_uniform2(siminf_stream *s, double *u0, double *u1)
{
    uint32_t w[4];
    const uint64_t b = (s->j + 1u) >> 1;
    if (2u * b + 1u < s->npre) {
        *u0 = s->pre[2u * b];
        *u1 = s->pre[2u * b + 1u];
    } else {
        siminf_philox4x32_10((uint32_t)b, (uint32_t)(b >> 32), s->id0, s->id1,
                             s->key0, s->key1, w);
        *u0 = siminf_u01(w[0], w[1]);
        *u1 = siminf_u01(w[2], w[3]);
    }
    s->j = 2u * b + 2u;
}

/* Uniforms 2b and 2b+1 of a stream (one Philox block). */
SIMINF_FN void siminf_uniform_pair(const siminf_stream *s, uint64_t b,
                                   double *u0, double *u1)
{
    uint32_t w[4];
    siminf_philox4x32_10((uint32_t)b, (uint32_t)(b >> 32), s->id0, s->id1,
                         s->key0, s->key1, w);
    *u0 = siminf_u01(w[0], w[1]);
    *u1 = siminf_u01(w[2], w[3]);
}

/* The j-th uniform of a stream without touching any stream state. */
SIMINF_FN double siminf_uniform_at(const siminf_stream *s, uint64_t j)
{
    uint32_t w[4];
    siminf_philox4x32_10((uint32_t)(j >> 1), (uint32_t)(j >> 33),
                         s->id0, s->id1, s->key0, s->key1, w);
    return (j & 1u) ? siminf_u01(w[2], w[3]) : siminf_u01(w[0], w[1]);
}

SIMINF_FN uint64_t siminf_bits(double x)
{
    union { double d; uint64_t u; } c;
    c.d = x;
    return c.u;
}

SIMINF_FN double siminf_from_bits(uint64_t u)
{
    union { double d; uint64_t u; } c;
    c.u = u;
    return c.d;
}

/* The arithmetic of siminf_log for the bits b of a positive normal
 * finite number, plus e to its exponent. */
SIMINF_FN double siminf_log_core(uint64_t b, int e)
{
    const double LN2_HI = 0x1.62e42fef00000p-1;
    const double LN2_LO = 0x1.473de6af278edp-34;
    e += (int)(b >> 52) - 1023;
    b &= 0x000FFFFFFFFFFFFFull;
    {
        /* m >= sqrt(2): use m/2.  Selects, not a branch, so that the
         * series is straight-line code. */
        const int half = b >= 0x6A09E667F3BCDull;
        b |= half ? 0x3FE0000000000000ull : 0x3FF0000000000000ull;
        e += half;
    }

    {
        const double m = siminf_from_bits(b);
        const double f = m - 1.0;
        const double s = f / (2.0 + f);
        const double z = s * s;
        const double two_s = s + s;
        double r = 1.0 / 25.0;
        r = 1.0 / 23.0 + z * r;
        r = 1.0 / 21.0 + z * r;
        r = 1.0 / 19.0 + z * r;
        r = 1.0 / 17.0 + z * r;
        r = 1.0 / 15.0 + z * r;
        r = 1.0 / 13.0 + z * r;
        r = 1.0 / 11.0 + z * r;
        r = 1.0 / 9.0 + z * r;
        r = 1.0 / 7.0 + z * r;
        r = 1.0 / 5.0 + z * r;
        r = 1.0 / 3.0 + z * r;
        r = z * r;
        {
            const double de = (double)e;
            return de * LN2_HI + (two_s + (two_s * r + de * LN2_LO));
        }
    }
}

/* siminf_log of a uniform in (0, 1) (always normal): the same arithmetic
 * without the special cases, so that the SSA's waiting time is
 * straight-line code. */
SIMINF_FN double siminf_log_u01(double x)
{
    return siminf_log_core(siminf_bits(x), 0);
}

/*
 * Natural logarithm.  x = 2^e * m with m in [sqrt(2)/2, sqrt(2)); with
 * f = m - 1 (exact) and s = f / (2 + f),
 *     log(m) = 2 atanh(s) = 2s (1 + s^2/3 + s^4/5 + ...),
 * |s| <= 0.1716, and the series is summed to s^24 (truncation below
 * 2^-60 relative).  ln2 is split so that e*LN2_HI is exact.
 */
SIMINF_FN double siminf_log(double x)
{
    uint64_t b = siminf_bits(x);
    int e;

    if (x != x)
        return x;                          /* NaN */
    if (x < 0.0)
        return (x - x) / (x - x);          /* NaN */
    if (x == 0.0)
        return -1.0 / (x * x);             /* -inf */
    if (b >= 0x7FF0000000000000ull)
        return x;                          /* +inf */

    e = 0;
    if (b < 0x0010000000000000ull) {       /* subnormal: scale by 2^54 */
        x = x * 0x1p54;
        b = siminf_bits(x);
        e = -54;
    }
    return siminf_log_core(b, e);
}

/* Square root by Newton's iteration from an exponent-halving start
 * (within a factor 1.5 of the root); six steps reach full precision.
 * Subnormal inputs are scaled by 2^600 first. */
SIMINF_FN double siminf_sqrt(double x)
{
    double y, scale = 1.0;
    uint64_t b;
    if (!(x > 0.0))
        return (x == 0.0) ? x : (x - x) / (x - x);
    if (siminf_bits(x) >= 0x7FF0000000000000ull)
        return x;
    if (x < 0x1p-1000) {
        x = x * 0x1p600;
        scale = 0x1p-300;
    }
    b = siminf_bits(x);
    y = siminf_from_bits((b >> 1) + 0x1FF8000000000000ull);
    for (int i = 0; i < 6; i++)
        y = 0.5 * (y + x / y);
    return y * scale;
}

/* x^n for integer n >= 0 by binary exponentiation. */
SIMINF_FN double siminf_powi(double x, int n)
{
    double r = 1.0;
    while (n > 0) {
        if (n & 1)
            r = r * x;
        x = x * x;
        n >>= 1;
    }
    return r;
}

/* log(k!) - [(k + 1/2) log(k + 1) - (k + 1) + log(2 pi) / 2], the
 * Stirling-series remainder used by BTRD. */
SIMINF_COLD_FN double siminf_stirling_fc(int k)
{
    if (k < 10) {
        const double HALF_LOG_2PI = 0.91893853320467274178;
        double lf = 0.0;
        for (int i = 2; i <= k; i++)
            lf = lf + siminf_log((double)i);
        return lf - (((double)k + 0.5) * siminf_log((double)k + 1.0)
                     - ((double)k + 1.0) + HALF_LOG_2PI);
    } else {
        const double kp1 = (double)k + 1.0;
        const double r2 = 1.0 / (kp1 * kp1);
        return (1.0 / 12.0 - (1.0 / 360.0 - (1.0 / 1260.0) * r2) * r2) / kp1;
    }
}

/* Binomial(n, p) with 0 <= p <= 0.5 and n*p < 30: sequential inversion. */
SIMINF_FN int siminf_binomial_inv(siminf_stream *s, double p, int n)
{
    const double q = 1.0 - p;
    const double ratio = p / q;
    const double a = (double)(n + 1) * ratio;
    const double q_n = siminf_powi(q, n);
    for (;;) {
        double u = siminf_uniform(s);
        double px = q_n;
        int x = 0;
        for (;;) {
            if (u <= px)
                return x;
            u = u - px;
            x++;
            if (x > n || x > 110)
                break;                      /* numerical tail: restart */
            px = px * (a / (double)x - ratio);
        }
    }
}

/* Binomial(n, p) with 0 < p <= 0.5 and n*p >= 30: BTRD. */
SIMINF_FN int siminf_binomial_btrd(siminf_stream *s, double p, int n)
{
    const double q = 1.0 - p;
    const double dn = (double)n;
    const int m = (int)((dn + 1.0) * p);
    const double r = p / q;
    const double nr = (dn + 1.0) * r;
    const double npq = dn * p * q;
    const double sqrt_npq = siminf_sqrt(npq);
    const double b = 1.15 + 2.53 * sqrt_npq;
    const double a = -0.0873 + 0.0248 * b + 0.01 * p;
    const double c = dn * p + 0.5;
    const double alpha = (2.83 + 5.1 / b) * sqrt_npq;
    const double v_r = 0.92 - 4.2 / b;
    const double u_rv_r = 0.86 * v_r;

    for (;;) {
        double u, v, us;
        int k, km;

        v = siminf_uniform(s);
        if (v <= u_rv_r) {
            u = v / v_r - 0.43;
            us = 0.5 - (u < 0.0 ? -u : u);
            return (int)((2.0 * a / us + b) * u + c);
        }

        if (v >= v_r) {
            u = siminf_uniform(s) - 0.5;
        } else {
            u = v / v_r - 0.93;
            u = (u < 0.0 ? -0.5 : 0.5) - u;
            v = siminf_uniform(s) * v_r;
        }

        us = 0.5 - (u < 0.0 ? -u : u);
        {
            const double kd = (2.0 * a / us + b) * u + c;
            if (kd < 0.0 || kd >= dn + 1.0)
                continue;
            k = (int)kd;
        }
        v = v * alpha / (a / (us * us) + b);
        km = k > m ? k - m : m - k;

        if (km <= 15) {
            double f = 1.0;
            if (m < k) {
                for (int i = m + 1; i <= k; i++)
                    f = f * (nr / (double)i - r);
            } else if (m > k) {
                for (int i = k + 1; i <= m; i++)
                    v = v * (nr / (double)i - r);
            }
            if (v <= f)
                return k;
            continue;
        }

        {
            const double dkm = (double)km;
            const double lv = siminf_log(v);
            const double rho = (dkm / npq) *
                (((dkm / 3.0 + 0.625) * dkm + 1.0 / 6.0) / npq + 0.5);
            const double t = -dkm * dkm / (2.0 * npq);
            double h, nm, nk;
            if (lv < t - rho)
                return k;
            if (lv > t + rho)
                continue;
            nm = dn - (double)m + 1.0;
            h = ((double)m + 0.5) * siminf_log(((double)m + 1.0) / (r * nm))
                + siminf_stirling_fc(m) + siminf_stirling_fc(n - m);
            nk = dn - (double)k + 1.0;
            if (lv <= h + (dn + 1.0) * siminf_log(nm / nk)
                + ((double)k + 0.5) * siminf_log(nk * r / ((double)k + 1.0))
                - siminf_stirling_fc(k) - siminf_stirling_fc(n - k))
                return k;
        }
    }
}

/* BTRD is rarely taken, so it is kept out of line to spare the callers'
 * registers.  The stream goes in and out by value: a pointer argument
 * would force the caller's stream, which the SSA loop uses on every
 * firing, out of registers into memory. */
typedef struct siminf_btrd_out {
    siminf_stream s;
    int x;
} siminf_btrd_out;

SIMINF_COLD_FN siminf_btrd_out siminf_binomial_btrd_cold(siminf_stream s,
                                                         double p, int n)
{
    siminf_btrd_out r;
    r.x = siminf_binomial_btrd(&s, p, n);
    r.s = s;
    return r;
}

/* Binomial(n, p) variate for 0 <= p <= 1, n >= 0. */
SIMINF_FN int siminf_binomial(siminf_stream *s, double p, int n)
{
    int flip = 0, x;
    if (n <= 0 || p <= 0.0)
        return 0;
    if (p >= 1.0)
        return n;
    if (p > 0.5) {
        p = 1.0 - p;
        flip = 1;
    }
    if ((double)n * p < 30.0) {
        x = siminf_binomial_inv(s, p, n);
    } else {
        /* Only the position and the cached uniform come back: the other
         * fields stay what the caller knows them to be (for a node's
         * stream, no uniforms computed ahead). */
        const siminf_btrd_out r = siminf_binomial_btrd_cold(*s, p, n);
        s->j = r.s.j;
        s->cache = r.s.cache;
        x = r.x;
    }
    return flip ? n - x : x;
}

/* Hypergeometric: the number of 'good' items in a sample of t drawn
 * without replacement from n1 good and n2 bad items.  The urn is drawn
 * item by item; when t exceeds half the urn the complement (the items
 * left behind) is drawn instead. */
SIMINF_FN int siminf_hypergeometric(siminf_stream *s, int n1, int n2, int t)
{
    const int total = n1 + n2;
    int draws, good = n1, left = total, k = 0;
    if (t > total)
        t = total;
    draws = (2 * t <= total) ? t : total - t;
    if (draws > 0 && draws <= SIMINF_PRE_MAX && s->j + (uint64_t)draws <= s->npre) {
        /* Every uniform the urn can need was computed ahead: the same
         * draws as below, without the stream's branches.  Until the urn
         * runs out of good items, draw i sees left = total - i, so every
         * product u_i * left is known up front, and with good an integer
         * u_i * left < good is floor(u_i * left) < good: the draws then
         * chain through integer additions only.  Once no good item is
         * left, no product is below 0, so k stops growing exactly where
         * the draw-by-draw urn stops drawing. */
        const double *pre = s->pre + s->j;
        const int full = s->j + (uint64_t)SIMINF_PRE_MAX <= s->npre;
        double v[SIMINF_PRE_MAX];
        int fl[SIMINF_PRE_MAX];
        int used = draws;
        /* Unconditional loads, issued together (entries past draws are
         * read from a valid slot and not used). */
        if (full) {
            SIMINF_UNROLL
            for (int i = 0; i < SIMINF_PRE_MAX; i++)
                v[i] = pre[i];
        } else {
            SIMINF_UNROLL
            for (int i = 0; i < SIMINF_PRE_MAX; i++)
                v[i] = pre[i < draws ? i : 0];
        }
        SIMINF_UNROLL
        for (int i = 0; i < SIMINF_PRE_MAX; i++)
            fl[i] = i < draws ? (int)(v[i] * (double)(left - i)) : 0;
        SIMINF_UNROLL
        for (int i = 0; i < SIMINF_PRE_MAX; i++) {
            if (i < draws) {
                const int take = fl[i] + k < good;
                k += take;
                if (take && k == good && used == draws)
                    used = i + 1;
            }
        }
        s->j += (uint64_t)used;
        return (2 * t <= total) ? k : n1 - k;
    }
    for (int i = 0; i < draws; i++) {
        const double u = siminf_uniform(s);
        if (u * (double)left < (double)good) {
            k++;
            good--;
            if (good == 0)
                break;
        }
        left--;
    }
    return (2 * t <= total) ? k : n1 - k;
}

#endif /* SIMINF_RAND_H */
