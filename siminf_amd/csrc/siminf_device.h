/*
 * siminf_device.h -- gfx950 device code of the solver core: per-node
 * direct-method SSA and all-events-method day steps, scheduled-event
 * application and the kernel parameter block.
 *
 * One lane owns one (replicate, node).  Its compartment state u[NC], its
 * continuous state v[ND], its propensities rate[NT] and the running sum a0
 * live in VGPRs; every index into them is a compile-time constant, so the
 * model's rate functions inline into straight-line code.  The selected
 * transition is dispatched through a compile-time chain of cases.  The
 * dependency graph G and the stoichiometry S arrive at run time as
 * per-transition bit masks and a dense value table, read with uniform
 * (scalar) loads.
 *
 * Reference: src/solvers/SimInf_solver_ssm.c:144-302 (SSA step, PTS and
 * refresh), src/solvers/SimInf_solver_aem.c:61-384 (aem), and
 * src/solvers/SimInf_solver.c:62-324, :658-919 (sampling, events).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <utility>

#include "siminf_rand.h"

#include <type_traits>
#include "siminf_models.h"
#include "../../include/siminf_hip.h"

namespace siminf {

/* ------------------------------------------------------------------ */
/* Compile-time iteration.                                             */
/* ------------------------------------------------------------------ */

template <int N, class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}

template <int N, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl<N>(f, std::make_integer_sequence<int, N>{});
}

/* ------------------------------------------------------------------ */
/* Kernel parameters.                                                  */
/* ------------------------------------------------------------------ */

/* Error keys order the errors the way a one-thread reference run meets
 * them: replicate, day, phase, position, with the code in the low bits.
 * The smallest key over all lanes is the reported error. */
enum Phase { PH_INIT = 0, PH_SSA = 1, PH_E1 = 2, PH_E2 = 3, PH_PTS = 4 };

__host__ __device__ inline uint64_t error_key(int rep, int day, int phase,
                                              long long pos, int code)
{
    const uint64_t r = rep < 0xFFFF ? (uint64_t)rep : 0xFFFFull;
    const uint64_t d = day < 0x3FFFF ? (uint64_t)day : 0x3FFFFull;
    const uint64_t p = pos < 0x1FFFFF ? (uint64_t)pos : 0x1FFFFFull;
    return (r << 48) | (d << 30) | ((uint64_t)phase << 27) | (p << 6) |
           (uint64_t)(-code & 63);
}

/* Barrier words per replicate of a grid-wide coupled launch. */
#define SIMINF_GSYNC_WORDS 4

struct KParams {
    /* Dimensions. */
    int Nn, Nrep, Nld, tlen, len, Nshift;
    long long lanes;                 /* Nn * Nrep */
    unsigned long long seed;
    /* When not null, the run's seed is read from here instead (a run
     * captured in a graph takes a new seed on every replay). */
    const unsigned long long *seed_dev;
    double tt0, next0;               /* tspan[0], floor(tspan[0]) + 1 */

    /* Model inputs. */
    const int *u0;                   /* Nc x (Nn*Nrep)  */
    const double *v0;                /* Nd x (Nn*Nrep)  */
    const double *ldata;             /* Nld x Nn, or {INFINITY} */
    const double *gdata;
    const double *tspan;
    /* Parameters per replicate (batches of proposals): ld_rep 1 when
     * ldata holds Nld x Nn values for each replicate in turn, 2 when it
     * holds Nld values per replicate for all its nodes (0: Nld x Nn shared
     * by the replicates); gd_stride the gdata values per replicate (0: one
     * gdata for all). */
    int ld_rep, gd_stride;

    /* S and G: Smask[k] has bit c set when S[c,k] is stored, Sval[k*NC+c]
     * holds its value; Gmask[k] has bit j set when G[j,k] is stored. */
    const uint32_t *Smask;
    const int *Sval;
    const uint32_t *Gmask;

    /* E and N: Emask[s] has bit c set when E[c,s] is stored, Ew[s*NC+c]
     * holds its weight; Nshift_mat is N (Nc x Nshift) or NULL. */
    const uint32_t *Emask;
    const double *Ew;
    const int *Nmat;

    /* Events.  ev_rank[k] is event k's position in the caller's list
     * (error reports).  The kernels read packed records, 32 bytes each
     * plus the proportion:
     *   E1 (exit, enter, internal transfer), grouped by node in list
     *   order: node i's are e1_rec[2j], e1_rec[2j+1], e1_prop[j] for j in
     *   e1_ptr[i] .. e1_ptr[i+1], with
     *     e1_rec[2j]   = (time, type, n, select)
     *     e1_rec[2j+1] = (shift, k, 0, 0);
     *   E2 (external transfer), in level order: day d's conflict-free
     *   levels are e2_day[d] .. e2_day[d+1], level l holds positions
     *   e2_lvl[l] .. e2_lvl[l+1], and position i is
     *     e2_rec[2i]   = (k, n, select, shift)
     *     e2_rec[2i+1] = (src slot, dst slot, src node, dest node)
     *   where the slots index the herds of the day's touched list
     *   e2_tn[e2_tn_ptr[d] .. e2_tn_ptr[d+1]) (dst slot -1 for an invalid
     *   destination).  select and shift are zero-based, shift -1 for
     *   none; an event the fast path takes has shift SIMINF_E2_FAST and
     *   its E column's rows packed in place of the select (first row |
     *   second row << 8 | (two rows) << 16). */
    const int *ev_rank;
    const int *e1_ptr;
    const int4 *e1_rec;
    const double *e1_prop;
    const int *e2_day, *e2_lvl;
    const int4 *e2_rec;
    const double *e2_prop;
    const int *e2_tn_ptr, *e2_tn;
    /* Per plan day: 1 when the day's E2 work fits the LDS buffers, every
     * event takes the fast path and there are under 128 levels; such a
     * day's second record words hold the byte offsets, in the staged herd
     * state, of the source's and the destination's counts in the event's
     * rows (src row a, src row b, dst row a, dst row b). */
    const int *e2_day_fast;
    /* Per node, the plan days on which external transfers touch it, in
     * order: e2_hday[e2_hptr[i] .. e2_hptr[i+1]).  The coupled kernel's
     * herds publish and reload their state only on those days. */
    const int *e2_hptr, *e2_hday;
    int ndays;
    int Nselect;
    /* A run over part of the plan's days (the intervals of a particle
     * filter): the plan's day index of the run's first day, E1 events at
     * or before time ev_floor skipped, E2 days at or before plan day
     * e2_floor skipped (INT_MIN / -1 / 0 for a whole run). */
    int day0, ev_floor, e2_floor;

    /* LDS layout of the coupled kernel (byte offsets into the dynamic
     * LDS block) for one day's E2 work, used when the day has at most
     * e2_cap_t touched herds, e2_cap_ev events and e2_cap_lv levels
     * (other days run from global memory). */
    int e2_cap_t, e2_cap_ev, e2_cap_lv;
    int lds_u, lds_upd, lds_tn, lds_rec, lds_prop, lds_lvl, lds_pre;
    /* E and N for the E2 work, staged once per kernel (zero caps when
     * they do not fit). */
    int lds_Ew, lds_Emask, lds_N;

    /* Output: dense U/V or sparse (irU, jcU, prU) / (irV, jcV, prV). */
    int *U;
    double *V;
    const int *irU, *jcU;
    double *prU;
    const int *irV, *jcV;
    double *prV;

    /* Coupled kernel: compartment state published for the E2 phase and
     * the update flags it sets.  A grid-wide launch runs the replicates
     * rep0, rep0 + 1, ... with rep_blocks workgroups each (0: one
     * replicate over the whole grid), the workgroups of each replicate
     * joined by barrier words of their own; rep_err (one per replicate,
     * or null) keeps each replicate's first error, so that a replicate's
     * exit decision never depends on another's errors. */
    int *u_work;                     /* Nc x (Nn*Nrep) */
    int *upd_work;                   /* Nn*Nrep */
    /* Models whose post-time-step function reads other nodes: every
     * node's v at the start of the day, published with its u (Nd x
     * (Nn*Nrep)); null otherwise. */
    double *v_work;
    int rep0;
    int rep_blocks;
    unsigned long long *rep_err;
    /* The coupled kernel runs the all-events method instead of the SSA
     * (one replicate, grid launch). */
    int coupled_aem;
    /* Grid launch: the node each lane runs (-1: none), or null for lane
     * = node; and, if not null, each node's transition count of the last
     * run (Nn). */
    const int *lane_node;
    unsigned long long *node_fired;
    /* Grid launch: the barrier words of each replicate of the launch
     * (SIMINF_GSYNC_WORDS each: arrivals, last day whose E2 work is done,
     * arrivals after the post-time-step phase of models that read other
     * nodes), cleared before each launch. */
    unsigned *gsync;

    /* Node kernels with a dense trajectory: a node whose state can no
     * longer change records the first column it has not written
     * (frozen_from, tlen when it never froze) and its state; a fill
     * kernel then writes those columns in trajectory order.
     * frozen_cols counts the node-columns left to it.  Null when
     * unused. */
    int *frozen_from;
    int *frozen_u;
    double *frozen_v;
    unsigned long long *frozen_cols;

    /* Node kernels run in one of three modes (NODES_*): every lane
     * through the whole run; or a classify pass over every lane, which
     * hands nodes frozen from the start to the fill kernel and lists the
     * others in active (count in *active_count), then an active pass
     * over the listed lanes, packed 1 << lpw_shift to a wavefront
     * (lpw_shift -1: the power of two that gives about two wavefronts per
     * SIMD of the device's simds, at most 64 lanes).  The active pass
     * reads the count on the device: its grid is fixed, and each
     * wavefront takes the wave slots w, w + (wavefronts in the grid),
     * ... below the slots the count needs. */
    int mode;
    int lpw_shift;
    int simds;
    int *active;
    int *active_count;
    /* The fill kernel's trajectory-flush schedule: the columns one of its
     * workgroups writes per 256 lanes (0: SIMINF_FILL_COLS). */
    int fill_block;

    /* Results. */
    unsigned long long *err_key;
    unsigned long long *transitions;

    /* Fused prevalence (node plans; null when off): instead of writing
     * U, every column's sums of the cases and population compartments
     * (fp_cases, fp_pop masks) go to fp_sum[2 (rep tlen + it) + 0/1], and a
     * node whose state stops changing at column it adds its counts once to
     * the difference array fp_delta[2 (rep (tlen + 1) + it) + 0/1], summed
     * over the columns when the prevalence is read. */
    long long *fp_sum;
    long long *fp_delta;
    unsigned fp_cases, fp_pop;
};

__device__ __forceinline__ void record_error(const KParams &p, int rep, int day,
                                             int phase, long long pos, int code)
{
    const unsigned long long key = error_key(rep, day, phase, pos, code);
    atomicMin(p.err_key, key);
    if (p.rep_err)
        atomicMin(p.rep_err + rep, key);
}

__device__ __forceinline__ bool rate_ok(double r)
{
    return siminf_isfinite(r) && !(r < 0.0);
}

/* ------------------------------------------------------------------ */
/* Register-array helpers with a run-time index.                       */
/* ------------------------------------------------------------------ */

template <int N, class T>
__device__ __forceinline__ T sel(const T (&a)[N], int i)
{
    T r = a[0];
    static_for<N>([&](auto K) {
        constexpr int k = decltype(K)::value;
        if (k > 0 && i == k)
            r = a[k];
    });
    return r;
}

/* Holds x in a register at this point of the program: the work that
 * computes x stays before the branches that follow (the compiler would
 * otherwise sink it past them, next to its use, where it no longer runs
 * beside the independent work of the same block). */
__device__ __forceinline__ void pin(double &x)
{
    __asm__ volatile("" : "+v"(x));
}

/* a[i] for a run-time i as a sum of masked elements: a chain of selects
 * between elements can be turned back into one indexed load, which moves
 * the array to scratch memory. */
template <int N>
__device__ __forceinline__ unsigned long long pick_u64(const unsigned long long (&a)[N],
                                                       int i)
{
    unsigned long long r = 0;
    static_for<N>([&](auto K) {
        constexpr int k = decltype(K)::value;
        r += a[k] & (0ull - (unsigned long long)(i == k));
    });
    return r;
}

/* a[i] += x for a run-time i, written as an unconditional masked update
 * of every element so that the array stays in registers (a guarded
 * store per element can be merged back into one indexed store). */
template <int N>
__device__ __forceinline__ void add_at(int (&a)[N], int i, int x)
{
    static_for<N>([&](auto K) {
        constexpr int k = decltype(K)::value;
        a[k] += x & -(int)(i == k);
    });
}

/* ------------------------------------------------------------------ */
/* Direct-method SSA for one node over [t, next) (ssm.c:144-235).      */
/* ------------------------------------------------------------------ */

template <class M>
struct NodeState {
    int u[M::NC];
    double v[M::ND > 0 ? M::ND : 1];
    double vnew[M::ND > 0 ? M::ND : 1];
    double rate[M::NT];
    double a0;
    siminf_stream rng;
    /* The SSA's next uniform pair (Philox block nb of rng), computed one
     * attempt ahead so that Philox runs beside the current attempt's
     * log and division instead of before them. */
    double nu0, nu1;
    uint64_t nb;
};

/* What a lane reads while it fires transitions: its node's ldata and the
 * gdata copied into registers once (the model reads M::NLD and M::NGD
 * values; ldata beyond Nld reads as INFINITY, like the reference's
 * placeholder when Nld == 0), and S and G.  With CANON the run's S and G
 * equal the model's own tables (M::S[k][c] = S(c,k), M::G[k][j] = G(j,k))
 * and fold into the code; otherwise their masks and values are loaded
 * once from the kernel arguments. */
template <class M, bool CANON>
struct Ctx {
    static constexpr bool canon = CANON;
    double ld[M::NLD > 0 ? M::NLD : 1];
    double gd[M::NGD > 0 ? M::NGD : 1];
    uint32_t smask[CANON ? 1 : M::NT];
    uint32_t gmask[CANON ? 1 : M::NT];
    int sval[CANON ? 1 : M::NT * M::NC];
    unsigned long long seed;

    __device__ __forceinline__ void load(const KParams &p, int node, int rep)
    {
        const long long li = p.ld_rep == 2 ? (long long)rep
                             : (p.ld_rep ? (long long)rep * p.Nn : 0LL) + node;
        const double *l = p.ldata + li * (long long)p.Nld;
        const double *gsrc = p.gdata + (long long)rep * p.gd_stride;
        static_for<(M::NLD > 0 ? M::NLD : 1)>([&](auto I) {
            constexpr int i = decltype(I)::value;
            ld[i] = i < p.Nld ? l[i] : INFINITY;
        });
        static_for<(M::NGD > 0 ? M::NGD : 1)>([&](auto I) {
            constexpr int i = decltype(I)::value;
            gd[i] = i < M::NGD ? gsrc[i] : 0.0;
        });
        if constexpr (!CANON) {
            static_for<M::NT>([&](auto K) {
                constexpr int k = decltype(K)::value;
                smask[k] = p.Smask[k];
                gmask[k] = p.Gmask[k];
                static_for<M::NC>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    sval[k * M::NC + c] = p.Sval[k * M::NC + c];
                });
            });
        }
        seed = p.seed;
    }

    /* Defined values for a lane without a node, which runs a wavefront's
     * straight-line code with results it discards. */
    __device__ __forceinline__ void clear(const KParams &p)
    {
        static_for<(M::NLD > 0 ? M::NLD : 1)>([&](auto I) { ld[decltype(I)::value] = 0.0; });
        static_for<(M::NGD > 0 ? M::NGD : 1)>([&](auto I) { gd[decltype(I)::value] = 0.0; });
        if constexpr (!CANON) {
            static_for<M::NT>([&](auto K) {
                constexpr int k = decltype(K)::value;
                smask[k] = 0;
                gmask[k] = 0;
                static_for<M::NC>([&](auto C) { sval[k * M::NC + decltype(C)::value] = 0; });
            });
        }
        seed = p.seed;
    }

    template <int K, int C>
    __device__ __forceinline__ bool s_stored() const
    {
        if constexpr (CANON)
            return M::S[K][C] != 0;
        else
            return (smask[K] >> C) & 1u;
    }

    template <int K, int C>
    __device__ __forceinline__ int s_val() const
    {
        if constexpr (CANON)
            return M::S[K][C];
        else
            return sval[K * M::NC + C];
    }

    template <int K, int J>
    __device__ __forceinline__ bool g_stored() const
    {
        if constexpr (CANON)
            return M::G[K][J] != 0;
        else
            return (gmask[K] >> J) & 1u;
    }
};

/* Fire transition K: apply S(:,K), then recompute the rates listed in
 * G(:,K) in row order and add their changes to a0.  Returns an error
 * code or 0; like the reference the step completes after an error. */
template <class M, int K, class CT>
__device__ __forceinline__ int fire(NodeState<M> &s, const CT &cx, double t)
{
    int err = 0;
    static_for<M::NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        if (cx.template s_stored<K, c>()) {
            s.u[c] += cx.template s_val<K, c>();
            if (s.u[c] < 0)
                err = SIMINF_ERR_NEGATIVE_STATE;
        }
    });
    double delta = 0.0;
    static_for<M::NT>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if (cx.template g_stored<K, j>()) {
            const double old = s.rate[j];
            const double r = M::template rate<j>(s.u, s.v, cx.ld, cx.gd, t);
            s.rate[j] = r;
            delta += r - old;
            if (!rate_ok(r))
                err = SIMINF_ERR_INVALID_RATE;
        }
    });
    s.a0 += delta;
    return err;
}

/* Bits k of G's row j: the transitions whose firing makes rate j
 * stale. */
template <class M, int J>
__device__ __forceinline__ constexpr unsigned g_row_mask()
{
    unsigned m = 0;
    for (int k = 0; k < M::NT; k++)
        if (M::G[k][J] != 0)
            m |= 1u << k;
    return m;
}

/* fire<M, K> for a run-time transition tr in one code path, for models
 * whose S and G are folded into the code: the state moves by S's column
 * tr through selects, and every rate G marks is recomputed in G's row
 * order (the others are evaluated and discarded, which changes nothing:
 * rates are pure functions).  Lanes of a wavefront that picked different
 * transitions then run the same instructions instead of one variant
 * after another. */
template <class M, class CT>
__device__ __forceinline__ int fire_any(NodeState<M> &s, const CT &cx, int tr,
                                        double t)
{
    int err = 0;
    static_for<M::NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        int d = 0;
        static_for<M::NT>([&](auto K) {
            constexpr int k = decltype(K)::value;
            if constexpr (M::S[k][c] != 0)
                d += tr == k ? M::S[k][c] : 0;
        });
        s.u[c] += d;
        if (s.u[c] < 0)
            err = SIMINF_ERR_NEGATIVE_STATE;
    });
    double delta = 0.0;
    static_for<M::NT>([&](auto J) {
        constexpr int j = decltype(J)::value;
        constexpr unsigned gm = g_row_mask<M, j>();
        if constexpr (gm != 0) {
            const bool in = (gm >> tr) & 1u;
            const double old = s.rate[j];
            const double r = M::template rate<j>(s.u, s.v, cx.ld, cx.gd, t);
            s.rate[j] = in ? r : old;
            delta += in ? r - old : 0.0;
            if (in && !rate_ok(r))
                err = SIMINF_ERR_INVALID_RATE;
        }
    });
    s.a0 += delta;
    return err;
}

/* fire_any computed on copies and kept only where on: every lane runs
 * the same straight-line code whether it fires or not, so that the
 * compiler can interleave it with independent work (the next attempt's
 * Philox block and logarithm).  Returns an error code (0 where !on). */
template <class M, class CT>
__device__ __forceinline__ int fire_sel(NodeState<M> &s, const CT &cx, int tr,
                                        double t, bool on)
{
    int err = 0;
    int un[M::NC];
    static_for<M::NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        int d = 0;
        static_for<M::NT>([&](auto K) {
            constexpr int k = decltype(K)::value;
            if constexpr (M::S[k][c] != 0)
                d += tr == k ? M::S[k][c] : 0;
        });
        un[c] = s.u[c] + d;
        if (un[c] < 0)
            err = SIMINF_ERR_NEGATIVE_STATE;
    });
    double delta = 0.0;
    static_for<M::NT>([&](auto J) {
        constexpr int j = decltype(J)::value;
        constexpr unsigned gm = g_row_mask<M, j>();
        if constexpr (gm != 0) {
            const bool in = (gm >> tr) & 1u;
            const double old = s.rate[j];
            const double r = M::template rate<j>(un, s.v, cx.ld, cx.gd, t);
            s.rate[j] = on && in ? r : old;
            delta += in ? r - old : 0.0;
            if (in && !rate_ok(r))
                err = SIMINF_ERR_INVALID_RATE;
        }
    });
    static_for<M::NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        s.u[c] = on ? un[c] : s.u[c];
    });
    s.a0 = on ? s.a0 + delta : s.a0;
    return on ? err : 0;
}

/* One SSA firing at time t: the transition u_tr picks by a linear scan
 * of the cumulative rates (ssm.c:160-189) is applied.  Returns false when
 * the node stops: a0 fell to zero, or the step failed (err).  few (the
 * same on every lane of the wavefront) says that at most two lanes are
 * firing: each then runs only its transition's code, where otherwise all
 * lanes run the one code path of fire_any. */
template <class M, class CT>
__device__ __forceinline__ bool ssa_fire(NodeState<M> &s, const CT &cx, double t,
                                         double u_tr, unsigned long long &fired,
                                         int &err, bool few = false)
{
    constexpr int NT = M::NT;
    /* The scan's stopping point without branches: rates are non-negative,
     * so the cumulative sums do not decrease and the transitions the
     * scan passes are those with rnd above their sum (the additions are
     * the scan's, in its order).  Past the end means the last one. */
    const double rnd = u_tr * s.a0;
    int tr = 0;
    double cum = s.rate[0];
    static_for<NT - 1>([&](auto K) {
        constexpr int k = decltype(K)::value;
        tr += rnd > cum ? 1 : 0;
        if constexpr (k + 1 < NT)
            cum += s.rate[k + 1];
    });
    double rt = s.rate[0];
    static_for<NT>([&](auto K) {
        constexpr int k = decltype(K)::value;
        if constexpr (k > 0)
            rt = tr == k ? s.rate[k] : rt;
    });
    /* Floating point fix (ssm.c:170-189): a zero rate walks back to the
     * nearest non-zero one below it. */
    if (rt == 0.0) {
        static_for<NT>([&](auto K) {
            constexpr int k = NT - 1 - decltype(K)::value;
            if (k > 0 && tr == k && s.rate[k] == 0.0) {
                tr = k - 1;
                rt = s.rate[k - 1];
            }
        });
        if (rt == 0.0) {
            s.a0 = 0.0;
            return false;
        }
    }

    fired++;
    auto each = [&]() {
        static_for<NT>([&](auto K) {
            constexpr int k = decltype(K)::value;
            if (tr == k) {
                const int e = fire<M, k>(s, cx, t);
                if (e)
                    err = e;
            }
        });
    };
    if constexpr (CT::canon && NT <= 8) {
        if (few)
            each();
        else
            err = fire_any<M>(s, cx, tr, t);
    } else {
        each();
    }
    /* The node stops at its first failing step. */
    return err == 0;
}

template <class M, class CT>
__device__ __forceinline__ int ssa_day(NodeState<M> &s, const CT &cx,
                                       double tt, double next,
                                       unsigned long long &fired)
{
    int err = 0;
    double t = tt;
    for (;;) {
        if (s.a0 <= 0.0)
            break;
        /* siminf_uniform2 for a node stream (never computed ahead), with
         * the pair of the next attempt started here. */
        double u_tau, u_tr;
        const uint64_t b = (s.rng.j + 1u) >> 1;
        if (b == s.nb) {
            u_tau = s.nu0;
            u_tr = s.nu1;
        } else {
            siminf_uniform_pair(&s.rng, b, &u_tau, &u_tr);
        }
        s.rng.j = 2u * b + 2u;
        siminf_uniform_pair(&s.rng, b + 1u, &s.nu0, &s.nu1);
        s.nb = b + 1u;
        const double tau = -siminf_log_u01(u_tau) / s.a0;
        if ((tau + t) >= next)
            break;
        t += tau;
        if (!ssa_fire<M>(s, cx, t, u_tr, fired, err))
            break;
    }
    return err;
}

/* Lane l's value v from lane src of the wavefront. */
__device__ __forceinline__ uint32_t lane_pull(int src, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}

__device__ __forceinline__ double lane_pull(int src, double v)
{
    const uint64_t b = siminf_bits(v);
    const uint32_t lo = lane_pull(src, (uint32_t)b);
    const uint32_t hi = lane_pull(src, (uint32_t)(b >> 32));
    return siminf_from_bits(((uint64_t)hi << 32) | lo);
}

/* Values of the lane l ^ 1 or l ^ 2 of the same quad (DPP quad_perm
 * moves; every lane of the quad must be active). */
template <int X>
__device__ __forceinline__ uint32_t quad_xor(uint32_t v)
{
    static_assert(X == 1 || X == 2, "quad lanes");
    constexpr int ctrl = X == 1 ? 0xB1 : 0x4E;   /* quad_perm(1,0,3,2), (2,3,0,1) */
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, ctrl, 0xF, 0xF, false);
}

template <int X>
__device__ __forceinline__ double quad_xor(double v)
{
    const uint64_t b = siminf_bits(v);
    const uint32_t lo = quad_xor<X>((uint32_t)b);
    const uint32_t hi = quad_xor<X>((uint32_t)(b >> 32));
    return siminf_from_bits(((uint64_t)hi << 32) | lo);
}

/* The earliest (t, k) of the quad: smaller t, the smaller k among exact
 * ties (the first channel wins, as aem_day's scan). */
__device__ __forceinline__ void quad_argmin(double &t, int &k)
{
    {
        const double t1 = quad_xor<1>(t);
        const int k1 = (int)quad_xor<1>((uint32_t)k);
        const bool take = t1 < t || (t1 == t && k1 < k);
        t = take ? t1 : t;
        k = take ? k1 : k;
    }
    {
        const double t1 = quad_xor<2>(t);
        const int k1 = (int)quad_xor<2>((uint32_t)k);
        const bool take = t1 < t || (t1 == t && k1 < k);
        t = take ? t1 : t;
        k = take ? k1 : k;
    }
}

/* OR of a flag over the quad. */
__device__ __forceinline__ bool quad_any(bool b)
{
    uint32_t x = b ? 1u : 0u;
    x |= quad_xor<1>(x);
    x |= quad_xor<2>(x);
    return x != 0u;
}

/* Position of the (x+1)-th set bit of m (x < popcount(m)). */
__device__ __forceinline__ int nth_set_bit(uint64_t m, int x)
{
    int p = 0;
    uint32_t w = (uint32_t)m;
    int c = __popc(w);
    if (x >= c) {
        x -= c;
        p = 32;
        w = (uint32_t)(m >> 32);
    }
    c = __popc(w & 0xFFFFu);
    if (x >= c) { x -= c; p += 16; w >>= 16; }
    c = __popc(w & 0xFFu);
    if (x >= c) { x -= c; p += 8; w >>= 8; }
    c = __popc(w & 0xFu);
    if (x >= c) { x -= c; p += 4; w >>= 4; }
    c = __popc(w & 0x3u);
    if (x >= c) { x -= c; p += 2; w >>= 2; }
    return p + (x >= (int)(w & 1u) ? 1 : 0);
}

/* ssa_day for a whole wavefront: every lane enters, live says whether it
 * holds a node.  The attempts' Philox blocks and waiting-time logarithms
 * are computed by all 64 lanes together in refills: with A lanes still in
 * their day, lane l computes attempt k = l / A of the running lane of
 * rank l % A, so each running lane holds up to 64 / A attempts after one
 * refill and fetches them from the lanes that computed them
 * (ds_bpermute).  A refill runs when some running lane has used its
 * share.  A wavefront's day lasts as long as its busiest lane's, and most
 * of that lane's attempts come when few lanes are left, so the Philox
 * block and the logarithm, which are most of an attempt's instructions,
 * leave the busiest lane's chain.  Stream positions and every value are
 * those of ssa_day. */
template <class M, class CT>
__device__ __forceinline__ int ssa_day_wave(NodeState<M> &s, const CT &cx,
                                            double tt, double next,
                                            unsigned long long &fired,
                                            bool live)
{
    int err = 0;
    double t = tt;
    bool run = live && s.a0 > 0.0;
    const uint64_t b0 = live ? (s.rng.j + 1u) >> 1 : 0u;
    uint64_t b = b0;               /* this lane's next block */
    double hE = 0.0, hU = 0.0;     /* the attempt this lane computed */
    int A = 1, rank = 0, pos = 0, have = 0;
    const int lane = (int)__lane_id();
    for (;;) {
        const uint64_t running = __ballot(run);
        if (running == 0)
            break;
        if (__ballot(run && pos >= have) != 0) {
            A = __popcll(running);
            /* l / A and l % A for l < 64 by a float reciprocal: (l + 0.5) / A
             * is at least 0.5 / A away from an integer, far beyond the
             * rounding of the product. */
            const float rA = __builtin_amdgcn_rcpf((float)A);
            const int k = (int)(((float)lane + 0.5f) * rA);
            const int c = nth_set_bit(running, lane - k * A);
            siminf_stream st;
            st.key0 = lane_pull(c, s.rng.key0);
            st.key1 = lane_pull(c, s.rng.key1);
            st.id0 = lane_pull(c, s.rng.id0);
            st.id1 = lane_pull(c, s.rng.id1);
            const uint64_t bc = ((uint64_t)lane_pull(c, (uint32_t)(b >> 32)) << 32) |
                                lane_pull(c, (uint32_t)b);
            double u0, u1;
            siminf_uniform_pair(&st, bc + (uint64_t)k, &u0, &u1);
            hE = -siminf_log_u01(u0);
            hU = u1;
            rank = (int)__builtin_amdgcn_mbcnt_hi(
                (uint32_t)(running >> 32),
                __builtin_amdgcn_mbcnt_lo((uint32_t)running, 0u));
            /* ceil((64 - rank) / A) = floor((63 - rank) / A) + 1 */
            have = (int)(((float)(63 - rank) + 0.5f) * rA) + 1;
            pos = 0;
        }
        const bool few = __popcll(running) <= 2;
        /* Every lane takes part in the pulls; only running lanes use
         * them. */
        const int src = run ? pos * A + rank : lane;
        const double E = lane_pull(src, hE);
        const double u_tr = lane_pull(src, hU);
        if (run) {
            b++;
            pos++;
            const double tau = E / s.a0;
            if ((tau + t) >= next) {
                run = false;
            } else {
                t += tau;
                /* As in ssa_day, the node stops drawing once its total rate
                 * is no longer positive. */
                run = ssa_fire<M>(s, cx, t, u_tr, fired, err, few) && s.a0 > 0.0;
            }
        }
    }
    if (live && b != b0)
        s.rng.j = 2u * b;
    return err;
}

/* ------------------------------------------------------------------ */
/* All-events method for one node over [tt, next) (aem.c:181-300).     */
/* ------------------------------------------------------------------ */

template <class M>
struct AemState {
    int u[M::NC];
    double v[M::ND > 0 ? M::ND : 1];
    double vnew[M::ND > 0 ? M::ND : 1];
    double rate[M::NT];
    double time[M::NT];     /* next firing time of each transition */
    double inf[M::NT];      /* reactInf */
    unsigned long long j[M::NT];
    /* Uniform cj[J] of channel J's stream, the other half of the Philox
     * block of its last draw (~0 when none): a channel's draws run through
     * consecutive positions, so every second one is already computed. */
    double cu[M::NT];
    unsigned long long cj[M::NT];
    siminf_stream rng;      /* node stream: E1 events */
};

/* Uniform at position j of channel J's stream c, from the channel's
 * cached half block when it holds j; the other half of a computed block
 * is cached. */
template <class M, int J>
__device__ __forceinline__ double aem_uniform(AemState<M> &s, const siminf_stream &c,
                                              unsigned long long j)
{
    if (s.cj[J] == j)
        return s.cu[J];
    double u0, u1;
    siminf_uniform_pair(&c, (uint64_t)(j >> 1), &u0, &u1);
    s.cj[J] = j ^ 1ull;
    s.cu[J] = (j & 1ull) ? u0 : u1;
    return (j & 1ull) ? u1 : u0;
}

/* calcTimes (aem.c:61-86) for channel J. */
template <class M, int J>
__device__ __forceinline__ void calc_times(AemState<M> &s,
                                           unsigned long long seed,
                                           uint64_t chan_base, double tt,
                                           double old_rate, double new_rate)
{
    const double oldtime = s.time[J];
    siminf_stream c = siminf_stream_make(seed, SIMINF_DOMAIN_AEM,
                                         chan_base + (uint64_t)J);
    c.j = s.j[J];
    if (isinf(oldtime)) {
        if (s.inf[J] == 0.0) {
            s.time[J] = -siminf_log(aem_uniform<M, J>(s, c, c.j)) / new_rate + tt;
            s.j[J] = c.j + 1;
        } else if (new_rate > 0.0) {
            s.time[J] = tt + (s.inf[J] / new_rate);
        }
    } else if (new_rate >= 2.2250738585072014e-308) {
        if (oldtime == tt) {
            s.time[J] = -siminf_log(aem_uniform<M, J>(s, c, c.j)) / new_rate + tt;
            s.j[J] = c.j + 1;
        } else {
            s.time[J] = ((old_rate / new_rate) * (oldtime - tt)) + tt;
        }
    } else {
        s.inf[J] = (oldtime - tt) * old_rate;
        s.time[J] = INFINITY;
    }
}

template <class M, int K, class CT>
__device__ __forceinline__ int aem_fire(AemState<M> &s, const CT &cx, double t,
                                        uint64_t chan_base)
{
    int err = 0;
    static_for<M::NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        if (cx.template s_stored<K, c>()) {
            s.u[c] += cx.template s_val<K, c>();
            if (s.u[c] < 0)
                err = SIMINF_ERR_NEGATIVE_STATE;
        }
    });
    static_for<M::NT>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if (j != K && cx.template g_stored<K, j>()) {
            const double old = s.rate[j];
            const double r = M::template rate<j>(s.u, s.v, cx.ld, cx.gd, t);
            s.rate[j] = r;
            if (!rate_ok(r))
                err = SIMINF_ERR_INVALID_RATE;
            calc_times<M, j>(s, cx.seed, chan_base, t, old, r);
        }
    });
    {
        const double old = s.rate[K];
        const double r = M::template rate<K>(s.u, s.v, cx.ld, cx.gd, t);
        s.rate[K] = r;
        if (!rate_ok(r))
            err = SIMINF_ERR_INVALID_RATE;
        calc_times<M, K>(s, cx.seed, chan_base, t, old, r);
    }
    return err;
}

template <class M, class CT>
__device__ __forceinline__ int aem_day(AemState<M> &s, const CT &cx,
                                       double next, uint64_t chan_base,
                                       unsigned long long &fired)
{
    constexpr int NT = M::NT;
    int err = 0;
    for (;;) {
        /* Earliest channel; the first index wins among exact ties. */
        int tr = 0;
        double t = s.time[0];
        static_for<NT>([&](auto K) {
            constexpr int k = decltype(K)::value;
            if (k > 0 && s.time[k] < t) {
                t = s.time[k];
                tr = k;
            }
        });
        if (isinf(t) || t >= next)
            break;
        fired++;
        static_for<NT>([&](auto K) {
            constexpr int k = decltype(K)::value;
            if (tr == k) {
                const int e = aem_fire<M, k>(s, cx, t, chan_base);
                if (e)
                    err = e;
            }
        });
        if (err)
            break;
    }
    return err;
}

/* calcTimes (aem.c:61-86) for channel J with the logarithm of the
 * channel's next uniform given (logu: siminf_log of the uniform at
 * position s.j[J] of the channel's stream, used only if the channel
 * draws). */
template <class M, int J>
__device__ __forceinline__ void calc_times_pre(AemState<M> &s, double logu, double tt,
                                               double old_rate, double new_rate)
{
    const double oldtime = s.time[J];
    if (isinf(oldtime)) {
        if (s.inf[J] == 0.0) {
            s.time[J] = -logu / new_rate + tt;
            s.j[J] = s.j[J] + 1;
        } else if (new_rate > 0.0) {
            s.time[J] = tt + (s.inf[J] / new_rate);
        }
    } else if (new_rate >= 2.2250738585072014e-308) {
        if (oldtime == tt) {
            s.time[J] = -logu / new_rate + tt;
            s.j[J] = s.j[J] + 1;
        } else {
            s.time[J] = ((old_rate / new_rate) * (oldtime - tt)) + tt;
        }
    } else {
        s.inf[J] = (oldtime - tt) * old_rate;
        s.time[J] = INFINITY;
    }
}

/* calc_times_pre for channel J without branches, applied only where on:
 * the three cases that divide (draw, wake-up, rescale) share one
 * division, num / new_rate with num = -logu, reactInf or the old rate,
 * and the results are selected; tt + q and q + tt are the same IEEE sum.
 * Returns whether the channel drew (used logu: its stream position then
 * advances by one, which the caller applies). */
__device__ __forceinline__ bool calc_times_one(double &time, double &rinf, double logu,
                                               double tt, double old_rate, double new_rate,
                                               bool on)
{
    const double oldtime = time;
    const bool oi = isinf(oldtime);
    const bool big = new_rate >= 2.2250738585072014e-308;
    const bool draw = oi ? rinf == 0.0 : (big && oldtime == tt);
    const bool wake = oi && rinf != 0.0 && new_rate > 0.0;
    const bool resc = !oi && big && oldtime != tt;
    const bool sleep = !oi && !big;
    const double num = draw ? -logu : (wake ? rinf : old_rate);
    const double q = num / new_rate;
    const double tn = resc ? ((q * (oldtime - tt)) + tt) : (q + tt);
    double tnew = oldtime;
    tnew = (draw || wake || resc) ? tn : tnew;
    tnew = sleep ? INFINITY : tnew;
    time = on ? tnew : oldtime;
    rinf = on && sleep ? (oldtime - tt) * old_rate : rinf;
    return on && draw;
}

template <class M, int J>
__device__ __forceinline__ bool calc_times_sel(AemState<M> &s, double logu, double tt,
                                               double old_rate, double new_rate, bool on)
{
    return calc_times_one(s.time[J], s.inf[J], logu, tt, old_rate, new_rate, on);
}

/* aem_fire with every channel's next logarithm given (lv). */
template <class M, int K, class CT>
__device__ __forceinline__ int aem_fire_pre(AemState<M> &s, const CT &cx, double t,
                                            const double (&lv)[M::NT])
{
    int err = 0;
    static_for<M::NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        if (cx.template s_stored<K, c>()) {
            s.u[c] += cx.template s_val<K, c>();
            if (s.u[c] < 0)
                err = SIMINF_ERR_NEGATIVE_STATE;
        }
    });
    static_for<M::NT>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if (j != K && cx.template g_stored<K, j>()) {
            const double old = s.rate[j];
            const double r = M::template rate<j>(s.u, s.v, cx.ld, cx.gd, t);
            s.rate[j] = r;
            if (!rate_ok(r))
                err = SIMINF_ERR_INVALID_RATE;
            calc_times_pre<M, j>(s, lv[j], t, old, r);
        }
    });
    {
        const double old = s.rate[K];
        const double r = M::template rate<K>(s.u, s.v, cx.ld, cx.gd, t);
        s.rate[K] = r;
        if (!rate_ok(r))
            err = SIMINF_ERR_INVALID_RATE;
        calc_times_pre<M, K>(s, lv[K], t, old, r);
    }
    return err;
}

/* ------------------------------------------------------------------ */
/* Sampling individuals for an event (SimInf_solver.c:62-324).         */
/* ------------------------------------------------------------------ */

/* First / last stored row of an E column, as compartment indices. */
__device__ __forceinline__ int mask_first(uint32_t m) { return __builtin_ctz(m); }
__device__ __forceinline__ int mask_last(uint32_t m) { return 31 - __builtin_clz(m); }
__device__ __forceinline__ int mask_prev(uint32_t m, int c)
{
    const uint32_t below = m & ((1u << c) - 1u);
    return below ? 31 - __builtin_clz(below) : -1;
}

/* Inversion over the stored rows of an E column with weights wv: the
 * first row whose cumulative weight reaches rnd, or the last row when
 * rounding leaves rnd above the total (SimInf_solver.c:185-193). */
template <int NC>
__device__ __forceinline__ int scan_rows(uint32_t mask, const double (&wv)[NC],
                                         double rnd)
{
    int i = -1;
    double cum = 0.0;
    bool searching = true, first = true;
    static_for<NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        if (mask & (1u << c)) {
            if (first) {
                cum = wv[c];
                first = false;
                i = c;
                searching = rnd > cum;
            } else if (searching) {
                i = c;
                cum += wv[c];
                searching = rnd > cum;
            }
        }
    });
    return i;
}

/* The largest stored row at or below i whose weight is non-zero, or -1:
 * the reference's walk back from i (SimInf_solver.c:194-205, :308-316). */
template <int NC>
__device__ __forceinline__ int last_nonzero_at_or_below(
    uint32_t mask, const double (&wv)[NC], int i)
{
    int j = -1;
    static_for<NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        if ((mask & (1u << c)) && c <= i && wv[c] != 0.0)
            j = c;
    });
    return j;
}

/* The run-time-indexed element a[i], as a masked sum (see add_at). */
template <int NC>
__device__ __forceinline__ int value_at(const int (&a)[NC], int i)
{
    int x = 0;
    static_for<NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        x |= a[c] & -(int)(i == c);
    });
    return x;
}

/* Is the run-time-indexed element a[i] negative?  Written as a masked
 * sum, like add_at: a chain of compares would be folded back into a
 * run-time index, which moves the array to scratch memory. */
template <int NC>
__device__ __forceinline__ bool negative_at(const int (&a)[NC], int i)
{
    int x = 0;
    static_for<NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        x |= a[c] & -(int)(i == c);
    });
    return x < 0;
}

/* sample_select for a node whose state is the register array u.  The E
 * column is (mask, w).  Fills ind and returns 0 or an error code. */
template <int NC>
__device__ int sample_select(const int (&u)[NC], uint32_t mask,
                             const double (&w)[NC], int n, double proportion,
                             int (&ind)[NC], siminf_stream *rng)
{
    int Nind = 0, Nkinds = 0;
    const int Nstates = __builtin_popcount(mask);

    static_for<NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        ind[c] = 0;
        if (mask & (1u << c)) {
            if (u[c] > 0)
                Nkinds++;
            Nind += u[c];
        }
    });

    if (n == 0) {
        if (proportion < 0 || proportion > 1)
            return SIMINF_ERR_INVALID_PROPORTION;
        n = siminf_binomial(rng, proportion, Nind);
    }

    if (Nstates <= 0 || n > Nind || n < 0)
        return SIMINF_ERR_SAMPLE_SELECT;

    if (n == 0)
        return 0;

    if (Nind == n) {
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            ind[c] = u[c] & -(int)((mask >> c) & 1u);
        });
        return 0;
    }

    if (Nstates == 1) {
        add_at(ind, mask_first(mask), n);
        return 0;
    }

    if (Nkinds == 1) {
        int first_kind = -1;
        static_for<NC>([&](auto C) {
            constexpr int c = NC - 1 - decltype(C)::value;
            if ((mask & (1u << c)) && u[c] > 0)
                first_kind = c;
        });
        add_at(ind, first_kind, n);
        return 0;
    }

    /* Are all weights in the column equal? */
    bool equal = true;
    {
        bool have_prev = false;
        double prev = 0.0;
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            if (mask & (1u << c)) {
                if (have_prev && w[c] != prev)
                    equal = false;
                prev = w[c];
                have_prev = true;
            }
        });
    }

    if (equal) {
        /* Conditional hypergeometric draws over the rows in order; the
         * last row takes the remainder (once n reaches zero the rest get
         * nothing).  A run-time loop over the rows, so that every lane
         * runs the same single copy of the urn whichever rows its column
         * holds. */
        const int last = mask_last(mask);
        uint32_t rows = mask & ~(1u << last);
        while (rows && n > 0) {
            const int c = mask_first(rows);
            rows &= rows - 1u;
            const int uc = value_at(u, c);
            const int x = siminf_hypergeometric(rng, uc, Nind - uc, n);
            add_at(ind, c, x);
            Nind -= uc;
            n -= x;
        }
        add_at(ind, last, n);
        return 0;
    }

    /* Wallenius urn, one individual per draw. */
    while (n > 0) {
        double wv[NC];
        double cum = 0.0;
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            wv[c] = w[c] * (u[c] - ind[c]);
            if (mask & (1u << c))
                cum += wv[c];
        });
        const double rnd = siminf_uniform(rng) * cum;
        const int i = scan_rows<NC>(mask, wv, rnd);
        /* Floating point fix (SimInf_solver.c:191-206): the nearest row at
         * or below i with weight left. */
        const int j = last_nonzero_at_or_below<NC>(mask, wv, i);
        if (j < 0)
            return SIMINF_ERR_SAMPLE_SELECT;
        add_at(ind, j, 1);
        n--;
    }
    return 0;
}

/* sample_select_enter (SimInf_solver.c:243-324). */
template <int NC>
__device__ int sample_select_enter(const int (&u)[NC], uint32_t mask,
                                   const double (&w)[NC], int n,
                                   double proportion, int (&ind)[NC],
                                   siminf_stream *rng)
{
    const int Nstates = __builtin_popcount(mask);
    static_for<NC>([&](auto C) { ind[decltype(C)::value] = 0; });

    if (n == 0) {
        if (proportion < 0 || proportion > 1)
            return SIMINF_ERR_INVALID_PROPORTION;
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            if (mask & (1u << c))
                n += u[c];
        });
        n = siminf_binomial(rng, proportion, n);
    }

    if (Nstates <= 0 || n < 0)
        return SIMINF_ERR_SAMPLE_SELECT;
    if (n == 0)
        return 0;
    if (Nstates == 1) {
        add_at(ind, mask_first(mask), n);
        return 0;
    }

    double w_cum = 0.0;
    static_for<NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        if (mask & (1u << c))
            w_cum += w[c];
    });

    while (n > 0) {
        const double rnd = siminf_uniform(rng) * w_cum;
        const int i = scan_rows<NC>(mask, w, rnd);
        const int j = last_nonzero_at_or_below<NC>(mask, w, i);
        if (j < 0)
            return SIMINF_ERR_SAMPLE_SELECT;
        add_at(ind, j, 1);
        n--;
    }
    return 0;
}

/* E and N as the event code reads them: Emask[s] has bit c set when
 * E[c,s] is stored and Ew[s*NC+c] holds its weight; N[shift*NC+c] is
 * N(c, shift), or N is null when the model has no shift matrix.  Small
 * and read by every event, they stay in the caches. */
struct EvTab {
    const uint32_t *Emask;
    const double *Ew;
    const int *N;
};

__device__ __forceinline__ EvTab global_tables(const KParams &p)
{
    return EvTab{p.Emask, p.Ew, p.Nmat};
}

/* Load E column s as (mask, dense weights). */
template <int NC>
__device__ __forceinline__ uint32_t load_E(const EvTab &tab, int s, double (&w)[NC])
{
    static_for<NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        w[c] = tab.Ew[s * NC + c];
    });
    return tab.Emask[s];
}

/* Apply an E1 event (exit, enter, internal transfer) to the register
 * state u of its node (SimInf_solver.c:680-822).  External transfers are
 * handled by apply_e2. */
template <int NC>
__device__ int apply_e1(const EvTab &tab, int type, int n, int select, int shift,
                        double prop, int (&u)[NC], siminf_stream *rng)
{
    double w[NC];
    int ind[NC];
    int err;
    const uint32_t mask = load_E<NC>(tab, select, w);

    if (type == SIMINF_EXIT_EVENT) {
        err = sample_select<NC>(u, mask, w, n, prop, ind, rng);
        if (err)
            return err;
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            if (!err && (mask & (1u << c))) {
                u[c] -= ind[c];
                if (u[c] < 0)
                    err = SIMINF_ERR_NEGATIVE_STATE;
            }
        });
        return err;
    }

    if (type == SIMINF_ENTER_EVENT) {
        err = sample_select_enter<NC>(u, mask, w, n, prop, ind, rng);
        if (err)
            return err;
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            if (!err && (mask & (1u << c))) {
                if (shift < 0) {
                    u[c] += ind[c];
                    if (u[c] < 0)
                        err = SIMINF_ERR_NEGATIVE_STATE;
                } else if (!tab.N) {
                    err = SIMINF_ERR_EVENTS_N;
                } else {
                    const int ll = tab.N[shift * NC + c];
                    if (c + ll < 0 || c + ll >= NC) {
                        err = SIMINF_ERR_SHIFT_OUT_OF_BOUNDS;
                    } else {
                        add_at(u, c + ll, ind[c]);
                        if (negative_at(u, c + ll))
                            err = SIMINF_ERR_NEGATIVE_STATE;
                    }
                }
            }
        });
        return err;
    }

    if (type == SIMINF_INTERNAL_TRANSFER_EVENT) {
        if (!tab.N)
            return SIMINF_ERR_EVENTS_N;
        if (shift < 0)
            return SIMINF_ERR_EVENT_SHIFT;
        err = sample_select<NC>(u, mask, w, n, prop, ind, rng);
        if (err)
            return err;
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            if (!err && (mask & (1u << c))) {
                const int ll = tab.N[shift * NC + c];
                if (c + ll < 0 || c + ll >= NC) {
                    err = SIMINF_ERR_SHIFT_OUT_OF_BOUNDS;
                } else {
                    add_at(u, c + ll, ind[c]);
                    if (negative_at(u, c + ll)) {
                        err = SIMINF_ERR_NEGATIVE_STATE;
                    } else {
                        u[c] -= ind[c];
                        if (u[c] < 0)
                            err = SIMINF_ERR_NEGATIVE_STATE;
                    }
                }
            }
        });
        return err;
    }

    return SIMINF_UNDEFINED_EVENT;
}

/* Apply an external transfer (SimInf_solver.c:824-902) to a state array
 * uw (NC ints per slot, in LDS or global memory): the source herd sits in
 * slot src, the destination in slot dst; upd receives the update flags of
 * both slots.  dest is the destination node, checked against the node
 * count before any slot is touched. */
/* Shift field of an external transfer that the host found to need no
 * shift and to select one row, or two rows of equal weight, of E, with a
 * count n > 0, a valid destination other than the source. */
#define SIMINF_E2_FAST (-2)

/* apply_e2 for those events: sample_select's branches for one or two
 * rows (take all, a single row, a single non-empty row, one conditional
 * hypergeometric draw for the first row with the second taking the rest)
 * over counts read straight from the state array, without the per-
 * compartment loops of the general case; the same draws, the same
 * results and errors. */
template <int NC>
__device__ __forceinline__ int apply_e2_fast(int n, int rows, int *uw, int *upd,
                                             int src, int dst, siminf_stream *rng)
{
    const int ra = rows & 0xFF, rb = (rows >> 8) & 0xFF;
    const bool two = (rows >> 16) != 0;
    int *sa = uw + src * NC + ra, *sb = uw + src * NC + rb;
    int *da = uw + dst * NC + ra, *db = uw + dst * NC + rb;
    const int ua = *sa, ub = two ? *sb : 0;
    const int nind = ua + ub;
    if (n > nind)
        return SIMINF_ERR_SAMPLE_SELECT;
    int xa;
    if (n == nind)
        xa = ua;
    else if (!two || ub <= 0)
        xa = n;
    else if (ua <= 0)
        xa = 0;
    else
        xa = siminf_hypergeometric(rng, ua, ub, n);
    const int xb = n - xa;
    const int va = *da + xa, wa = ua - xa;
    *da = va;
    *sa = wa;
    bool neg = (va < 0) | (wa < 0);
    if (two) {
        const int vb = *db + xb, wb = ub - xb;
        *db = vb;
        *sb = wb;
        neg = neg | (vb < 0) | (wb < 0);
    }
    if (neg)
        return SIMINF_ERR_NEGATIVE_STATE;
    upd[dst] = 1;
    upd[src] = 1;
    return 0;
}

/* apply_e2_fast for an event of a day prepared by the host (offsets of the
 * four counts in the staged state lu, in bytes) whose stream's first
 * SIMINF_PRE_MAX uniforms are in pre.  The counts and the uniforms are
 * loaded together and the urn is computed before any decision (its result
 * is used only where siminf_hypergeometric would be called), as in that
 * function's computed-ahead branch: with g_i = ua - floor(u_i * (total -
 * i)), draw i takes a good item exactly when k < g_i, so after the
 * independent g_i the chain is one add and one med3 per draw.  The
 * choice among the cases is made with selects, and a failing event writes
 * back the counts it read, which leaves the state as skipping it would.
 * The event's stream is used by nothing else, so its position is not
 * kept; an urn of more draws than were computed ahead takes the stream.
 * Returns the error code or 0. */
__device__ __forceinline__ int apply_e2_fast_pre(int n, int rows, unsigned char *lu,
                                                 int4 off, const double *pre,
                                                 unsigned long long seed, uint64_t stream_id)
{
    const bool two = (rows >> 16) != 0;
    int *sa = (int *)(lu + off.x), *sb = (int *)(lu + off.y);
    int *da = (int *)(lu + off.z), *db = (int *)(lu + off.w);
    const int ua = *sa, ub0 = *sb, dva = *da, dvb = *db;
    double v[SIMINF_PRE_MAX];
    SIMINF_UNROLL
    for (int i = 0; i < SIMINF_PRE_MAX; i++)
        v[i] = pre[i];
    /* Keep the loads here, ahead of the decisions below (the compiler
     * would otherwise sink the uniforms' loads into the urn's branch, a
     * second round trip after the counts'). */
    SIMINF_UNROLL
    for (int i = 0; i < SIMINF_PRE_MAX; i++)
        asm volatile("" ::"v"(v[i]));
    const int ub = two ? ub0 : 0;
    const int total = ua + ub;
    const bool low = 2 * n <= total;
    const int draws = low ? n : total - n;
    const double dt = (double)total;
    /* k + (k < g) as med3(k, k + 1, g), four draws at a time; draws past
     * every lane's count change nothing and are skipped.  The chain runs
     * in single precision (v_med3_f32): k, k + 1 and g are integers below
     * 2^24, exact there, so the result is the integer chain's. */
    float kf = 0.0f;
    auto four = [&](auto I0) {
        constexpr int i0 = decltype(I0)::value;
        float g[4];
        static_for<4>([&](auto J) {
            constexpr int i = i0 + decltype(J)::value;
            g[i - i0] = i < draws ? (float)(ua - (int)(v[i] * (dt - (double)i))) : -1.0f;
        });
        static_for<4>([&](auto J) {
            kf = __builtin_amdgcn_fmed3f(kf, kf + 1.0f, g[decltype(J)::value]);
        });
    };
    static_assert(SIMINF_PRE_MAX == 12, "the urn runs in three groups of four draws");
    /* Only an event that takes from both rows draws from the urn; the
     * groups run when some lane of the wavefront needs them. */
    const bool urn = n < total && two && ub > 0 && ua > 0;
    if (__ballot(urn) != 0) {
        four(std::integral_constant<int, 0>{});
        if (__ballot(urn && draws > 4) != 0) {
            four(std::integral_constant<int, 4>{});
            if (__ballot(urn && draws > 8) != 0)
                four(std::integral_constant<int, 8>{});
        }
    }
    const int k = (int)kf;
    int xa = n == total ? ua : (!two || ub <= 0) ? n : ua <= 0 ? 0 : low ? k : ua - k;
    if (urn && draws > SIMINF_PRE_MAX) {
        siminf_stream rng = siminf_stream_make(seed, SIMINF_DOMAIN_E2, stream_id);
        rng.pre = pre;
        rng.npre = SIMINF_PRE_MAX;
        xa = siminf_hypergeometric(&rng, ua, ub, n);
    }
    const bool ok = n <= total;
    const int xb = n - xa;
    const int va = ok ? dva + xa : dva, wa = ok ? ua - xa : ua;
    const int vb = ok ? dvb + xb : dvb, wb = ok ? ub - xb : ub0;
    *da = va;
    *sa = wa;
    if (two) {
        *db = vb;
        *sb = wb;
    }
    const bool neg = (va < 0) | (wa < 0) | (two & ((vb < 0) | (wb < 0)));
    return !ok ? SIMINF_ERR_SAMPLE_SELECT : neg ? SIMINF_ERR_NEGATIVE_STATE : 0;
}

template <int NC>
__device__ int apply_e2(const EvTab &tab, int n, int select, int shift,
                        double prop, int dest, int Nn, int *uw, int *upd,
                        int src, int dst, siminf_stream *rng)
{
    double w[NC];
    int us[NC], ind[NC];
    int err;

    if (dest < 0 || dest >= Nn)
        return SIMINF_ERR_DEST_OUT_OF_BOUNDS;

    const uint32_t mask = load_E<NC>(tab, select, w);
    static_for<NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        us[c] = uw[src * NC + c];
    });
    err = sample_select<NC>(us, mask, w, n, prop, ind, rng);
    if (err)
        return err;

    if (shift < 0 && src != dst) {
        /* No shift and two distinct herds: the destination's counts are
         * loaded together and both herds written back in one pass. */
        int ud[NC];
        bool neg = false;
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            if (mask & (1u << c))
                ud[c] = uw[dst * NC + c];
        });
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            if (mask & (1u << c)) {
                ud[c] += ind[c];
                us[c] -= ind[c];
                neg = neg | (ud[c] < 0) | (us[c] < 0);
                uw[dst * NC + c] = ud[c];
                uw[src * NC + c] = us[c];
            }
        });
        if (neg)
            return SIMINF_ERR_NEGATIVE_STATE;
        upd[dst] = 1;
        upd[src] = 1;
        return 0;
    }

    static_for<NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        if (!err && (mask & (1u << c))) {
            int kd = dst * NC + c;
            if (shift >= 0) {
                if (!tab.N) {
                    err = SIMINF_ERR_EVENTS_N;
                    return;
                }
                const int ll = tab.N[shift * NC + c];
                if (c + ll < 0 || c + ll >= NC) {
                    err = SIMINF_ERR_SHIFT_OUT_OF_BOUNDS;
                    return;
                }
                kd += ll;
            }
            uw[kd] += ind[c];
            if (uw[kd] < 0) {
                err = SIMINF_ERR_NEGATIVE_STATE;
                return;
            }
            uw[src * NC + c] -= ind[c];
            if (uw[src * NC + c] < 0)
                err = SIMINF_ERR_NEGATIVE_STATE;
        }
    });
    if (err)
        return err;
    upd[dst] = 1;
    upd[src] = 1;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Writeback (ssm.c:318-334, mssm.c:323-371, SimInf_solver.c:930-950). */
/* ------------------------------------------------------------------ */

/* Sparse column writeback for one node: the rows of column col that
 * belong to node are contiguous; find them by binary search. */
template <int NROW, class T>
__device__ __forceinline__ void store_sparse(const int *ir, const int *jc,
                                             double *pr, long long col,
                                             int node, const T (&x)[NROW])
{
    int lo = jc[col], hi = jc[col + 1];
    const int first = node * NROW;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ir[mid] < first)
            lo = mid + 1;
        else
            hi = mid;
    }
    const int end = jc[col + 1];
    for (int j = lo; j < end && ir[j] < first + NROW; j++) {
        const int r = ir[j] - first;
        /* The row's value by selects (an if per row lets the compiler
         * index x at run time, which moves x to scratch memory). */
        T val = x[0];
        static_for<NROW>([&](auto C) {
            constexpr int c = decltype(C)::value;
            if constexpr (c > 0)
                val = r == c ? x[c] : val;
        });
        pr[j] = (double)val;
    }
}

template <int NC, int ND>
__device__ __forceinline__ void store_columns(
    const KParams &p, int rep, int node, double tt, int &it, double &ts,
    const int (&u)[NC], const double (&vnew)[ND > 0 ? ND : 1])
{
    /* U and V columns advance together (both while tt > tspan[it]); ts
     * holds tspan[it], loaded when it advances, so that the load's
     * latency overlaps the next day instead of opening this one. */
    while (it < p.tlen && tt > ts) {
        const long long col = (long long)rep * p.tlen + it;
        if (p.U) {
            int *dst = p.U + (col * p.Nn + node) * NC;
            static_for<NC>([&](auto C) {
                constexpr int c = decltype(C)::value;
                dst[c] = u[c];
            });
        } else {
            store_sparse<NC>(p.irU, p.jcU, p.prU, col, node, u);
        }
        if constexpr (ND > 0) {
            if (p.V) {
                double *dst = p.V + (col * p.Nn + node) * ND;
                static_for<ND>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    dst[c] = vnew[c];
                });
            } else {
                store_sparse<ND>(p.irV, p.jcV, p.prV, col, node, vnew);
            }
        }
        it++;
        ts = it < p.tlen ? p.tspan[it] : INFINITY;
    }
}

/* A node's (cases, population) counts for the fused prevalence. */
template <int NC>
__device__ __forceinline__ void fused_counts(const KParams &p, const int (&u)[NC],
                                             long long &x, long long &y)
{
    x = 0;
    y = 0;
    static_for<NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        x += ((p.fp_cases >> c) & 1u) ? (long long)u[c] : 0LL;
        y += ((p.fp_pop >> c) & 1u) ? (long long)u[c] : 0LL;
    });
}

__device__ __forceinline__ void fused_add(long long *dst, long long x, long long y)
{
    atomicAdd((unsigned long long *)dst, (unsigned long long)x);
    atomicAdd((unsigned long long *)dst + 1, (unsigned long long)y);
}

/* Column it of replicate rep in the fused sums. */
template <int NC>
__device__ __forceinline__ void fused_column(const KParams &p, int rep, int it,
                                             const int (&u)[NC])
{
    long long x, y;
    fused_counts<NC>(p, u, x, y);
    fused_add(p.fp_sum + 2 * ((long long)rep * p.tlen + it), x, y);
}

/* Columns it .. tlen - 1 of replicate rep, all equal to u. */
template <int NC>
__device__ __forceinline__ void fused_from(const KParams &p, int rep, int it,
                                           const int (&u)[NC])
{
    long long x, y;
    fused_counts<NC>(p, u, x, y);
    fused_add(p.fp_delta + 2 * ((long long)rep * (p.tlen + 1) + it), x, y);
}

/* Largest tspan the free-running node kernels copy into LDS (64 KiB). */
#define SIMINF_TS_LDS_MAX 8192

/* tspan[i] from the workgroup's copy in LDS (TSL) or from global memory.
 * The copy matters beyond latency: vector-memory loads and stores share
 * one completion counter, so waiting for a global tspan load also waits
 * for every trajectory store the wavefront has in flight, while an LDS
 * read completes on its own counter. */
template <bool TSL>
__device__ __forceinline__ double tspan_at(const KParams &p, const double *ts_sh, int i)
{
    if constexpr (TSL)
        return ts_sh[i];
    else
        return p.tspan[i];
}

/* Copy tspan into the workgroup's LDS (every thread of the workgroup
 * calls this once, before any of them diverges). */
template <bool TSL>
__device__ __forceinline__ void tspan_to_lds(const KParams &p, double *ts_sh)
{
    if constexpr (TSL) {
        for (int i = (int)threadIdx.x; i < p.tlen; i += (int)blockDim.x)
            ts_sh[i] = p.tspan[i];
        __syncthreads();
    }
}

/* store_columns with the time point after ts also held (ts1): a day that
 * writes one column compares tt with values loaded a day or more before,
 * so the load of the next point never sits on the day's end (a lane
 * stepping through its own days reaches its next day's end within a few
 * hundred cycles, far less than a load's latency). */
template <int NC, int ND, bool TSL>
__device__ __forceinline__ void store_columns_ahead(
    const KParams &p, int rep, int node, double tt, int &it, double &ts, double &ts1,
    const int (&u)[NC], const double (&vnew)[ND > 0 ? ND : 1], const double *ts_sh,
    bool writer = true)
{
    while (it < p.tlen && tt > ts) {
        const long long col = (long long)rep * p.tlen + it;
        if (!writer) {
            /* another lane of the node writes */
        } else if (p.fp_sum) {
            fused_column<NC>(p, rep, it, u);
        } else if (p.U) {
            int *dst = p.U + (col * p.Nn + node) * NC;
            static_for<NC>([&](auto C) {
                constexpr int c = decltype(C)::value;
                dst[c] = u[c];
            });
        } else {
            store_sparse<NC>(p.irU, p.jcU, p.prU, col, node, u);
        }
        if constexpr (ND > 0) {
            if (!writer || p.fp_sum) {
            } else if (p.V) {
                double *dst = p.V + (col * p.Nn + node) * ND;
                static_for<ND>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    dst[c] = vnew[c];
                });
            } else {
                store_sparse<ND>(p.irV, p.jcV, p.prV, col, node, vnew);
            }
        }
        it++;
        ts = ts1;
        ts1 = it + 1 < p.tlen ? tspan_at<TSL>(p, ts_sh, it + 1) : INFINITY;
    }
}

/* The end of a day for a free-running lane of a model without events,
 * post-time-step work or continuous state, dense U: where the day writes
 * at most one column, the node is not frozen (alive) and the column is
 * not the last, the column is stored and tt, next and the time points
 * advance; returns whether it did (the caller's general day end takes the
 * others).  Straight-line code: the store is masked, the rest selects. */
template <int NC, bool TSL>
__device__ __forceinline__ bool day_end_fast(const KParams &p, int rep, int node, bool dend,
                                             bool alive, const int (&u)[NC], double &tt,
                                             double &next, int &it, double &ts, double &ts1,
                                             const double *ts_sh, bool writer = true)
{
    const double tn = next;
    const bool col = tn > ts;
    const bool fast = dend && alive && !(tn > ts1) && it + 1 < p.tlen;
    if (fast && col && writer) {
        if (p.fp_sum) {
            fused_column<NC>(p, rep, it, u);
        } else {
            int *dst = p.U + (((long long)rep * p.tlen + it) * p.Nn + node) * NC;
            static_for<NC>([&](auto C) {
                constexpr int c = decltype(C)::value;
                dst[c] = u[c];
            });
        }
    }
    const bool adv = fast && col;
    it += adv ? 1 : 0;
    ts = adv ? ts1 : ts;
    const int i1 = it + 1 < p.tlen ? it + 1 : p.tlen - 1;
    const double t1 = tspan_at<TSL>(p, ts_sh, i1);
    ts1 = adv ? (it + 1 < p.tlen ? t1 : INFINITY) : ts1;
    tt = fast ? tn : tt;
    next = fast ? tn + 1.0 : next;
    return fast;
}

/* Does the model declare that its post-time-step function never changes
 * anything (PTS_NOOP)?  Models that do not say are assumed to. */
template <class M, class = void>
struct pts_noop : std::false_type {};
template <class M>
struct pts_noop<M, std::void_t<decltype(M::PTS_NOOP)>>
    : std::bool_constant<M::PTS_NOOP> {};

/* Does the model's post-time-step function read other nodes' u and v
 * (NEIGHBORS, the spatial models)?  Such a model runs in the coupled
 * kernel, which passes the function pointers into the replicate's
 * node-major state arrays in global memory. */
template <class M, class = void>
struct model_neighbors : std::false_type {};
template <class M>
struct model_neighbors<M, std::void_t<decltype(M::NEIGHBORS)>>
    : std::bool_constant<M::NEIGHBORS> {};

/* A node whose state can no longer change (all rates zero, no events
 * left, a post-time-step function that does nothing) holds the same
 * state in every remaining column: record it for the fill kernel
 * instead of stepping the remaining days. */
template <int NC, int ND>
__device__ __forceinline__ void record_frozen(const KParams &p, long long g,
                                              int U_it, const int (&u)[NC],
                                              const double (&vnew)[ND > 0 ? ND : 1])
{
    p.frozen_from[g] = U_it;
    static_for<NC>([&](auto C) {
        constexpr int c = decltype(C)::value;
        p.frozen_u[g * NC + c] = u[c];
    });
    if constexpr (ND > 0) {
        static_for<ND>([&](auto C) {
            constexpr int c = decltype(C)::value;
            p.frozen_v[g * ND + c] = vnew[c];
        });
    }
}

enum { NODES_ALL = 0, NODES_CLASSIFY = 1, NODES_ACTIVE = 2 };

/* The seed of the run. */
__device__ __forceinline__ unsigned long long run_seed(const KParams &p)
{
    return p.seed_dev ? *p.seed_dev : p.seed;
}

/* The wave slots a wavefront of a node kernel runs: in the active pass
 * slot w holds the listed lanes w << shift .. (w + 1) << shift (at most
 * n of them), and the wavefront takes w, w + stride, ... below waves;
 * in the other modes one slot, the thread's own lane. */
struct WaveSlots {
    long long w, waves, stride;
    int shift, n;
};

__device__ __forceinline__ WaveSlots wave_slots(const KParams &p)
{
    WaveSlots a{0, 1, 1, 0, 0};
    if (p.mode != NODES_ACTIVE)
        return a;
    a.n = __hip_atomic_load(p.active_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int shift = p.lpw_shift;
    if (shift < 0) {
        shift = 0;
        while (shift < 6 && ((long long)a.n >> shift) > 2LL * p.simds)
            shift++;
    }
    a.shift = shift;
    a.waves = ((long long)a.n + (1 << shift) - 1) >> shift;
    a.w = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    a.stride = ((long long)gridDim.x * blockDim.x) >> 6;
    return a;
}

/* The lane (replicate * Nn + node) a thread of a node kernel runs in
 * slot a.w, or -1. */
__device__ __forceinline__ long long node_lane(const KParams &p, const WaveSlots &a)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p.mode != NODES_ACTIVE)
        return t < p.lanes ? t : -1;
    const int l = (int)(t & 63);
    if (l >= (1 << a.shift))
        return -1;
    const long long i = (a.w << a.shift) + l;
    return i < a.n ? (long long)p.active[i] : -1;
}

/* node_lane for kernels that give each node a quad of lanes (the
 * channel-parallel aem): lane l runs node l / 4 of its wave slot. */
__device__ __forceinline__ long long node_lane_quad(const KParams &p, const WaveSlots &a)
{
    const int q = (int)((threadIdx.x & 63) >> 2);
    if (q >= (1 << a.shift))
        return -1;
    const long long i = (a.w << a.shift) + q;
    return i < a.n ? (long long)p.active[i] : -1;
}

/* Append lane g to the active list (one atomic per wavefront). */
__device__ __forceinline__ void push_active(const KParams &p, long long g, bool push)
{
    const unsigned long long m = __ballot(push);
    const int lane = (int)(threadIdx.x & 63);
    int base = 0;
    const int first = m ? __ffsll((long long)m) - 1 : 0;
    if (m && lane == first)
        base = atomicAdd(p.active_count, __popcll(m));
    base = __shfl(base, first);
    if (push)
        p.active[base + __popcll(m & ((1ull << lane) - 1ull))] = (int)g;
}

/* A frozen node of the active pass writes its remaining columns itself
 * (the fill kernel has already been given the nodes frozen from the
 * start). */
template <int NC, int ND>
__device__ __forceinline__ void store_frozen(const KParams &p, long long g, int U_it,
                                             const int (&u)[NC],
                                             const double (&vnew)[ND > 0 ? ND : 1])
{
    if (p.fp_sum) {
        if (U_it < p.tlen)
            fused_from<NC>(p, (int)(g / p.Nn), U_it, u);
        return;
    }
    const int rep = (int)(g / p.Nn);
    const int node = (int)(g - (long long)rep * p.Nn);
    for (int it = U_it; it < p.tlen; it++) {
        const long long q = ((long long)rep * p.tlen + it) * p.Nn + node;
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            p.U[q * NC + c] = u[c];
        });
        if constexpr (ND > 0) {
            static_for<ND>([&](auto C) {
                constexpr int c = decltype(C)::value;
                p.V[q * ND + c] = vnew[c];
            });
        }
    }
}

/* Sum of a per-lane count into *dst, one atomic per wavefront. */
__device__ __forceinline__ void add_count(unsigned long long *dst,
                                          unsigned long long n)
{
    for (int off = 32; off > 0; off >>= 1)
        n += __shfl_xor(n, off);
    if ((threadIdx.x & 63) == 0 && n)
        atomicAdd(dst, n);
}

/* Full refresh of all NT rates after the post-time-step function
 * (ssm.c:272-301); returns an error code or 0. */
template <class M, class CT>
__device__ __forceinline__ int refresh_rates(NodeState<M> &s, const CT &cx,
                                             double tt)
{
    int err = 0;
    double delta = 0.0;
    static_for<M::NT>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const double old = s.rate[j];
        const double r = M::template rate<j>(s.u, s.vnew, cx.ld, cx.gd, tt);
        s.rate[j] = r;
        delta += r - old;
        if (!rate_ok(r))
            err = SIMINF_ERR_INVALID_RATE;
    });
    s.a0 += delta;
    return err;
}

__device__ __forceinline__ void add_transitions(const KParams &p, unsigned long long n)
{
    /* Wavefront reduction, one atomic per wave. */
    for (int off = 32; off > 0; off >>= 1)
        n += __shfl_xor(n, off);
    if ((threadIdx.x & 63) == 0 && n)
        atomicAdd(p.transitions, n);
}

} /* namespace siminf */
