/*
 * siminf_kernels.h -- the solver kernels as templates over a model
 * plugin, and the plugin record the library dispatches through.  Built-in
 * models instantiate them in siminf_hip.hip; a model generated by the
 * parser (siminf_amd/mparse.py) instantiates them in its own shared object
 * and hands the library a SimInf_hip_model.
 *
 * A model plugin M provides NC, NT, ND, NLD, NGD (ldata and gdata values
 * its functions read), its canonical tables S[NT][NC] (S[k][c] = S(c,k))
 * and G[NT][NT] (G[k][j] = G(j,k)), rate<J>(u, v, ldata, gdata, t) and
 * pts(v_new, u, v, ldata, gdata, node, t) as __device__ functions.
 */
#pragma once

#include "siminf_device.h"

#include <hip/hip_runtime.h>
#include <algorithm>

namespace siminf {

/* Model functions are always inlined: an out-of-line call would take
 * the addresses of the node's state arrays and move them from registers
 * to scratch memory. */
#define SIMINF_MODEL_FN __device__ __forceinline__ static

#define SIMINF_RATE_ARGS const int *u, const double *v, const double *ld, \
    const double *gd, double t
#define SIMINF_PTS_ARGS double *vn, const int *u, const double *v, \
    const double *ld, const double *gd, int node, double t

/* ------------------------------------------------------------------ */
/* Kernels.                                                            */
/* ------------------------------------------------------------------ */

/* A node's next E1 event, held in registers.  The record is loaded when
 * the previous one has been applied, so the load overlaps the following
 * SSA work instead of sitting on the day's critical path. */
struct E1Cursor {
    int cur, end;
    int4 a, b;         /* (time, type, n, select), (shift, k, 0, 0) */
    double prop;

    __device__ __forceinline__ void start(const KParams &p, int node)
    {
        cur = p.e1_ptr[node];
        end = p.e1_ptr[node + 1];
        while (cur < end && p.e1_rec[2 * cur].x <= p.ev_floor)
            cur++;
        fetch(p);
    }
    __device__ __forceinline__ void fetch(const KParams &p)
    {
        if (cur < end) {
            a = p.e1_rec[2 * cur];
            b = p.e1_rec[2 * cur + 1];
            prop = p.e1_prop[cur];
        }
    }
};

/* E1 events of one node due at tt, in list order, drawing from the
 * node's stream.  Returns 0 or an error code; *bad gets the failing
 * event. */
template <int NC>
__device__ __forceinline__ int e1_day(const KParams &p, const EvTab &tab,
                                      E1Cursor &ec, double tt, int (&u)[NC],
                                      siminf_stream *rng, bool &upd, int &bad)
{
    while (ec.cur < ec.end) {
        if ((double)ec.a.x > tt)
            break;
        const int err = apply_e1<NC>(tab, ec.a.y, ec.a.z, ec.a.w, ec.b.x,
                                     ec.prop, u, rng);
        if (err) {
            bad = ec.b.y;
            return err;
        }
        upd = true;
        ec.cur++;
        ec.fetch(p);
    }
    return 0;
}

/* One lane's ssm / mssm run in a node kernel after the classify pass
 * (modes NODES_ALL and NODES_ACTIVE; g: replicate * Nn + node, or -1 for
 * a lane without a node), with free-running lanes: each lane steps
 * through its own days, one SSA
 * attempt or one day's end per trip round the loop, so a wavefront lasts
 * as long as its busiest node's run, where a loop over days lasts the
 * sum over days of each day's busiest lane.  Every attempt consumes one
 * Philox block (ssa_day's schedule), so the next attempt's block and
 * waiting-time logarithm are computed during the current one, in the
 * same straight-line code as the attempt (fire_sel for models whose S and
 * G fold into the code).  Stream positions and every value are those of
 * ssa_day. */
template <class M, bool EVENTS, bool CANON, bool HELP, bool TSL>
__device__ __forceinline__ void ssm_lane_free(const KParams &p, long long g, int shift,
                                              const double *ts_sh)
{
    constexpr int NC = M::NC, NT = M::NT, ND = M::ND;
    const bool live = g >= 0;
    unsigned long long fired = 0;

    const int rep = live ? (int)(g / p.Nn) : 0;
    const int node = live ? (int)(g - (long long)rep * p.Nn) : 0;
    Ctx<M, CANON> cx;
    NodeState<M> s;
    s.a0 = 0.0;
    int err = 0;
    double tt = p.tt0, next = p.next0;
    const EvTab tab = global_tables(p);
    E1Cursor ec;
    auto frozen = [&]() __attribute__((always_inline)) -> bool {
        if constexpr (pts_noop<M>::value)
            return s.a0 <= 0.0 && (!EVENTS || ec.cur >= ec.end);
        else
            return false;
    };
    bool going = false;
    if (live) {
        cx.load(p, node, rep);
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            s.u[c] = p.u0[g * NC + c];
        });
        if constexpr (ND > 0) {
            static_for<ND>([&](auto C) {
                constexpr int c = decltype(C)::value;
                s.v[c] = p.v0[g * ND + c];
                s.vnew[c] = s.v[c];
            });
        }
        s.rng = siminf_stream_make(p.seed, SIMINF_DOMAIN_NODE, (uint64_t)g);
        static_for<NT>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const double r = M::template rate<j>(s.u, s.v, cx.ld, cx.gd, tt);
            s.rate[j] = r;
            s.a0 += r;
            if (!rate_ok(r))
                err = SIMINF_ERR_INVALID_RATE;
        });
        if constexpr (EVENTS)
            ec.start(p, node);
        if (err)
            record_error(p, rep, 0, PH_INIT, node, err);
        else
            going = true;
    } else {
        cx.clear(p);
        static_for<NC>([&](auto C) { s.u[decltype(C)::value] = 0; });
        static_for<(ND > 0 ? ND : 1)>([&](auto C) {
            s.v[decltype(C)::value] = 0.0;
            s.vnew[decltype(C)::value] = 0.0;
        });
        static_for<NT>([&](auto J) { s.rate[decltype(J)::value] = 0.0; });
        s.rng = siminf_stream_make(p.seed, SIMINF_DOMAIN_NODE, 0);
    }

    int U_it = 0, day = 0;
    double ts = tspan_at<TSL>(p, ts_sh, 0);
    double ts1 = p.tlen > 1 ? tspan_at<TSL>(p, ts_sh, 1) : INFINITY;
    double t = tt;
    /* b: the next attempt's block (the stream's position is 2b once an
     * attempt of the day has consumed one); bday: b at the day's start;
     * wE, wU: block b's waiting-time numerator -log(u0) and u1. */
    uint64_t b = 0, bday = 0;
    double wE = 1.0, wU = 0.5;
    auto load_block = [&]() __attribute__((always_inline)) {
        double u0, u1;
        siminf_uniform_pair(&s.rng, b, &u0, &u1);
        wE = -siminf_log_u01(u0);
        wU = u1;
    };
    if (going)
        load_block();
    if (going && p.mode == NODES_ACTIVE && frozen()) {
        store_frozen<NC, ND>(p, g, U_it, s.u, s.vnew);
        going = false;
    }
    /* With HELP the wavefront's A = 2^shift lanes hold nodes and all 64
     * lanes compute their attempts' values ahead into a table: lane l
     * holds block tb + l / A of the node on lane l % A (Q = 64 / A blocks
     * per node, refilled by the whole wavefront when some node's next
     * block is past its share), and each node fetches its next block's
     * values from the lane that computed them (ds_bpermute), one trip
     * ahead of their use. */
    const int lane = (int)__lane_id();
    const int A = 1 << shift, Q = 64 >> shift;
    const int own = lane & (A - 1), slot = lane >> shift;
    uint64_t tb = 0;
    double hE = 0.0, hU = 0.0;
    bool filled = false;

    for (;;) {
        if (__ballot(going) == 0)
            break;
        /* The attempt with block b; block b + 1's values beside it. */
        const bool att = going && s.a0 > 0.0;
        double nE, n1;
        if constexpr (HELP) {
            if (__ballot(going && (!filled || b + 1u - tb >= (uint64_t)Q)) != 0) {
                tb = b + 1u;
                siminf_stream st = s.rng;
                st.id0 = lane_pull(own, s.rng.id0);
                st.id1 = lane_pull(own, s.rng.id1);
                const uint64_t t0 = ((uint64_t)lane_pull(own, (uint32_t)(tb >> 32)) << 32) |
                                    lane_pull(own, (uint32_t)tb);
                double h0, h1;
                siminf_uniform_pair(&st, t0 + (uint64_t)slot, &h0, &h1);
                hE = -siminf_log_u01(h0);
                hU = h1;
                filled = true;
            }
            const int src = going ? (int)(b + 1u - tb) * A + own : lane;
            nE = lane_pull(src, hE);
            n1 = lane_pull(src, hU);
        } else {
            double n0;
            siminf_uniform_pair(&s.rng, b + 1u, &n0, &n1);
            nE = -siminf_log_u01(n0);
        }
        const double tau = wE / s.a0;
        bool fires = att && (tau + t) < next;
        t = fires ? t + tau : t;
        /* Transition by the scan's stopping point (ssa_fire). */
        const double rnd = wU * s.a0;
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
        /* Floating point fix (ssm.c:170-189) by selects: a zero rate
         * walks back to the nearest non-zero one below it; none left
         * stops the node for the day with a0 = 0. */
        static_for<NT>([&](auto K) {
            constexpr int k = NT - 1 - decltype(K)::value;
            if constexpr (k > 0) {
                const bool back = tr == k && s.rate[k] == 0.0;
                tr = back ? k - 1 : tr;
                rt = back ? s.rate[k - 1] : rt;
            }
        });
        const bool none = fires && rt == 0.0;
        s.a0 = none ? 0.0 : s.a0;
        fires = fires && !none;
        int e = 0;
        if constexpr (CANON && NT <= 8) {
            e = fire_sel<M>(s, cx, tr, t, fires);
        } else {
            if (fires) {
                static_for<NT>([&](auto K) {
                    constexpr int k = decltype(K)::value;
                    if (tr == k)
                        e = fire<M, k>(s, cx, t);
                });
            }
        }
        fired += fires ? 1u : 0u;
        if (att) {
            b++;
            wE = nE;
            wU = n1;
        }
        pin(wE);
        pin(wU);
        if (e) {
            record_error(p, rep, day, PH_SSA, node, e);
            going = false;
        }
        /* The day ends here unless the lane fired and keeps a positive
         * total rate (ssa_day_wave). */
        const bool dend = going && !(fires && s.a0 > 0.0);
        bool slow = dend;
        if constexpr (!EVENTS && pts_noop<M>::value && ND == 0) {
            /* Most days end with at most one dense column written and
             * nothing else to do: that runs as selects and one masked
             * store, beside the next trip's work; the rest (a node that
             * freezes or finishes, several columns in one day, sparse
             * output) takes the general code below. */
            if (p.U || p.fp_sum)
                slow = dend && !day_end_fast<NC, TSL>(p, rep, node, dend, s.a0 > 0.0, s.u,
                                                       tt, next, U_it, ts, ts1, ts_sh);
            const bool ended = dend && !slow;
            t = ended ? tt : t;
            bday = ended ? b : bday;
            day += ended ? 1 : 0;
        }
        if (__ballot(slow) == 0 || !slow)
            continue;
        bool upd = false;
        if constexpr (EVENTS) {
            int bad = 0;
            if (b != bday)
                s.rng.j = 2u * b;
            const uint64_t j0 = s.rng.j;
            err = e1_day<NC>(p, tab, ec, tt, s.u, &s.rng, upd, bad);
            if (err) {
                record_error(p, rep, day, PH_E1, p.ev_rank[bad], err);
                going = false;
                continue;
            }
            if (s.rng.j != j0) {
                b = (s.rng.j + 1u) >> 1;
                load_block();
            }
        }
        const int rc = M::pts(s.vnew, s.u, s.v, cx.ld, cx.gd, node, tt);
        if (rc < 0) {
            record_error(p, rep, day, PH_PTS, node, rc);
            going = false;
            continue;
        }
        if (rc > 0 || upd) {
            err = refresh_rates<M>(s, cx, tt);
            if (err) {
                record_error(p, rep, day, PH_PTS, node, err);
                going = false;
                continue;
            }
        }
        tt = next;
        next += 1.0;
        t = tt;
        bday = b;
        store_columns_ahead<NC, ND, TSL>(p, rep, node, tt, U_it, ts, ts1, s.u, s.vnew, ts_sh);
        if constexpr (ND > 0) {
            static_for<ND>([&](auto C) {
                constexpr int c = decltype(C)::value;
                const double x = s.v[c];
                s.v[c] = s.vnew[c];
                s.vnew[c] = x;
            });
        }
        if (U_it >= p.tlen)
            going = false;
        day++;
        if (going && p.mode == NODES_ACTIVE && frozen()) {
            store_frozen<NC, ND>(p, g, U_it, s.u, s.vnew);
            going = false;
        }
    }
    add_transitions(p, fired);
}

/* The classify pass of the node kernels: one lane per (replicate, node)
 * computes the initial rates (an invalid one is the run's PH_INIT error),
 * hands a node that can never change (every rate zero, or a0 <= 0 for
 * ssm; no events; a post-time-step function that does nothing) to the
 * fill kernel and lists the others for the active pass.  A kernel of its
 * own, small enough to fill the device. */
template <class M, bool EVENTS, bool CANON, bool AEM>
__global__ __launch_bounds__(1024) void k_classify(KParams p)
{
    constexpr int NC = M::NC, NT = M::NT, ND = M::ND;
    /* The active list and the frozen-column count take one atomic per
     * workgroup (the listed lanes are written at offsets from a prefix
     * over the workgroup's wavefronts): one per wavefront on two counters
     * serialised at the cache and cost as much as the pass itself. */
    __shared__ unsigned s_cnt[16];
    __shared__ unsigned long long s_frz[16];
    __shared__ int s_base;
    const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int nw = (int)(blockDim.x >> 6);
    unsigned long long frozen_cols = 0;
    for (long long g0 = (long long)blockIdx.x * blockDim.x; g0 < p.lanes;
         g0 += (long long)gridDim.x * blockDim.x) {
        const long long g = g0 + threadIdx.x;
        const bool live = g < p.lanes;
        bool push = false;
        long long fx = 0, fy = 0;   /* a frozen node's fused counts */
        if (live) {
            const int rep = (int)(g / p.Nn);
            const int node = (int)(g - (long long)rep * p.Nn);
            Ctx<M, CANON> cx;
            cx.load(p, node, rep);
            int u[NC];
            double v[ND > 0 ? ND : 1];
            static_for<NC>([&](auto C) {
                constexpr int c = decltype(C)::value;
                u[c] = p.u0[g * NC + c];
            });
            if constexpr (ND > 0) {
                static_for<ND>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    v[c] = p.v0[g * ND + c];
                });
            } else {
                v[0] = 0.0;
            }
            int err = 0;
            double a0 = 0.0;
            bool zero = true;
            static_for<NT>([&](auto J) {
                constexpr int j = decltype(J)::value;
                const double r = M::template rate<j>(u, v, cx.ld, cx.gd, p.tt0);
                a0 += r;
                zero = zero && r == 0.0;
                if (!rate_ok(r))
                    err = SIMINF_ERR_INVALID_RATE;
            });
            bool idle = AEM ? zero : a0 <= 0.0;
            if constexpr (EVENTS) {
                E1Cursor ec;
                ec.start(p, node);
                idle = idle && ec.cur >= ec.end;
            }
            bool frz = false;
            if (err) {
                record_error(p, rep, 0, PH_INIT, node, err);
            } else if (pts_noop<M>::value && idle) {
                if (p.fp_sum)
                    fused_counts<NC>(p, u, fx, fy);
                else
                    record_frozen<NC, ND>(p, g, 0, u, v);
                frozen_cols += (unsigned long long)p.tlen;
                frz = true;
            } else {
                push = true;
            }
            if (!frz)
                p.frozen_from[g] = p.tlen;
        }
        if (p.fp_sum) {
            /* Frozen nodes' counts into column 0 of their replicate's
             * difference array: summed over the wavefront's lanes of the
             * first lane's replicate and, separately, of the others (a
             * wavefront of consecutive lanes spans two replicates at most
             * when Nn >= 64; with smaller Nn each lane adds its own). */
            const long long gl = live ? g : p.lanes - 1;
            const int rep = (int)(gl / p.Nn);
            const int rep0 = __shfl(rep, 0);
            if (p.Nn >= 64) {
                const bool first = rep == rep0;
                long long a0 = first ? fx : 0, b0 = first ? fy : 0;
                long long a1 = first ? 0 : fx, b1 = first ? 0 : fy;
                for (int off = 32; off > 0; off >>= 1) {
                    a0 += __shfl_xor(a0, off);
                    b0 += __shfl_xor(b0, off);
                    a1 += __shfl_xor(a1, off);
                    b1 += __shfl_xor(b1, off);
                }
                const unsigned long long other = __ballot(!first);
                if (lane == 0 && (a0 || b0))
                    fused_add(p.fp_delta + 2 * ((long long)rep0 * (p.tlen + 1)), a0, b0);
                if (other && lane == __ffsll((long long)other) - 1 && (a1 || b1))
                    fused_add(p.fp_delta + 2 * ((long long)rep * (p.tlen + 1)), a1, b1);
            } else if (fx || fy) {
                fused_add(p.fp_delta + 2 * ((long long)rep * (p.tlen + 1)), fx, fy);
            }
        }
        const unsigned long long m = __ballot(push);
        if (lane == 0)
            s_cnt[wave] = (unsigned)__popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned tot = 0;
            for (int w = 0; w < nw; w++) {
                const unsigned c = s_cnt[w];
                s_cnt[w] = tot;
                tot += c;
            }
            s_base = tot ? atomicAdd(p.active_count, (int)tot) : 0;
        }
        __syncthreads();
        if (push)
            p.active[s_base + (int)s_cnt[wave] +
                     __popcll(m & ((1ull << lane) - 1ull))] = (int)g;
        __syncthreads();
    }
    for (int off = 32; off > 0; off >>= 1)
        frozen_cols += __shfl_xor(frozen_cols, off);
    if (lane == 0)
        s_frz[wave] = frozen_cols;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int w = 0; w < nw; w++)
            tot += s_frz[w];
        if (tot)
            atomicAdd(p.frozen_cols, tot);
    }
}

/* ssm / mssm over nodes that never interact after the classify pass
 * (modes NODES_ALL and NODES_ACTIVE): free-running lanes. */
template <class M, bool EVENTS, bool CANON, bool TSL>
__global__ __launch_bounds__(256) void k_ssm_nodes_free(KParams p)
{
    extern __shared__ double ts_sh[];
    tspan_to_lds<TSL>(p, ts_sh);
    p.seed = run_seed(p);
    for (WaveSlots a = wave_slots(p); a.w < a.waves; a.w += a.stride) {
        /* Lanes without a node help when the active pass packs fewer
         * than 64 nodes to a wavefront. */
        if (p.mode == NODES_ACTIVE && a.shift < 6)
            ssm_lane_free<M, EVENTS, CANON, true, TSL>(p, node_lane(p, a), a.shift, ts_sh);
        else
            ssm_lane_free<M, EVENTS, CANON, false, TSL>(p, node_lane(p, a), 6, ts_sh);
    }
}

/* One lane's aem run in a node kernel after the classify pass, with
 * free-running lanes as ssm_lane_free: one firing or one day's end per
 * trip round the loop (aem.c:181-300, the day's end :329-419).
 * Each channel keeps the logarithms of the uniforms at its next two
 * stream positions (l0, l1): a firing that draws from a channel takes l0
 * and shifts l1 down, and the next trip computes the channel's new l1
 * (one Philox block and one logarithm) in the same straight-line code as
 * its firing, which needs only l0.  A firing recomputes every rate G
 * marks and runs calcTimes for those channels and the fired one
 * (calc_times_sel, one division per channel), so lanes that fired
 * different transitions run the same instructions.  Values, stream
 * positions and results are aem_day's. */
template <class M, bool EVENTS, bool CANON, bool HELP, bool TSL>
__device__ __forceinline__ void aem_lane_free(const KParams &p, long long g, int shift,
                                              const double *ts_sh)
{
    constexpr int NC = M::NC, NT = M::NT, ND = M::ND;
    unsigned long long fired = 0;
    const bool live = g >= 0;
    const int node = live ? (int)g : 0;
    Ctx<M, CANON> cx;
    const uint64_t chan_base = (uint64_t)node * (uint64_t)NT;
    AemState<M> s;
    int err = 0;
    double tt = p.tt0, next = p.next0;
    const EvTab tab = global_tables(p);
    E1Cursor ec;
    auto frozen = [&]() __attribute__((always_inline)) -> bool {
        if constexpr (pts_noop<M>::value) {
            bool idle = !EVENTS || ec.cur >= ec.end;
            static_for<NT>([&](auto J) {
                idle = idle && s.rate[decltype(J)::value] == 0.0;
            });
            return idle;
        } else {
            return false;
        }
    };
    /* l0[k], l1[k]: siminf_log of channel k's uniforms at positions
     * s.j[k] and s.j[k] + 1; pend: channels whose l1 is still to be
     * computed (bit k). */
    double l0[NT], l1[NT];
    unsigned pend = 0;
    auto chan_log = [&](int k, unsigned long long j)
        __attribute__((always_inline)) -> double {
        const siminf_stream c = siminf_stream_make(cx.seed, SIMINF_DOMAIN_AEM,
                                                   chan_base + (uint64_t)k);
        return siminf_log_u01(siminf_uniform_at(&c, j));
    };
    bool going = false;
    if (live) {
        cx.load(p, node, 0);
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            s.u[c] = p.u0[g * NC + c];
        });
        if constexpr (ND > 0) {
            static_for<ND>([&](auto C) {
                constexpr int c = decltype(C)::value;
                s.v[c] = p.v0[g * ND + c];
                s.vnew[c] = s.v[c];
            });
        }
        s.rng = siminf_stream_make(p.seed, SIMINF_DOMAIN_NODE, (uint64_t)g);
        /* Initial rates and reaction times (aem.c:117-153). */
        static_for<NT>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const double r = M::template rate<j>(s.u, s.v, cx.ld, cx.gd, tt);
            s.rate[j] = r;
            if (!rate_ok(r))
                err = SIMINF_ERR_INVALID_RATE;
            const siminf_stream c = siminf_stream_make(
                p.seed, SIMINF_DOMAIN_AEM, chan_base + (uint64_t)j);
            double u0, u1, u2, u3;
            siminf_uniform_pair(&c, 0, &u0, &u1);
            siminf_uniform_pair(&c, 1, &u2, &u3);
            double tj = -siminf_log(u0) / r + tt;
            if (tj <= 0.0)
                tj = INFINITY;
            s.time[j] = tj;
            s.inf[j] = 0.0;
            s.j[j] = 1;
            l0[j] = siminf_log_u01(u1);
            l1[j] = siminf_log_u01(u2);
        });
        if constexpr (EVENTS)
            ec.start(p, node);
        if (err)
            record_error(p, 0, 0, PH_INIT, node, err);
        else
            going = true;
    } else {
        cx.clear(p);
        static_for<NC>([&](auto C) { s.u[decltype(C)::value] = 0; });
        static_for<(ND > 0 ? ND : 1)>([&](auto C) {
            s.v[decltype(C)::value] = 0.0;
            s.vnew[decltype(C)::value] = 0.0;
        });
        static_for<NT>([&](auto J) {
            constexpr int j = decltype(J)::value;
            s.rate[j] = 0.0;
            s.time[j] = INFINITY;
            s.inf[j] = 0.0;
            s.j[j] = 0;
            l0[j] = 0.0;
            l1[j] = 0.0;
        });
        s.rng = siminf_stream_make(p.seed, SIMINF_DOMAIN_NODE, 0);
    }
    int U_it = 0, day = 0;
    double ts = tspan_at<TSL>(p, ts_sh, 0);
    double ts1 = p.tlen > 1 ? tspan_at<TSL>(p, ts_sh, 1) : INFINITY;
    if (going && p.mode == NODES_ACTIVE && frozen()) {
        store_frozen<NC, ND>(p, g, U_it, s.u, s.vnew);
        going = false;
    }

    /* A draw from channel k: l1 moves down, the new l1 is pending. */
    auto took = [&](int k, bool drew) __attribute__((always_inline)) {
        static_for<NT>([&](auto K) {
            constexpr int c = decltype(K)::value;
            const bool d = drew && k == c;
            s.j[c] += d ? 1u : 0u;
            l0[c] = d ? l1[c] : l0[c];
            pend |= d ? 1u << c : 0u;
        });
    };
    /* The new l1 of the lowest pending channel (every lane computes one,
     * used where a channel is pending). */
    auto refill = [&]() __attribute__((always_inline)) {
        const int k = pend ? __builtin_ctz(pend) : 0;
        const unsigned long long jk = pick_u64(s.j, k);
        const double h = chan_log(k, jk + 1u);
        static_for<NT>([&](auto K) {
            constexpr int c = decltype(K)::value;
            l1[c] = pend && k == c ? h : l1[c];
        });
        pend &= pend - 1u;
    };

    /* With HELP the wavefront's A = 2^shift lanes hold nodes and all 64
     * lanes compute logarithms ahead into a table: lane l holds position
     * tb[c] + k of channel c of the node on lane l % A, with c and k the
     * remainder and quotient of (l / A) / NT (Q = 64 / (A NT) positions
     * per channel).  A pending l1 is then fetched from the table
     * (ds_bpermute) instead of computed; the wavefront refills the table
     * when some node's pending position is past its share, from each
     * channel's next needed position. */
    const int lane = (int)__lane_id();
    const int A = 1 << shift, Q = 64 / (A * NT);
    const int own = lane & (A - 1);
    const int tch = (lane >> shift) % NT, tk = (lane >> shift) / NT;
    unsigned long long tb[NT];
    static_for<NT>([&](auto K) { tb[decltype(K)::value] = 0; });
    double th = 0.0;
    bool filled = false;

    for (;;) {
        if (__ballot(going) == 0)
            break;
        /* At most one channel is pending here: its new l1, computed (or
         * fetched) beside the firing. */
        const unsigned pend0 = pend;
        const int kp = pend0 ? __builtin_ctz(pend0) : 0;
        const unsigned long long jp = pick_u64(s.j, kp);
        double hp;
        if constexpr (HELP) {
            const bool need = going && pend0 &&
                              jp + 1u - pick_u64(tb, kp) >= (unsigned long long)Q;
            if (__ballot(going && (!filled || need)) != 0) {
                static_for<NT>([&](auto K) {
                    constexpr int c = decltype(K)::value;
                    tb[c] = s.j[c] + (((pend0 >> c) & 1u) ? 1u : 2u);
                });
                const uint64_t cb =
                    ((uint64_t)lane_pull(own, (uint32_t)(chan_base >> 32)) << 32) |
                    lane_pull(own, (uint32_t)chan_base);
                unsigned long long to[NT];
                static_for<NT>([&](auto K) {
                    constexpr int c = decltype(K)::value;
                    to[c] = ((unsigned long long)lane_pull(own, (uint32_t)(tb[c] >> 32)) << 32) |
                            lane_pull(own, (uint32_t)tb[c]);
                });
                const siminf_stream c = siminf_stream_make(cx.seed, SIMINF_DOMAIN_AEM,
                                                           cb + (uint64_t)tch);
                th = siminf_log_u01(siminf_uniform_at(&c, pick_u64(to, tch) + (uint64_t)tk));
                filled = true;
            }
            const int src = going && pend0
                ? ((int)(jp + 1u - pick_u64(tb, kp)) * NT + kp) * A + own : lane;
            hp = lane_pull(src, th);
        } else {
            hp = chan_log(kp, jp + 1u);
        }

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
        const bool fires = going && !(isinf(t) || t >= next);
        int e = 0;
        unsigned drew = 0;
        if constexpr (CANON && NT <= 8) {
            int un[NC];
            static_for<NC>([&](auto C) {
                constexpr int c = decltype(C)::value;
                int d = 0;
                static_for<NT>([&](auto K) {
                    constexpr int k = decltype(K)::value;
                    if constexpr (M::S[k][c] != 0)
                        d += tr == k ? M::S[k][c] : 0;
                });
                un[c] = s.u[c] + d;
                if (un[c] < 0)
                    e = SIMINF_ERR_NEGATIVE_STATE;
            });
            static_for<NT>([&](auto J) {
                constexpr int j = decltype(J)::value;
                constexpr unsigned gm = g_row_mask<M, j>() | (1u << j);
                const bool in = fires && ((gm >> tr) & 1u);
                const double old = s.rate[j];
                const double r = M::template rate<j>(un, s.v, cx.ld, cx.gd, t);
                s.rate[j] = in ? r : old;
                if (in && !rate_ok(r))
                    e = SIMINF_ERR_INVALID_RATE;
                drew |= calc_times_sel<M, j>(s, l0[j], t, old, r, in) ? 1u << j : 0u;
            });
            static_for<NC>([&](auto C) {
                constexpr int c = decltype(C)::value;
                s.u[c] = fires ? un[c] : s.u[c];
            });
            e = fires ? e : 0;
        } else if (fires) {
            static_for<NT>([&](auto K) {
                constexpr int k = decltype(K)::value;
                if (tr == k) {
                    static_for<NC>([&](auto C) {
                        constexpr int c = decltype(C)::value;
                        if (cx.template s_stored<k, c>()) {
                            s.u[c] += cx.template s_val<k, c>();
                            if (s.u[c] < 0)
                                e = SIMINF_ERR_NEGATIVE_STATE;
                        }
                    });
                    static_for<NT>([&](auto J) {
                        constexpr int j = decltype(J)::value;
                        if (j == k || cx.template g_stored<k, j>()) {
                            const double old = s.rate[j];
                            const double r = M::template rate<j>(s.u, s.v, cx.ld, cx.gd, t);
                            s.rate[j] = r;
                            if (!rate_ok(r))
                                e = SIMINF_ERR_INVALID_RATE;
                            drew |= calc_times_sel<M, j>(s, l0[j], t, old, r, true)
                                    ? 1u << j : 0u;
                        }
                    });
                }
            });
        }
        fired += fires ? 1u : 0u;
        /* The pending l1 lands before this firing's draws move l1 down. */
        static_for<NT>([&](auto K) {
            constexpr int c = decltype(K)::value;
            l1[c] = pend0 && kp == c ? hp : l1[c];
        });
        pend &= pend - 1u;
        static_for<NT>([&](auto K) {
            constexpr int c = decltype(K)::value;
            took(c, (drew >> c) & 1u);
            pin(l0[c]);
            pin(l1[c]);
        });
        if (e) {
            record_error(p, 0, day, PH_SSA, node, e);
            going = false;
        }
        const bool dend = going && !fires;
        bool slow = dend;
        if constexpr (!EVENTS && pts_noop<M>::value && ND == 0) {
            /* The fast day end of ssm_lane_free. */
            if (p.U || p.fp_sum) {
                bool live_rates = false;
                static_for<NT>([&](auto J) {
                    live_rates = live_rates || s.rate[decltype(J)::value] != 0.0;
                });
                slow = dend && !day_end_fast<NC, TSL>(p, 0, node, dend, live_rates, s.u, tt,
                                                       next, U_it, ts, ts1, ts_sh);
                day += dend && !slow ? 1 : 0;
            }
        }
        if (__ballot(slow) != 0 && slow) {
            bool upd = false;
            bool ok = true;
            if constexpr (EVENTS) {
                int bad = 0;
                err = e1_day<NC>(p, tab, ec, tt, s.u, &s.rng, upd, bad);
                if (err) {
                    record_error(p, 0, day, PH_E1, p.ev_rank[bad], err);
                    going = false;
                    ok = false;
                }
            }
            /* Post time step; refresh through calcTimes at the next unit
             * of time (aem.c:329-384).  A refresh may draw from several
             * channels: their l1 are computed below. */
            if (ok) {
                const int rc = M::pts(s.vnew, s.u, s.v, cx.ld, cx.gd, node, tt);
                if (rc < 0) {
                    record_error(p, 0, day, PH_PTS, node, rc);
                    going = false;
                    ok = false;
                } else if (rc > 0 || upd) {
                    unsigned dr = 0;
                    static_for<NT>([&](auto J) {
                        constexpr int j = decltype(J)::value;
                        const double old = s.rate[j];
                        const double r = M::template rate<j>(s.u, s.vnew, cx.ld, cx.gd, tt);
                        s.rate[j] = r;
                        if (!rate_ok(r))
                            err = SIMINF_ERR_INVALID_RATE;
                        dr |= calc_times_sel<M, j>(s, l0[j], next, old, r, true) ? 1u << j : 0u;
                    });
                    static_for<NT>([&](auto K) {
                        constexpr int c = decltype(K)::value;
                        took(c, (dr >> c) & 1u);
                    });
                    if (err) {
                        record_error(p, 0, day, PH_PTS, node, err);
                        going = false;
                        ok = false;
                    }
                }
            }
            if (ok) {
                tt = next;
                next += 1.0;
                store_columns_ahead<NC, ND, TSL>(p, 0, node, tt, U_it, ts, ts1, s.u, s.vnew, ts_sh);
                if constexpr (ND > 0) {
                    static_for<ND>([&](auto C) {
                        constexpr int c = decltype(C)::value;
                        const double x = s.v[c];
                        s.v[c] = s.vnew[c];
                        s.vnew[c] = x;
                    });
                }
                if (U_it >= p.tlen)
                    going = false;
                day++;
                if (going && p.mode == NODES_ACTIVE && frozen()) {
                    store_frozen<NC, ND>(p, g, U_it, s.u, s.vnew);
                    going = false;
                }
            }
        }
        /* Leave at most one pending channel for the next trip. */
        while (__ballot(going && (pend & (pend - 1u)) != 0) != 0) {
            if (going && (pend & (pend - 1u)) != 0)
                refill();
        }
    }
    add_transitions(p, fired);
}

/* aem with each node's channels on lanes of their own, for models whose S
 * and G fold into the code and that have at most four transitions: a
 * node holds a quad of lanes, lane c of the quad runs channel c (its
 * rate, next firing time, reactInf and stream) and every lane of the quad
 * keeps the node's state.  A firing is the quad's earliest channel (DPP
 * moves, the first channel winning exact ties as in aem_day); each lane
 * then applies the transition, recomputes its own channel's rate if G
 * marks it and runs calcTimes for it, so the channels' divisions run side
 * by side instead of one after another.  A channel's next logarithms (l0,
 * l1) are kept as in aem_lane_free; with HELP the wavefront's lanes
 * compute l1 values ahead into a table: for A = 2^shift nodes a channel
 * lane s = 4 q + c owns positions tb .. tb + Q - 1 (Q = 16 / A) held by
 * lanes s, s + 4A, ...  Values, stream positions and results are
 * aem_day's. */
template <class M, bool EVENTS, bool TSL, bool HELP>
__device__ __forceinline__ void aem_lane_quad(const KParams &p, long long g, int shift,
                                              const double *ts_sh)
{
    constexpr int NC = M::NC, NT = M::NT, ND = M::ND;
    static_assert(NT <= 4, "a quad of lanes per node");
    const int lane = (int)__lane_id();
    const int c = lane & 3;
    const bool chan = c < NT;
    const bool lead = c == 0;
    unsigned long long fired = 0;
    const bool live = g >= 0;
    const int node = live ? (int)g : 0;
    Ctx<M, true> cx;
    const uint64_t chan_base = (uint64_t)node * (uint64_t)NT;
    int u[NC];
    double v[ND > 0 ? ND : 1], vnew[ND > 0 ? ND : 1];
    /* This lane's channel: rate, next firing time, reactInf, stream
     * position and the logarithms of the uniforms at j and j + 1. */
    double rate = 0.0, ctime = INFINITY, rinf = 0.0, l0 = 0.0, l1 = 0.0;
    unsigned long long j = 0;
    bool pend = false;
    siminf_stream rng = siminf_stream_make(p.seed, SIMINF_DOMAIN_NODE, live ? (uint64_t)g : 0);
    int err = 0;
    double tt = p.tt0, next = p.next0;
    const EvTab tab = global_tables(p);
    E1Cursor ec;
    /* The rate of this lane's channel at the given state. */
    auto own_rate = [&](const int (&uu)[NC], const double *vv, double t)
        __attribute__((always_inline)) -> double {
        double r = 0.0;
        static_for<NT>([&](auto K) {
            constexpr int k = decltype(K)::value;
            const double rk = M::template rate<k>(uu, vv, cx.ld, cx.gd, t);
            r = c == k ? rk : r;
        });
        return r;
    };
    auto chan_log = [&](unsigned long long pos) __attribute__((always_inline)) -> double {
        const siminf_stream cs = siminf_stream_make(cx.seed, SIMINF_DOMAIN_AEM,
                                                    chan_base + (uint64_t)(chan ? c : 0));
        return siminf_log_u01(siminf_uniform_at(&cs, pos));
    };
    auto frozen = [&]() __attribute__((always_inline)) -> bool {
        if constexpr (pts_noop<M>::value)
            return (!EVENTS || ec.cur >= ec.end) && !quad_any(chan && rate != 0.0);
        else
            return false;
    };
    /* A draw by this lane's channel: l1 moves down, the new l1 pends. */
    auto took = [&](bool drew) __attribute__((always_inline)) {
        j += drew ? 1u : 0u;
        l0 = drew ? l1 : l0;
        pend = pend || drew;
    };

    bool going = false;
    if (live) {
        cx.load(p, node, 0);
        static_for<NC>([&](auto C) {
            constexpr int cc = decltype(C)::value;
            u[cc] = p.u0[g * NC + cc];
        });
        if constexpr (ND > 0) {
            static_for<ND>([&](auto C) {
                constexpr int cc = decltype(C)::value;
                v[cc] = p.v0[g * ND + cc];
                vnew[cc] = v[cc];
            });
        }
    } else {
        cx.clear(p);
        static_for<NC>([&](auto C) { u[decltype(C)::value] = 0; });
        static_for<(ND > 0 ? ND : 1)>([&](auto C) {
            v[decltype(C)::value] = 0.0;
            vnew[decltype(C)::value] = 0.0;
        });
    }
    {
        /* Initial rates and reaction times (aem.c:117-153). */
        bool bad = false;
        static_for<NT>([&](auto K) {
            constexpr int k = decltype(K)::value;
            bad = bad || !rate_ok(M::template rate<k>(u, v, cx.ld, cx.gd, tt));
        });
        rate = chan ? own_rate(u, v, tt) : 0.0;
        const siminf_stream cs = siminf_stream_make(p.seed, SIMINF_DOMAIN_AEM,
                                                    chan_base + (uint64_t)(chan ? c : 0));
        double u0, u1, u2, u3;
        siminf_uniform_pair(&cs, 0, &u0, &u1);
        siminf_uniform_pair(&cs, 1, &u2, &u3);
        double tj = -siminf_log(u0) / rate + tt;
        tj = tj <= 0.0 ? INFINITY : tj;
        ctime = chan ? tj : INFINITY;
        j = 1;
        l0 = siminf_log_u01(u1);
        l1 = siminf_log_u01(u2);
        if (live) {
            if constexpr (EVENTS)
                ec.start(p, node);
            if (bad) {
                if (lead)
                    record_error(p, 0, 0, PH_INIT, node, SIMINF_ERR_INVALID_RATE);
            } else {
                going = true;
            }
        }
    }
    int U_it = 0, day = 0;
    double ts = tspan_at<TSL>(p, ts_sh, 0);
    double ts1 = p.tlen > 1 ? tspan_at<TSL>(p, ts_sh, 1) : INFINITY;
    if (going && p.mode == NODES_ACTIVE && frozen()) {
        if (lead)
            store_frozen<NC, ND>(p, g, U_it, u, vnew);
        going = false;
    }

    const int A = 1 << shift, L = 4 << shift, Q = 16 >> shift;
    const int tslot = lane / L, towner = lane % L;
    unsigned long long tb = 0;
    double th = 0.0;
    bool filled = false;
    (void)A;

    for (;;) {
        if (__ballot(going) == 0)
            break;
        /* The pending l1 of this lane's channel, fetched or computed
         * beside the firing. */
        const bool pend0 = pend && going;
        double hp;
        if constexpr (HELP) {
            if (__ballot(pend0 && (!filled || j + 1u - tb >= (unsigned long long)Q)) != 0) {
                tb = j + (pend ? 1u : 2u);
                const uint64_t cb = ((uint64_t)lane_pull(towner, (uint32_t)(chan_base >> 32)) << 32) |
                                    lane_pull(towner, (uint32_t)chan_base);
                const uint64_t to = ((uint64_t)lane_pull(towner, (uint32_t)(tb >> 32)) << 32) |
                                    lane_pull(towner, (uint32_t)tb);
                const int oc = towner & 3;
                const siminf_stream cs = siminf_stream_make(cx.seed, SIMINF_DOMAIN_AEM,
                                                            cb + (uint64_t)(oc < NT ? oc : 0));
                th = siminf_log_u01(siminf_uniform_at(&cs, to + (uint64_t)tslot));
                filled = true;
            }
            const int src = pend0 ? lane + (int)(j + 1u - tb) * L : lane;
            hp = lane_pull(src, th);
        } else {
            hp = chan_log(j + 1u);
        }

        /* The quad's earliest channel. */
        double t = ctime;
        int tr = chan ? c : 4;
        quad_argmin(t, tr);
        const bool fires = going && !(isinf(t) || t >= next);
        /* The transition's state change and this lane's channel. */
        int un[NC];
        bool neg = false;
        static_for<NC>([&](auto C) {
            constexpr int cc = decltype(C)::value;
            int d = 0;
            static_for<NT>([&](auto K) {
                constexpr int k = decltype(K)::value;
                if constexpr (M::S[k][cc] != 0)
                    d += tr == k ? M::S[k][cc] : 0;
            });
            un[cc] = u[cc] + d;
            neg = neg || un[cc] < 0;
        });
        unsigned gm = 0;
        static_for<NT>([&](auto K) {
            constexpr int k = decltype(K)::value;
            gm = c == k ? (g_row_mask<M, k>() | (1u << k)) : gm;
        });
        const bool in = fires && chan && ((gm >> tr) & 1u);
        const double r = own_rate(un, v, t);
        const bool drew = calc_times_one(ctime, rinf, l0, t, rate, r, in);
        const bool badr = quad_any(in && !rate_ok(r));
        rate = in ? r : rate;
        static_for<NC>([&](auto C) {
            constexpr int cc = decltype(C)::value;
            u[cc] = fires ? un[cc] : u[cc];
        });
        fired += fires && lead ? 1u : 0u;
        /* The pending l1 lands before this firing's draw moves it down. */
        l1 = pend0 ? hp : l1;
        pend = pend && !pend0;
        took(drew);
        pin(l0);
        pin(l1);
        const int e = fires ? (badr ? SIMINF_ERR_INVALID_RATE
                                    : (neg ? SIMINF_ERR_NEGATIVE_STATE : 0)) : 0;
        if (e) {
            if (lead)
                record_error(p, 0, day, PH_SSA, node, e);
            going = false;
        }
        const bool dend = going && !fires;
        bool slow = dend;
        if constexpr (!EVENTS && pts_noop<M>::value && ND == 0) {
            if (p.U || p.fp_sum) {
                const bool alive = quad_any(chan && rate != 0.0);
                slow = dend && !day_end_fast<NC, TSL>(p, 0, node, dend, alive, u, tt, next,
                                                       U_it, ts, ts1, ts_sh, lead);
                day += dend && !slow ? 1 : 0;
            }
        }
        if (__ballot(slow) != 0 && slow) {
            bool upd = false;
            bool ok = true;
            if constexpr (EVENTS) {
                int bad = 0;
                err = e1_day<NC>(p, tab, ec, tt, u, &rng, upd, bad);
                if (err) {
                    if (lead)
                        record_error(p, 0, day, PH_E1, p.ev_rank[bad], err);
                    going = false;
                    ok = false;
                }
            }
            if (ok) {
                const int rc = M::pts(vnew, u, v, cx.ld, cx.gd, node, tt);
                if (rc < 0) {
                    if (lead)
                        record_error(p, 0, day, PH_PTS, node, rc);
                    going = false;
                    ok = false;
                } else if (rc > 0 || upd) {
                    bool bad = false;
                    static_for<NT>([&](auto K) {
                        constexpr int k = decltype(K)::value;
                        bad = bad || !rate_ok(M::template rate<k>(u, vnew, cx.ld, cx.gd, tt));
                    });
                    const double rr = own_rate(u, vnew, tt);
                    took(calc_times_one(ctime, rinf, l0, next, rate, rr, chan));
                    rate = chan ? rr : rate;
                    if (bad) {
                        if (lead)
                            record_error(p, 0, day, PH_PTS, node, SIMINF_ERR_INVALID_RATE);
                        going = false;
                        ok = false;
                    }
                }
            }
            if (ok) {
                tt = next;
                next += 1.0;
                store_columns_ahead<NC, ND, TSL>(p, 0, node, tt, U_it, ts, ts1, u, vnew, ts_sh,
                                                 lead);
                if constexpr (ND > 0) {
                    static_for<ND>([&](auto C) {
                        constexpr int cc = decltype(C)::value;
                        const double x = v[cc];
                        v[cc] = vnew[cc];
                        vnew[cc] = x;
                    });
                }
                if (U_it >= p.tlen)
                    going = false;
                day++;
                if (going && p.mode == NODES_ACTIVE && frozen()) {
                    if (lead)
                        store_frozen<NC, ND>(p, g, U_it, u, vnew);
                    going = false;
                }
            }
        }
        /* A refresh's draw may leave a second l1 pending: none is, as
         * each lane runs one channel and its pending l1 landed above. */
    }
    add_transitions(p, fired);
}

/* aem over nodes that never interact after the classify pass, by wave
 * slots as k_ssm_nodes_free. */
template <class M, bool EVENTS, bool CANON, bool TSL>
__global__ __launch_bounds__(256) void k_aem_nodes_free(KParams p)
{
    extern __shared__ double ts_sh[];
    tspan_to_lds<TSL>(p, ts_sh);
    p.seed = run_seed(p);
    for (WaveSlots a = wave_slots(p); a.w < a.waves; a.w += a.stride) {
        if constexpr (CANON && M::NT <= 4) {
            /* A quad of lanes per node when the active pass packs at most
             * 16 nodes to a wavefront. */
            if (p.mode == NODES_ACTIVE && a.shift <= 4) {
                if (a.shift <= 3)
                    aem_lane_quad<M, EVENTS, TSL, true>(p, node_lane_quad(p, a), a.shift, ts_sh);
                else
                    aem_lane_quad<M, EVENTS, TSL, false>(p, node_lane_quad(p, a), a.shift, ts_sh);
                continue;
            }
        }
        /* Lanes without a node help when the active pass packs few
         * enough nodes to a wavefront for two positions per channel. */
        if (p.mode == NODES_ACTIVE && (64 >> a.shift) / M::NT >= 2)
            aem_lane_free<M, EVENTS, CANON, true, TSL>(p, node_lane(p, a), a.shift, ts_sh);
        else
            aem_lane_free<M, EVENTS, CANON, false, TSL>(p, node_lane(p, a), 6, ts_sh);
    }
}

/* ------------------------------------------------------------------ */
/* Coupled models: external transfers (E2) link the nodes every day.    */
/* ------------------------------------------------------------------ */

/* Has an error been recorded at or before (rep, day, phase)?  Read after
 * a barrier, this gives every block the same answer: errors of later
 * phases, which other blocks may be writing meanwhile, are ignored. */
__device__ __forceinline__ bool error_by(const KParams &p, int rep, int day,
                                         int phase)
{
    const unsigned long long k = __hip_atomic_load(
        p.rep_err ? p.rep_err + rep : p.err_key, __ATOMIC_RELAXED,
        __HIP_MEMORY_SCOPE_AGENT);
    return (k >> 27) <= (error_key(rep, day, phase, 0, -1) >> 27);
}

/* Uniforms computed ahead per external transfer event (even). */
#define SIMINF_E2_PRE SIMINF_PRE_MAX

/* The levels of one day's external transfers over a state array su (NC
 * ints per slot) with update flags sf, event records rec/prop and level
 * bounds lvl (absolute positions; the day starts at i0).  Instantiated
 * once for LDS operands and once for global ones, so that each gets its
 * own memory instructions instead of generic ones. */
template <int NC, bool LDS>
__device__ __forceinline__ int e2_levels(const KParams &p, const EvTab &tab,
                                         int rep, int day, int *su, int *sf,
                                         const int4 *rec, const double *prop,
                                         const int *lvl, int nlv, int i0,
                                         const double *pre)
{
    int myerr = 0;
    const int tid = (int)threadIdx.x, nth = (int)blockDim.x;
    for (int l = 0; l < nlv; l++) {
        const int a = lvl[l] - i0, b = lvl[l + 1] - i0;
        for (int i = a + tid; i < b; i += nth) {
            const int4 r0 = rec[2 * i], r1 = rec[2 * i + 1];
            siminf_stream rng = siminf_stream_make(
                p.seed, SIMINF_DOMAIN_E2,
                (uint64_t)rep * (uint64_t)p.len + (uint64_t)r0.x);
            if constexpr (LDS) {
                rng.pre = pre + (long long)i * SIMINF_E2_PRE;
                rng.npre = SIMINF_E2_PRE;
            }
            const int err =
                r0.w == SIMINF_E2_FAST
                    ? apply_e2_fast<NC>(r0.y, r0.z, su, sf, LDS ? r1.x : r1.z,
                                        LDS ? r1.y : r1.w, &rng)
                    : apply_e2<NC>(tab, r0.y, r0.z, r0.w, prop[i], r1.w, p.Nn, su, sf,
                                   LDS ? r1.x : r1.z, LDS ? r1.y : r1.w, &rng);
            if (err) {
                record_error(p, rep, day, PH_E2, p.ev_rank[r0.x], err);
                myerr = 1;
            }
        }
        __syncthreads();
    }
    return myerr;
}

/* The levels of a staged day prepared for the fast path (every event a
 * count, no shift, one E row or two equal-weight rows; the records' second
 * words hold the four counts' byte offsets in lu; levels of at most 64
 * events), run by one wavefront with lane = event within a level.  Every
 * lane runs every level: a lane without an event moves nothing between
 * four spare ints after the herds' counts, so each level is the same
 * straight sequence of instructions.  A wavefront's LDS accesses complete
 * in issue order, so a level reads what the one before it wrote without
 * a barrier.  The level bounds sit in two registers (read with
 * v_readlane), and the next level's records are loaded while the current
 * level runs, so the chain from level to level is one round of LDS loads
 * (both herds' counts and the event's uniforms), the urn and the stores.
 * The touched herds are all marked updated by the caller.  Needs
 * nlv < 128.  Returns 1 if this lane recorded an error. */
template <int NC>
__device__ __forceinline__ int e2_levels_fast(const KParams &p, int rep, int day,
                                              unsigned char *lu, const int4 *rec,
                                              const int *lvl, int nlv, int i0,
                                              const double *pre)
{
    const int lane = (int)threadIdx.x;
    int myerr = 0;
    const int bnd0 = lane <= nlv ? lvl[lane] - i0 : 0;
    const int bnd1 = 64 + lane <= nlv ? lvl[64 + lane] - i0 : 0;
    auto bound = [&](int l) -> int {
        return l < 64 ? __builtin_amdgcn_readlane(bnd0, l)
                      : __builtin_amdgcn_readlane(bnd1, l - 64);
    };
    const int spare = p.e2_cap_t * NC * (int)sizeof(int);
    const int4 idle0 = int4{0, 0, 0, 0}, idle1 = int4{spare, spare + 4, spare + 8, spare + 12};
    /* Lane's record of the level [a, b): loaded from a valid position in
     * any case, replaced by the idle one when the level has no event for
     * the lane. */
    auto load = [&](int a, int b, int4 &r0, int4 &r1, int &i) {
        const bool has = lane < b - a;
        i = has ? a + lane : a;
        r0 = rec[2 * i];
        r1 = rec[2 * i + 1];
        if (!has) {
            r0 = idle0;
            r1 = idle1;
        }
    };
    int a = bound(0), b = bound(1), i = 0;
    int4 c0, c1;
    load(a, b, c0, c1, i);
    for (int l = 0; l < nlv; l++) {
        /* This level's record, loaded during the level before, is taken
         * into registers before the next level's loads are issued: the
         * wait for it is then not a wait for those. */
        const int ev = c0.x, n = c0.y, rows = c0.z, ic = i;
        const int4 off = c1;
        asm volatile("" ::"v"(ev), "v"(n), "v"(rows), "v"(off.x), "v"(off.y), "v"(off.z),
                     "v"(off.w), "v"(ic)
                     : "memory");
        if (l + 1 < nlv) {
            a = bound(l + 1);
            b = bound(l + 2);
            load(a, b, c0, c1, i);
        }
        const int err = apply_e2_fast_pre(n, rows, lu, off,
                                          pre + (long long)ic * SIMINF_E2_PRE, p.seed,
                                          (uint64_t)rep * (uint64_t)p.len + (uint64_t)ev);
        if (err) {
            record_error(p, rep, day, PH_E2, p.ev_rank[ev], err);
            myerr = 1;
        }
    }
    return myerr;
}

/* A day's external transfers as staged for the workgroup that runs them:
 * the day's levels and touched herds, and where the levels run. */
struct E2Day {
    int l0, nlv, i0, nev, t0, nt;
    int mode; /* E2_NONE, E2_GLOBAL, E2_LDS (any events), E2_LDS_FAST */
};
enum { E2_NONE = 0, E2_GLOBAL = 1, E2_LDS = 2, E2_LDS_FAST = 3 };

/* The part of a day's E2 staging that does not depend on the herds'
 * state: the day's event records, proportions, level bounds and touched
 * herds copied to LDS with coalesced loads, and the first uniforms of
 * every event's stream computed in parallel (so that the levels, which
 * run one after another, find them in LDS instead of each running Philox
 * on the critical path).  A day larger than the LDS buffers runs from
 * global memory.  Called by the whole workgroup; in a grid launch the
 * workgroup that runs E2 does this while the others run the day's SSA. */
template <int NC>
__device__ __forceinline__ E2Day e2_stage(const KParams &p, int rep, int day,
                                          unsigned char *lds)
{
    E2Day d{0, 0, 0, 0, 0, 0, E2_NONE};
    /* The plan's day (tables) and the run's day (error order). */
    const int pd = day + p.day0;
    if (pd >= p.ndays || pd <= p.e2_floor)
        return d;
    d.l0 = p.e2_day[pd];
    d.nlv = p.e2_day[pd + 1] - d.l0;
    if (d.nlv == 0)
        return d;
    d.i0 = p.e2_lvl[d.l0];
    d.nev = p.e2_lvl[d.l0 + d.nlv] - d.i0;
    d.t0 = p.e2_tn_ptr[pd];
    d.nt = p.e2_tn_ptr[pd + 1] - d.t0;
    if (!(d.nt <= p.e2_cap_t && d.nev <= p.e2_cap_ev && d.nlv <= p.e2_cap_lv)) {
        d.mode = E2_GLOBAL;
        return d;
    }
    const int4 *rec = p.e2_rec + 2 * (long long)d.i0;
    const double *prop = p.e2_prop + d.i0;
    const int *lvl = p.e2_lvl + d.l0;
    const int tid = (int)threadIdx.x, nth = (int)blockDim.x;
    int4 *lrec = (int4 *)(lds + p.lds_rec);
    double *lprop = (double *)(lds + p.lds_prop);
    int *llvl = (int *)(lds + p.lds_lvl);
    int *ltn = (int *)(lds + p.lds_tn);
    double *lpre = (double *)(lds + p.lds_pre);
    for (int i = tid; i < 2 * d.nev; i += nth)
        lrec[i] = rec[i];
    for (int i = tid; i < d.nev; i += nth)
        lprop[i] = prop[i];
    for (int i = tid; i <= d.nlv; i += nth)
        llvl[i] = lvl[i];
    for (int i = tid; i < d.nt; i += nth)
        ltn[i] = p.e2_tn[d.t0 + i];
    for (int q = tid; q < d.nev * (SIMINF_E2_PRE / 2); q += nth) {
        const int i = q / (SIMINF_E2_PRE / 2), b = q % (SIMINF_E2_PRE / 2);
        const siminf_stream st = siminf_stream_make(
            p.seed, SIMINF_DOMAIN_E2,
            (uint64_t)rep * (uint64_t)p.len + (uint64_t)rec[2 * i].x);
        double u0, u1;
        siminf_uniform_pair(&st, (uint64_t)b, &u0, &u1);
        lpre[i * SIMINF_E2_PRE + 2 * b] = u0;
        lpre[i * SIMINF_E2_PRE + 2 * b + 1] = u1;
    }
    __syncthreads();
    d.mode = p.e2_day_fast[pd] ? E2_LDS_FAST : E2_LDS;
    return d;
}

/* One day's external transfers of replicate rep, run by one workgroup
 * after e2_stage: the touched herds are loaded into LDS, the
 * conflict-free levels run there (one wavefront without barriers when
 * every event takes the fast path, else a workgroup barrier between
 * levels; no atomics: events in one level touch disjoint herds, and
 * within a herd the events keep list order), and the herds are
 * scattered back with their update flags.  Returns 1 if this thread
 * recorded an error. */
template <int NC>
__device__ __forceinline__ int e2_run(const KParams &p, const EvTab &tab, int rep, int day,
                                      int *uw, int *updw, unsigned char *lds,
                                      const E2Day &d)
{
    if (d.mode == E2_NONE)
        return 0;
    if (d.mode == E2_GLOBAL)
        return e2_levels<NC, false>(p, tab, rep, day, uw, updw,
                                    p.e2_rec + 2 * (long long)d.i0, p.e2_prop + d.i0,
                                    p.e2_lvl + d.l0, d.nlv, d.i0, nullptr);

    const int tid = (int)threadIdx.x, nth = (int)blockDim.x;
    const EvTab ltab{(const uint32_t *)(lds + p.lds_Emask),
                     (const double *)(lds + p.lds_Ew),
                     p.Nmat ? (const int *)(lds + p.lds_N) : nullptr};
    int *lu = (int *)(lds + p.lds_u);
    int *lupd = (int *)(lds + p.lds_upd);
    const int4 *lrec = (const int4 *)(lds + p.lds_rec);
    const double *lprop = (const double *)(lds + p.lds_prop);
    const int *llvl = (const int *)(lds + p.lds_lvl);
    const int *ltn = (const int *)(lds + p.lds_tn);
    const double *lpre = (const double *)(lds + p.lds_pre);
    for (int i = tid; i < d.nt; i += nth) {
        const int node = ltn[i];
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            lu[i * NC + c] = uw[node * NC + c];
        });
        lupd[i] = 0;
    }
    if (tid < 4)
        lu[p.e2_cap_t * NC + tid] = 0; /* the idle lanes' spare counts */
    __syncthreads();
    int myerr;
    if (d.mode == E2_LDS_FAST) {
        myerr = tid < 64 ? e2_levels_fast<NC>(p, rep, day, (unsigned char *)lu, lrec, llvl,
                                              d.nlv, d.i0, lpre)
                         : 0;
        __syncthreads();
    } else {
        myerr = e2_levels<NC, true>(p, ltab, rep, day, lu, lupd, lrec, lprop, llvl,
                                    d.nlv, d.i0, lpre);
    }
    for (int i = tid; i < d.nt; i += nth) {
        const int node = ltn[i];
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            uw[node * NC + c] = lu[i * NC + c];
        });
        if (d.mode == E2_LDS_FAST || lupd[i])
            updw[node] = 1;
    }
    return myerr;
}

/* Day loop with one lane per node and the node's state resident in VGPRs
 * across days.  GRID: one replicate spread over the workgroups of a
 * cooperative launch, joined by grid barriers; one more workgroup, with
 * no herds, runs E2 and stages each day's event data while the others
 * run the day's SSA.  Otherwise: one workgroup per replicate, joined by
 * workgroup barriers.
 * Per day:  A  SSA (or, with AEM, the all-events method) + E1 of the
 *              lane's node, its u published to global memory (E2
 *              workgroup: e2_stage);                        barrier
 *           B  the day's E2 events (e2_run);              barrier
 *           C  reload u, post time step, refresh, writeback.
 * (SimInf_solver_ssm.c:134-362, SimInf_solver_mssm.c:158-383.) */
template <class M, bool CANON, bool GRID, int BLOCK, bool AEM = false>
__global__ __launch_bounds__(BLOCK) void k_coupled(KParams p)
{
    constexpr int NC = M::NC, NT = M::NT, ND = M::ND;
    p.seed = run_seed(p);
    extern __shared__ int4 lds4[];
    unsigned char *lds = (unsigned char *)lds4;
    /* Grid launch: G workgroups per replicate, the launch's replicate rb
     * of them (rep0 + rb), local the workgroup within its replicate. */
    const int G = GRID ? (p.rep_blocks > 0 ? p.rep_blocks : (int)gridDim.x) : 1;
    const int rb = GRID ? (int)blockIdx.x / G : 0;
    const int local = GRID ? (int)blockIdx.x - rb * G : 0;
    const int rep = GRID ? p.rep0 + rb : (int)blockIdx.x;
    const int lane_g = GRID ? (int)(local * blockDim.x + threadIdx.x) : (int)threadIdx.x;
    unsigned *gsync = GRID ? p.gsync + SIMINF_GSYNC_WORDS * rb : nullptr;
    /* A post-time-step function that reads other nodes gets pointers into
     * the replicate's published u and v (node-major) and the node's whole
     * ldata column. */
    constexpr bool NB = model_neighbors<M>::value;
    const int node = GRID && p.lane_node ? p.lane_node[lane_g] : lane_g;
    const bool active = node >= 0 && node < p.Nn;
    const long long base = (long long)rep * p.Nn;
    const long long g = base + node;
    int *uw = p.u_work + base * NC;
    int *updw = p.upd_work + base;
    double *vw = NB ? p.v_work + base * ND : nullptr;
    unsigned long long fired = 0;
    double tt = p.tt0, next = p.next0;
    int U_it = 0;
    double ts = p.tspan[0];
    int myerr = 0;

    std::conditional_t<AEM, AemState<M>, NodeState<M>> s;
    Ctx<M, CANON> cx;
    E1Cursor ec;
    int hcur = 0, hend = 0, hnext = INT_MAX;
    const uint64_t chan_base = (uint64_t)g * (uint64_t)NT;
    if constexpr (!AEM)
        s.a0 = 0.0;

    const EvTab tab = global_tables(p);

    /* The workgroup that runs the E2 levels: in a grid launch the last,
     * which holds no herds. */
    const bool e2wg = !GRID || local == G - 1;
    E2Day e2d{0, 0, 0, 0, 0, 0, E2_NONE};

    /* E and N into LDS for the workgroup that runs the E2 levels. */
    if (e2wg) {
        if (p.e2_cap_ev > 0) {
            uint32_t *lEmask = (uint32_t *)(lds + p.lds_Emask);
            double *lEw = (double *)(lds + p.lds_Ew);
            for (int i = threadIdx.x; i < p.Nselect; i += blockDim.x)
                lEmask[i] = p.Emask[i];
            for (int i = threadIdx.x; i < p.Nselect * NC; i += blockDim.x)
                lEw[i] = p.Ew[i];
            if (p.Nmat) {
                int *lN = (int *)(lds + p.lds_N);
                for (int i = threadIdx.x; i < p.Nshift * NC; i += blockDim.x)
                    lN[i] = p.Nmat[i];
            }
        }
        __syncthreads();
    }

    /* Grid launch: every workgroup is one wavefront.  gsync[0] counts
     * the replicate's arrivals (each of its G workgroups once at the
     * start, then each herd workgroup once a day after its SSA and E1);
     * gsync[1] is the last day whose E2 work is done.  A wavefront
     * publishes its writes device-wide before it arrives or signals, and
     * observes the others' after its wait. */
    auto gs_arrive = [&](unsigned *at) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (threadIdx.x == 0)
            __hip_atomic_fetch_add(at, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    auto gs_wait = [&](const unsigned *at, unsigned target) {
        if (threadIdx.x == 0)
            while (__hip_atomic_load(at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target)
                __builtin_amdgcn_s_sleep(1);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    };
    auto barrier = [&]() {
        if constexpr (GRID) {
            gs_arrive(gsync);
            gs_wait(gsync, (unsigned)G);
        } else {
            __syncthreads();
        }
    };
    /* Uniform exit decision after a barrier. */
    auto failed = [&](int day, int phase) -> bool {
        if constexpr (GRID)
            return error_by(p, rep, day, phase);
        else
            return __syncthreads_or(myerr) != 0;
    };

    if (active) {
        cx.load(p, node, rep);
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            s.u[c] = p.u0[g * NC + c];
            uw[node * NC + c] = s.u[c];
        });
        if constexpr (ND > 0) {
            static_for<ND>([&](auto C) {
                constexpr int c = decltype(C)::value;
                s.v[c] = p.v0[g * ND + c];
                s.vnew[c] = s.v[c];
            });
        }
        s.rng = siminf_stream_make(p.seed, SIMINF_DOMAIN_NODE, (uint64_t)g);
        int err = 0;
        if constexpr (AEM) {
            /* Initial rates and reaction times (aem.c:117-153). */
            static_for<NT>([&](auto J) {
                constexpr int j = decltype(J)::value;
                const double r = M::template rate<j>(s.u, s.v, cx.ld, cx.gd, tt);
                s.rate[j] = r;
                if (!rate_ok(r))
                    err = SIMINF_ERR_INVALID_RATE;
                const siminf_stream c = siminf_stream_make(
                    p.seed, SIMINF_DOMAIN_AEM, chan_base + (uint64_t)j);
                s.cj[j] = ~0ull;
                double tj = -siminf_log(aem_uniform<M, j>(s, c, 0)) / r + tt;
                if (tj <= 0.0)
                    tj = INFINITY;
                s.time[j] = tj;
                s.inf[j] = 0.0;
                s.j[j] = 1;
            });
        } else {
            s.nb = ~0ull;
            s.a0 = 0.0;
            static_for<NT>([&](auto J) {
                constexpr int j = decltype(J)::value;
                const double r = M::template rate<j>(s.u, s.v, cx.ld, cx.gd, tt);
                s.rate[j] = r;
                s.a0 += r;
                if (!rate_ok(r))
                    err = SIMINF_ERR_INVALID_RATE;
            });
        }
        updw[node] = 0;
        ec.start(p, node);
        /* The first transfer day at or after the run's first plan day. */
        hcur = p.e2_hptr[node];
        hend = p.e2_hptr[node + 1];
        while (hcur < hend && p.e2_hday[hcur] < p.day0)
            hcur++;
        hnext = hcur < hend ? p.e2_hday[hcur] : INT_MAX;
        if (err) {
            record_error(p, rep, 0, PH_INIT, node, err);
            myerr = 1;
        }
    }
    barrier();
    if (failed(0, PH_INIT))
        return;

    for (int day = 0;; day++) {
        /* A: SSA (the whole wavefront takes part) or aem, and E1
         * events. */
        int err = 0;
        if constexpr (AEM) {
            if (active)
                err = aem_day<M>(s, cx, next, chan_base, fired);
        } else {
            err = ssa_day_wave<M>(s, cx, tt, next, fired, active);
        }
        /* The plan day, and whether external transfers touch this herd
         * today: only then is its state published and reloaded. */
        const int pd = day + p.day0;
        const bool touched = active && hnext == pd;
        bool upd1 = false;
        if (active) {
            int bad = 0;
            if (err) {
                record_error(p, rep, day, PH_SSA, node, err);
                myerr = 1;
            } else {
                err = e1_day<NC>(p, tab, ec, tt, s.u, &s.rng, upd1, bad);
                if (err) {
                    record_error(p, rep, day, PH_E1, p.ev_rank[bad], err);
                    myerr = 1;
                }
            }
            if (touched || NB) {
                static_for<NC>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    uw[node * NC + c] = s.u[c];
                });
            }
            if constexpr (NB) {
                static_for<ND>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    vw[node * ND + c] = s.v[c];
                });
            }
        }
        if constexpr (GRID) {
            /* The herd workgroups arrive and wait for the day's E2 work;
             * the E2 workgroup stages the day during their SSA, waits for
             * all of them, runs the levels (whatever errors the day
             * recorded: nothing depends on its results then) and signals.
             * Every workgroup decides on errors after the signal, which
             * orders every error of the day's SSA, E1 and E2 before it. */
            if (e2wg) {
                e2d = e2_stage<NC>(p, rep, day, lds);
                gs_wait(gsync, (unsigned)G + (unsigned)(day + 1) * (unsigned)(G - 1));
                myerr |= e2_run<NC>(p, tab, rep, day, uw, updw, lds, e2d);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                if (threadIdx.x == 0)
                    __hip_atomic_store(gsync + 1, (unsigned)(day + 1), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            } else {
                gs_arrive(gsync);
                gs_wait(gsync + 1, (unsigned)(day + 1));
            }
        } else {
            __syncthreads();
            if (failed(day, PH_E1))
                break;
            /* B: external transfers in the workgroup. */
            e2d = e2_stage<NC>(p, rep, day, lds);
            myerr |= e2_run<NC>(p, tab, rep, day, uw, updw, lds, e2d);
            __syncthreads();
        }
        if (failed(day, PH_E2))
            break;

        /* C: post time step, refresh, writeback. */
        if (active) {
            bool upd = upd1;
            if (touched) {
                static_for<NC>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    s.u[c] = uw[node * NC + c];
                });
                upd = upd || updw[node] != 0;
                updw[node] = 0;
                hcur++;
                hnext = hcur < hend ? p.e2_hday[hcur] : INT_MAX;
            }
            int rc;
            if constexpr (NB) {
                const long long li = p.ld_rep == 2 ? (long long)rep
                                     : (p.ld_rep ? base : 0LL) + node;
                rc = M::pts(s.vnew, uw + node * NC, vw + node * ND,
                            p.ldata + li * (long long)p.Nld, cx.gd, node, tt);
            } else {
                rc = M::pts(s.vnew, s.u, s.v, cx.ld, cx.gd, node, tt);
            }
            if (rc < 0) {
                record_error(p, rep, day, PH_PTS, node, rc);
                myerr = 1;
            } else if (rc > 0 || upd) {
                int err = 0;
                if constexpr (AEM) {
                    /* Through calcTimes at the node's time, which is the
                     * next unit of time (aem.c:329-384). */
                    static_for<NT>([&](auto J) {
                        constexpr int j = decltype(J)::value;
                        const double old = s.rate[j];
                        const double r = M::template rate<j>(s.u, s.vnew, cx.ld, cx.gd, tt);
                        s.rate[j] = r;
                        if (!rate_ok(r))
                            err = SIMINF_ERR_INVALID_RATE;
                        calc_times<M, j>(s, cx.seed, chan_base, next, old, r);
                    });
                } else {
                    err = refresh_rates<M>(s, cx, tt);
                }
                if (err) {
                    record_error(p, rep, day, PH_PTS, node, err);
                    myerr = 1;
                }
            }
        }
        if constexpr (NB) {
            /* Every node's post-time-step function has read the day's
             * published state before any node publishes the next day's
             * (the E2 workgroup of a grid launch writes herd state only
             * after the next day's arrivals, which come after this). */
            if constexpr (GRID) {
                if (!e2wg) {
                    gs_arrive(gsync + 2);
                    gs_wait(gsync + 2, (unsigned)(day + 1) * (unsigned)(G - 1));
                }
            } else {
                __syncthreads();
            }
        }
        tt = next;
        next += 1.0;
        if (active) {
            store_columns<NC, ND>(p, rep, node, tt, U_it, ts, s.u, s.vnew);
            if constexpr (ND > 0) {
                static_for<ND>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    const double x = s.v[c];
                    s.v[c] = s.vnew[c];
                    s.vnew[c] = x;
                });
            }
        } else {
            while (U_it < p.tlen && tt > ts) {
                U_it++;
                ts = U_it < p.tlen ? p.tspan[U_it] : INFINITY;
            }
        }
        if (U_it >= p.tlen)
            break;
    }
    if (GRID && active && p.node_fired && rep == 0)
        p.node_fired[node] = fired;
    add_transitions(p, fired);
}

/* Columns per workgroup of the frozen-node fill. */
#define SIMINF_FILL_COLS 8

/* Non-temporal stores of a node's NC ints (one vector store for 2 or 4
 * compartments): the trajectory is written once and not read back by
 * this run, so the lines need not stay in the caches. */
template <int NC>
__device__ __forceinline__ void store_nt(int *dst, const int (&u)[NC])
{
    typedef int v2i __attribute__((ext_vector_type(2)));
    typedef int v4i __attribute__((ext_vector_type(4)));
    if constexpr (NC == 4) {
        const v4i x = {u[0], u[1], u[2], u[3]};
        __builtin_nontemporal_store(x, (v4i *)dst);
    } else if constexpr (NC == 2) {
        const v2i x = {u[0], u[1]};
        __builtin_nontemporal_store(x, (v2i *)dst);
    } else {
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            __builtin_nontemporal_store(u[c], dst + c);
        });
    }
}

/* The columns frozen nodes left unwritten (node kernels, dense U/V).
 * Workgroup b covers 256 lanes (replicate * Nn + node) of lane block
 * b % nb and the fb columns of column block b / nb (fb = the plan's
 * fill block, SIMINF_FILL_COLS by default), so the
 * workgroups resident together write neighbouring stretches of the same
 * few columns (U column-major, node-major within a column: one
 * wavefront's store is 64 consecutive nodes' counts, contiguous).  NC > 0
 * gives compile-time compartment counts; NC == 0 reads the count from
 * nc.  HASV adds V (nd doubles per node-column). */
template <int NC, bool HASV>
__device__ __forceinline__ void fill_block(const KParams &p, int nc, int nd, int fb,
                                           long long b)
{
    const long long nb = (p.lanes + 255) / 256;
    const long long lb = b % nb, cb = b / nb;
    const long long g = lb * 256 + threadIdx.x;
    if (g >= p.lanes)
        return;
    const int from = p.frozen_from[g];
    const int c0 = (int)(cb * fb);
    const int c1 = min(c0 + fb, p.tlen);
    const int first = max(from, c0);
    if (first >= c1)
        return;
    const int rep = (int)(g / p.Nn);
    const int node = (int)(g - (long long)rep * p.Nn);
    const int Nc = NC > 0 ? NC : nc;
    int u[NC > 0 ? NC : 1];
    if constexpr (NC > 0) {
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            u[c] = p.frozen_u[g * NC + c];
        });
    }
    const long long stride = (long long)p.Nn;
    long long q = ((long long)rep * p.tlen + first) * stride + node;
    for (int it = first; it < c1; it++, q += stride) {
        if constexpr (NC > 0) {
            store_nt<NC>(p.U + q * NC, u);
        } else {
            for (int c = 0; c < Nc; c++)
                __builtin_nontemporal_store(p.frozen_u[g * Nc + c], p.U + q * Nc + c);
        }
        if constexpr (HASV) {
            for (int c = 0; c < nd; c++)
                __builtin_nontemporal_store(p.frozen_v[g * nd + c], p.V + q * nd + c);
        }
    }
}

/* Columns per fill block: the plan's, or the default. */
__host__ __device__ inline int fill_cols_of(const KParams &p)
{
    return p.fill_block > 0 ? p.fill_block : SIMINF_FILL_COLS;
}

/* The fill's work is (lanes / 256) x (tlen / fb) blocks; a launch of at
 * most 2^20 workgroups takes them in order, each workgroup b,
 * b + gridDim.x, ... */
__host__ __device__ inline long long fill_blocks(long long lanes, int tlen, int fb)
{
    return ((lanes + 255) / 256) * ((tlen + fb - 1) / fb);
}

template <int NC, bool HASV>
__global__ __launch_bounds__(256) void k_fill_frozen(KParams p, int nc, int nd)
{
    const int fb = fill_cols_of(p);
    const long long n = fill_blocks(p.lanes, p.tlen, fb);
    for (long long b = blockIdx.x; b < n; b += gridDim.x)
        fill_block<NC, HASV>(p, nc, nd, fb, b);
}

inline unsigned fill_grid(const KParams &p)
{
    return (unsigned)std::min<long long>(fill_blocks(p.lanes, p.tlen, fill_cols_of(p)),
                                         1LL << 20);
}

/* Layout version of SimInf_hip_model, checked when a plugin is used. */
#define SIMINF_HIP_PLUGIN_ABI 3

/* Signature of the KParams layout a plugin was compiled against: the
 * library hands its KParams to the plugin's kernels, so a plugin built
 * from other headers than the library is refused instead of run. */
#define SIMINF_KPARAMS_SIG                                                 \
    ((long long)sizeof(siminf::KParams) * 1000003LL +                      \
     (long long)offsetof(siminf::KParams, e1_rec) * 10007LL +              \
     (long long)offsetof(siminf::KParams, e2_rec) * 1009LL +               \
     (long long)offsetof(siminf::KParams, lds_pre) * 101LL +               \
     (long long)offsetof(siminf::KParams, U) * 11LL +                      \
     (long long)offsetof(siminf::KParams, err_key))

} /* namespace siminf */

struct SimInf_hip_model {
    int abi;
    long long kparams_sig;
    const char *name;
    int Nc, Nt, Nd, Ngdata;
    const int *S;    /* canonical S: S[k*Nc + c] = S(c, k) */
    const int *G;    /* canonical G: G[k*Nt + j] = G(j, k) */
    void (*ssm_nodes)(const siminf::KParams &, bool, bool, unsigned, unsigned,
                      hipStream_t);
    void (*aem_nodes)(const siminf::KParams &, bool, bool, unsigned, unsigned,
                      hipStream_t);
    hipError_t (*coupled)(const siminf::KParams &, bool, bool, unsigned,
                          unsigned, size_t, hipStream_t, int *);
    /* Non-zero when the post-time-step function reads other nodes'
     * state: 1 + the ldata row where each node's (neighbour index,
     * value) pairs start, ended by an index below zero.  Every plan of
     * such a model runs the coupled kernel (ssm_nodes and aem_nodes are
     * null), and the host checks the lists before they reach the
     * device. */
    int neighbors;
};

namespace siminf {

/* Models without E1 events get a kernel without the event code, which
 * needs fewer registers. */
template <class M>
void launch_ssm_nodes(const KParams &p, bool events, bool canon,
                      unsigned grid, unsigned block, hipStream_t st)
{
    if (p.mode == NODES_CLASSIFY) {
        /* 1024-thread workgroups, each lane at most a few nodes. */
        const long long wgs = std::min<long long>((p.lanes + 1023) / 1024, 2048);
        const dim3 cg((unsigned)std::max<long long>(wgs, 1)), cb(1024);
        if (events && canon)
            hipLaunchKernelGGL((k_classify<M, true, true, false>), cg, cb, 0, st, p);
        else if (events)
            hipLaunchKernelGGL((k_classify<M, true, false, false>), cg, cb, 0, st, p);
        else if (canon)
            hipLaunchKernelGGL((k_classify<M, false, true, false>), cg, cb, 0, st, p);
        else
            hipLaunchKernelGGL((k_classify<M, false, false, false>), cg, cb, 0, st, p);
        return;
    }
    {
        /* tspan in LDS up to SIMINF_TS_LDS_MAX points. */
        const bool tsl = p.tlen <= SIMINF_TS_LDS_MAX;
        const size_t lds = tsl ? (size_t)p.tlen * sizeof(double) : 0;
#define SIMINF_FREE(E, C)                                                                \
    do {                                                                                 \
        if (tsl)                                                                         \
            hipLaunchKernelGGL((k_ssm_nodes_free<M, E, C, true>), dim3(grid), dim3(block),  \
                               lds, st, p);                                              \
        else                                                                             \
            hipLaunchKernelGGL((k_ssm_nodes_free<M, E, C, false>), dim3(grid), dim3(block), \
                               0, st, p);                                                \
    } while (0)
        if (events && canon)
            SIMINF_FREE(true, true);
        else if (events)
            SIMINF_FREE(true, false);
        else if (canon)
            SIMINF_FREE(false, true);
        else
            SIMINF_FREE(false, false);
#undef SIMINF_FREE
    }
}

template <class M>
void launch_aem_nodes(const KParams &p, bool events, bool canon,
                      unsigned grid, unsigned block, hipStream_t st)
{
    if (p.mode == NODES_CLASSIFY) {
        /* 1024-thread workgroups, each lane at most a few nodes. */
        const long long wgs = std::min<long long>((p.lanes + 1023) / 1024, 2048);
        const dim3 cg((unsigned)std::max<long long>(wgs, 1)), cb(1024);
        if (events && canon)
            hipLaunchKernelGGL((k_classify<M, true, true, true>), cg, cb, 0, st, p);
        else if (events)
            hipLaunchKernelGGL((k_classify<M, true, false, true>), cg, cb, 0, st, p);
        else if (canon)
            hipLaunchKernelGGL((k_classify<M, false, true, true>), cg, cb, 0, st, p);
        else
            hipLaunchKernelGGL((k_classify<M, false, false, true>), cg, cb, 0, st, p);
        return;
    }
    {
        /* tspan in LDS up to SIMINF_TS_LDS_MAX points. */
        const bool tsl = p.tlen <= SIMINF_TS_LDS_MAX;
        const size_t lds = tsl ? (size_t)p.tlen * sizeof(double) : 0;
#define SIMINF_FREE(E, C)                                                                \
    do {                                                                                 \
        if (tsl)                                                                         \
            hipLaunchKernelGGL((k_aem_nodes_free<M, E, C, true>), dim3(grid), dim3(block),  \
                               lds, st, p);                                              \
        else                                                                             \
            hipLaunchKernelGGL((k_aem_nodes_free<M, E, C, false>), dim3(grid), dim3(block), \
                               0, st, p);                                                \
    } while (0)
        if (events && canon)
            SIMINF_FREE(true, true);
        else if (events)
            SIMINF_FREE(true, false);
        else if (canon)
            SIMINF_FREE(false, true);
        else
            SIMINF_FREE(false, false);
#undef SIMINF_FREE
    }
}

/* Clears a grid launch's barrier words (a kernel, not a memset, so that
 * the launch sequence holds kernels only). */
static __global__ void k_clear_u32(unsigned *w, int n)
{
    for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x)
        w[i] = 0u;
}

/* Block sizes of the coupled kernel: 64 lanes per workgroup when one
 * replicate spans the grid, up to 512 lanes (one per node) when a
 * workgroup holds a replicate. */
#define SIMINF_GRID_BLOCK 64
#define SIMINF_REP_BLOCK 512

template <class M, bool CANON, bool GRID>
static const void *coupled_fn()
{
    return GRID ? (const void *)k_coupled<M, CANON, true, SIMINF_GRID_BLOCK>
                : (const void *)k_coupled<M, CANON, false, SIMINF_REP_BLOCK>;
}

/* The node kernels of a model's record: none for a model whose
 * post-time-step function reads other nodes (it never runs them, and
 * their register-resident state could not serve such reads). */
typedef void (*NodeLaunch)(const KParams &, bool, bool, unsigned, unsigned, hipStream_t);
template <class M, bool NB = model_neighbors<M>::value>
struct NodeLaunchers {
    static constexpr NodeLaunch ssm = launch_ssm_nodes<M>;
    static constexpr NodeLaunch aem = launch_aem_nodes<M>;
};
template <class M>
struct NodeLaunchers<M, true> {
    static constexpr NodeLaunch ssm = nullptr;
    static constexpr NodeLaunch aem = nullptr;
};

/* The ldata row where a neighbour-reading model's pairs start; -1 for
 * other models. */
template <class M>
constexpr int neighbor_row()
{
    if constexpr (model_neighbors<M>::value)
        return M::NEIGHBOR_ROW;
    else
        return -1;
}

/* The all-events method runs one replicate only: grid launches. */
template <class M, bool CANON>
static const void *coupled_aem_fn()
{
    return (const void *)k_coupled<M, CANON, true, SIMINF_GRID_BLOCK, true>;
}

/* Launch the coupled kernel.  With grid, a cooperative launch of `grid`
 * workgroups that must all be resident; max_blocks (out, optional)
 * reports how many can be, for the caller's planning. */
template <class M>
hipError_t launch_coupled(const KParams &p, bool canon, bool grid_mode,
                          unsigned grid, unsigned block, size_t lds,
                          hipStream_t st, int *max_blocks)
{
    if (p.coupled_aem && !grid_mode)
        return hipErrorInvalidValue;
    const void *fn = p.coupled_aem ? (canon ? coupled_aem_fn<M, true>()
                                            : coupled_aem_fn<M, false>())
                     : grid_mode ? (canon ? coupled_fn<M, true, true>()
                                          : coupled_fn<M, false, true>())
                                 : (canon ? coupled_fn<M, true, false>()
                                          : coupled_fn<M, false, false>());
    hipError_t e;
    if (lds > 0) {
        e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
        if (e != hipSuccess)
            return e;
    }
    if (max_blocks) {
        int per_cu = 0, dev = 0, cus = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, (int)block, lds);
        if (e != hipSuccess)
            return e;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        *max_blocks = per_cu * cus;
        return hipSuccess;
    }
    if (grid_mode) {
        if (!p.gsync)
            return hipErrorInvalidValue;
        /* The barrier words of each replicate of the launch. */
        const int nrep = p.rep_blocks > 0 ? (int)grid / p.rep_blocks : 1;
        hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(256), 0, st, p.gsync,
                           SIMINF_GSYNC_WORDS * nrep);
        KParams q = p;
        void *args[] = {&q};
        return hipLaunchCooperativeKernel(fn, dim3(grid), dim3(block), args,
                                          (unsigned)lds, st);
    }
    hipLaunchKernelGGL(((void (*)(KParams))fn), dim3(grid), dim3(block), lds, st, p);
    return hipGetLastError();
}

/* Canonical S and G of the model as flat tables for the host's check. */
template <class M>
struct CanonTables {
    int S[M::NT * M::NC];
    int G[M::NT * M::NT];
    constexpr CanonTables() : S(), G()
    {
        for (int k = 0; k < M::NT; k++) {
            for (int c = 0; c < M::NC; c++)
                S[k * M::NC + c] = M::S[k][c];
            for (int j = 0; j < M::NT; j++)
                G[k * M::NT + j] = M::G[k][j];
        }
    }
};

template <class M>
inline constexpr CanonTables<M> canon_tables{};

#define SIMINF_MODEL_ENTRY(NAME, M, NGDATA)                               \
    {SIMINF_HIP_PLUGIN_ABI, SIMINF_KPARAMS_SIG, NAME, M::NC, M::NT, M::ND, \
     NGDATA,                                                              \
     siminf::canon_tables<M>.S, siminf::canon_tables<M>.G,               \
     siminf::NodeLaunchers<M>::ssm, siminf::NodeLaunchers<M>::aem,        \
     siminf::launch_coupled<M>, siminf::neighbor_row<M>() + 1}

} /* namespace siminf */
