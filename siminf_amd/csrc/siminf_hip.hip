/*
 * siminf_hip.hip -- kernels, device plans and the C ABI of the MI355X
 * solver core (declarations in include/siminf_hip.h).
 *
 * Kernels
 *   k_ssm_nodes<M>  one lane per (replicate, node), all days in one launch:
 *                   ssm/mssm when no external-transfer (E2) event couples
 *                   the nodes (SimInf_solver_ssm.c:74-373,
 *                   SimInf_solver_mssm.c:81-402).
 *   k_aem_nodes<M>  the same for the all-events method
 *                   (SimInf_solver_aem.c:93-459).
 *   k_ssm_reps<M>   one workgroup per replicate, one barrier-separated day
 *                   at a time, for models with E2 events: the replicate's
 *                   compartment state sits in LDS while the day's E2
 *                   events run level by level; the levels are
 *                   conflict-free sets that keep each node's events in
 *                   list order, so no atomics are needed
 *                   (SimInf_solver.c:824-902 run serially in the
 *                   reference, SimInf_solver_ssm.c:244-250).
 */
#include "siminf_device.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace siminf;

/* ------------------------------------------------------------------ */
/* Built-in models as compile-time plugins.                            */
/* ------------------------------------------------------------------ */

#define SIMINF_RATE_ARGS const int *u, const double *v, const double *ld, \
    const double *gd, double t
#define SIMINF_PTS_ARGS double *vn, const int *u, const double *v, \
    const double *ld, const double *gd, int node, double t

struct ModelSIR {
    static constexpr int NC = 3, NT = 2, ND = 0;
    template <int J> __device__ static double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SIR_S_to_I(u, v, ld, gd, t);
        else return SIR_I_to_R(u, v, ld, gd, t);
    }
    __device__ static int pts(SIMINF_PTS_ARGS)
    {
        return siminf_pts_noop(vn, u, v, ld, gd, node, t);
    }
};

struct ModelSEIR {
    static constexpr int NC = 4, NT = 3, ND = 0;
    template <int J> __device__ static double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SEIR_S_to_E(u, v, ld, gd, t);
        else if constexpr (J == 1) return SEIR_E_to_I(u, v, ld, gd, t);
        else return SEIR_I_to_R(u, v, ld, gd, t);
    }
    __device__ static int pts(SIMINF_PTS_ARGS)
    {
        return siminf_pts_noop(vn, u, v, ld, gd, node, t);
    }
};

struct ModelSIS {
    static constexpr int NC = 2, NT = 2, ND = 0;
    template <int J> __device__ static double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SIS_S_to_I(u, v, ld, gd, t);
        else return SIS_I_to_S(u, v, ld, gd, t);
    }
    __device__ static int pts(SIMINF_PTS_ARGS)
    {
        return siminf_pts_noop(vn, u, v, ld, gd, node, t);
    }
};

struct ModelSISe {
    static constexpr int NC = 2, NT = 2, ND = 1;
    template <int J> __device__ static double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SISe_S_to_I(u, v, ld, gd, t);
        else return SISe_I_to_S(u, v, ld, gd, t);
    }
    __device__ static int pts(SIMINF_PTS_ARGS)
    {
        return SISe_pts(vn, u, v, ld, gd, node, t);
    }
};

struct ModelSISe3 {
    static constexpr int NC = 6, NT = 6, ND = 1;
    template <int J> __device__ static double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SISe3_S_1_to_I_1(u, v, ld, gd, t);
        else if constexpr (J == 1) return SISe3_I_1_to_S_1(u, v, ld, gd, t);
        else if constexpr (J == 2) return SISe3_S_2_to_I_2(u, v, ld, gd, t);
        else if constexpr (J == 3) return SISe3_I_2_to_S_2(u, v, ld, gd, t);
        else if constexpr (J == 4) return SISe3_S_3_to_I_3(u, v, ld, gd, t);
        else return SISe3_I_3_to_S_3(u, v, ld, gd, t);
    }
    __device__ static int pts(SIMINF_PTS_ARGS)
    {
        return SISe3_pts(vn, u, v, ld, gd, node, t);
    }
};

struct ModelSIRmparse {
    static constexpr int NC = 3, NT = 2, ND = 0;
    template <int J> __device__ static double rate(SIMINF_RATE_ARGS)
    {
        if constexpr (J == 0) return SIR_mparse_trFun1(u, v, ld, gd, t);
        else return SIR_mparse_trFun2(u, v, ld, gd, t);
    }
    __device__ static int pts(SIMINF_PTS_ARGS)
    {
        return siminf_pts_noop(vn, u, v, ld, gd, node, t);
    }
};

/* ------------------------------------------------------------------ */
/* Kernels.                                                            */
/* ------------------------------------------------------------------ */

/* E1 events of one node due at tt, in list order, drawing from the
 * node's stream.  Returns 0 or an error code; *bad gets the failing
 * event. */
template <int NC>
__device__ __forceinline__ int e1_day(const KParams &p, int &cur, int end,
                                      double tt, int (&u)[NC],
                                      siminf_stream *rng, bool &upd, int &bad)
{
    while (cur < end) {
        const int k = p.e1_idx[cur];
        if (p.ev_time[k] > tt)
            break;
        const int err = apply_e1<NC>(p, k, u, rng);
        if (err) {
            bad = k;
            return err;
        }
        upd = true;
        cur++;
    }
    return 0;
}

template <class M, bool EVENTS>
__global__ __launch_bounds__(256) void k_ssm_nodes(KParams p)
{
    constexpr int NC = M::NC, NT = M::NT, ND = M::ND;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long fired = 0;

    if (g < p.lanes) {
        const int rep = (int)(g / p.Nn);
        const int node = (int)(g - (long long)rep * p.Nn);
        const double *ld = p.ldata + (long long)node * p.Nld;
        NodeState<M> s;
        int err = 0;
        double tt = p.tt0, next = p.next0;

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

        /* Initial rates (ssm.c:96-117). */
        s.a0 = 0.0;
        static_for<NT>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const double r = M::template rate<j>(s.u, s.v, ld, p.gdata, tt);
            s.rate[j] = r;
            s.a0 += r;
            if (!rate_ok(r))
                err = SIMINF_ERR_INVALID_RATE;
        });

        if (err) {
            record_error(p, rep, 0, PH_INIT, node, err);
        } else {
            int cur = p.e1_ptr[node];
            const int end = p.e1_ptr[node + 1];
            int U_it = 0, V_it = 0, day = 0;
            for (;;) {
                bool upd = false;
                int bad = 0;

                err = ssa_day<M>(s, p, ld, tt, next, fired);
                if (err) {
                    record_error(p, rep, day, PH_SSA, node, err);
                    break;
                }
                if constexpr (EVENTS) {
                    err = e1_day<NC>(p, cur, end, tt, s.u, &s.rng, upd, bad);
                    if (err) {
                        record_error(p, rep, day, PH_E1, p.ev_rank[bad], err);
                        break;
                    }
                }

                /* Post time step and refresh (ssm.c:262-302). */
                const int rc = M::pts(s.vnew, s.u, s.v, ld, p.gdata, node, tt);
                if (rc < 0) {
                    record_error(p, rep, day, PH_PTS, node, rc);
                    break;
                }
                if (rc > 0 || upd) {
                    err = refresh_rates<M>(s, p, ld, tt);
                    if (err) {
                        record_error(p, rep, day, PH_PTS, node, err);
                        break;
                    }
                }

                tt = next;
                next += 1.0;
                store_columns<NC, ND>(p, rep, node, tt, U_it, V_it, s.u, s.vnew);
                if constexpr (ND > 0) {
                    static_for<ND>([&](auto C) {
                        constexpr int c = decltype(C)::value;
                        const double x = s.v[c];
                        s.v[c] = s.vnew[c];
                        s.vnew[c] = x;
                    });
                }
                if (U_it >= p.tlen)
                    break;
                day++;
            }
        }
    }
    add_transitions(p, fired);
}

template <class M, bool EVENTS>
__global__ __launch_bounds__(256) void k_aem_nodes(KParams p)
{
    constexpr int NC = M::NC, NT = M::NT, ND = M::ND;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long fired = 0;

    if (g < p.lanes) {
        const int node = (int)g;
        const double *ld = p.ldata + (long long)node * p.Nld;
        const uint64_t chan_base = (uint64_t)node * (uint64_t)NT;
        AemState<M> s;
        int err = 0;
        double tt = p.tt0, next = p.next0;

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
            const double r = M::template rate<j>(s.u, s.v, ld, p.gdata, tt);
            s.rate[j] = r;
            if (!rate_ok(r))
                err = SIMINF_ERR_INVALID_RATE;
            const siminf_stream c = siminf_stream_make(
                p.seed, SIMINF_DOMAIN_AEM, chan_base + (uint64_t)j);
            double tj = -siminf_log(siminf_uniform_at(&c, 0)) / r + tt;
            if (tj <= 0.0)
                tj = INFINITY;
            s.time[j] = tj;
            s.inf[j] = 0.0;
            s.j[j] = 1;
        });

        if (err) {
            record_error(p, 0, 0, PH_INIT, node, err);
        } else {
            int cur = p.e1_ptr[node];
            const int end = p.e1_ptr[node + 1];
            int U_it = 0, V_it = 0, day = 0;
            for (;;) {
                bool upd = false;
                int bad = 0;

                err = aem_day<M>(s, p, ld, next, chan_base, fired);
                if (err) {
                    record_error(p, 0, day, PH_SSA, node, err);
                    break;
                }
                if constexpr (EVENTS) {
                    err = e1_day<NC>(p, cur, end, tt, s.u, &s.rng, upd, bad);
                    if (err) {
                        record_error(p, 0, day, PH_E1, p.ev_rank[bad], err);
                        break;
                    }
                }

                /* Post time step; refresh through calcTimes at the node's
                 * time, which is the next unit of time (aem.c:329-384). */
                const int rc = M::pts(s.vnew, s.u, s.v, ld, p.gdata, node, tt);
                if (rc < 0) {
                    record_error(p, 0, day, PH_PTS, node, rc);
                    break;
                }
                if (rc > 0 || upd) {
                    static_for<NT>([&](auto J) {
                        constexpr int j = decltype(J)::value;
                        const double old = s.rate[j];
                        const double r = M::template rate<j>(s.u, s.vnew, ld,
                                                             p.gdata, tt);
                        s.rate[j] = r;
                        if (!rate_ok(r))
                            err = SIMINF_ERR_INVALID_RATE;
                        calc_times<M, j>(s, p, chan_base, next, old, r);
                    });
                    if (err) {
                        record_error(p, 0, day, PH_PTS, node, err);
                        break;
                    }
                }

                tt = next;
                next += 1.0;
                store_columns<NC, ND>(p, 0, node, tt, U_it, V_it, s.u, s.vnew);
                if constexpr (ND > 0) {
                    static_for<ND>([&](auto C) {
                        constexpr int c = decltype(C)::value;
                        const double x = s.v[c];
                        s.v[c] = s.vnew[c];
                        s.vnew[c] = x;
                    });
                }
                if (U_it >= p.tlen)
                    break;
                day++;
            }
        }
    }
    add_transitions(p, fired);
}

/* Replicate-per-workgroup ssm/mssm with external transfers.  Per day:
 *   A: each thread runs SSA + E1 for its nodes (state in registers while
 *      a node is processed, kept in global work arrays between days; u in
 *      LDS); barrier;
 *   B: the day's E2 events, one conflict-free level at a time, on the LDS
 *      copy of u; barrier after each level;
 *   C: post time step, refresh, writeback; merged with the next day's A
 *      for the same node, since both only touch that node. */
template <class M>
__global__ __launch_bounds__(1024) void k_ssm_reps(KParams p, int u_in_lds)
{
    constexpr int NC = M::NC, NT = M::NT, ND = M::ND;
    extern __shared__ int lds[];

    const int rep = blockIdx.x;
    const int Nn = p.Nn;
    const long long base = (long long)rep * Nn;
    int *uw = u_in_lds ? lds : p.u_work + base * NC;
    int *upd = u_in_lds ? lds + (long long)Nn * NC : p.e1_cursor + (long long)p.lanes + base;
    unsigned long long fired = 0;
    double tt = p.tt0, next = p.next0;
    int U_it = 0, V_it = 0;
    /* Errors are kept per thread and combined with __syncthreads_or at
     * the end of each phase, so every thread leaves the day loop at the
     * same barrier. */
    int myerr = 0;

    /* Initialisation. */
    for (int node = threadIdx.x; node < Nn; node += blockDim.x) {
        const long long g = base + node;
        const double *ld = p.ldata + (long long)node * p.Nld;
        NodeState<M> s;
        int err = 0;
        static_for<NC>([&](auto C) {
            constexpr int c = decltype(C)::value;
            s.u[c] = p.u0[g * NC + c];
            uw[node * NC + c] = s.u[c];
        });
        if constexpr (ND > 0) {
            static_for<ND>([&](auto C) {
                constexpr int c = decltype(C)::value;
                s.v[c] = p.v0[g * ND + c];
                p.v_work[g * ND + c] = s.v[c];
                p.vnew_work[g * ND + c] = s.v[c];
            });
        }
        s.a0 = 0.0;
        static_for<NT>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const double r = M::template rate<j>(s.u, s.v, ld, p.gdata, tt);
            p.rate_work[(long long)j * p.lanes + g] = r;
            s.a0 += r;
            if (!rate_ok(r))
                err = SIMINF_ERR_INVALID_RATE;
        });
        p.a0_work[g] = s.a0;
        p.rng_j[g] = 0;
        p.rng_cache[g] = 0.0;
        p.e1_cursor[g] = p.e1_ptr[node];
        upd[node] = 0;
        if (err) {
            record_error(p, rep, 0, PH_INIT, node, err);
            myerr = 1;
        }
    }
    if (__syncthreads_or(myerr))
        return;

    for (int day = 0;; day++) {
        /* The buffers holding v and v_new swap every day (ssm.c:350-353). */
        double *vbuf = (day & 1) ? p.vnew_work : p.v_work;
        double *vnbuf = (day & 1) ? p.v_work : p.vnew_work;

        /* Phase A: SSA and E1 events of this thread's nodes. */
        for (int node = threadIdx.x; node < Nn; node += blockDim.x) {
            const long long g = base + node;
            const double *ld = p.ldata + (long long)node * p.Nld;
            NodeState<M> s;
            static_for<NC>([&](auto C) {
                constexpr int c = decltype(C)::value;
                s.u[c] = uw[node * NC + c];
            });
            if constexpr (ND > 0) {
                static_for<ND>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    s.v[c] = vbuf[g * ND + c];
                });
            }
            static_for<NT>([&](auto J) {
                constexpr int j = decltype(J)::value;
                s.rate[j] = p.rate_work[(long long)j * p.lanes + g];
            });
            s.a0 = p.a0_work[g];
            s.rng = siminf_stream_make(p.seed, SIMINF_DOMAIN_NODE, (uint64_t)g);
            s.rng.j = p.rng_j[g];
            s.rng.cache = p.rng_cache[g];

            int err = ssa_day<M>(s, p, ld, tt, next, fired);
            if (err) {
                record_error(p, rep, day, PH_SSA, node, err);
                myerr = 1;
            } else {
                int cur = p.e1_cursor[g], bad = 0;
                bool e1 = false;
                err = e1_day<NC>(p, cur, p.e1_ptr[node + 1], tt, s.u, &s.rng,
                                 e1, bad);
                if (err) {
                    record_error(p, rep, day, PH_E1, p.ev_rank[bad], err);
                    myerr = 1;
                }
                p.e1_cursor[g] = cur;
                if (e1)
                    upd[node] = 1;
            }
            static_for<NC>([&](auto C) {
                constexpr int c = decltype(C)::value;
                uw[node * NC + c] = s.u[c];
            });
            static_for<NT>([&](auto J) {
                constexpr int j = decltype(J)::value;
                p.rate_work[(long long)j * p.lanes + g] = s.rate[j];
            });
            p.a0_work[g] = s.a0;
            p.rng_j[g] = s.rng.j;
            p.rng_cache[g] = s.rng.cache;
        }
        if (__syncthreads_or(myerr))
            break;

        /* Phase B: external transfers, level by level. */
        if (day < p.ndays) {
            for (int l = p.e2_day[day]; l < p.e2_day[day + 1]; l++) {
                for (int i = p.e2_lvl[l] + threadIdx.x; i < p.e2_lvl[l + 1];
                     i += blockDim.x) {
                    const int k = p.e2_idx[i];
                    siminf_stream rng = siminf_stream_make(
                        p.seed, SIMINF_DOMAIN_E2,
                        (uint64_t)rep * (uint64_t)p.len + (uint64_t)k);
                    const int err = apply_e2<NC>(p, k, uw, upd, &rng);
                    if (err) {
                        record_error(p, rep, day, PH_E2, p.ev_rank[k], err);
                        myerr = 1;
                    }
                }
                /* Later levels still run after a failure: an event that
                 * fails there can only be later in list order than the
                 * first failure of its own chain, so the smallest key is
                 * still the serial run's first failure. */
                __syncthreads();
            }
        }
        if (__syncthreads_or(myerr))
            break;

        /* Phase C: post time step, refresh, writeback. */
        const double tt_next = next;
        for (int node = threadIdx.x; node < Nn; node += blockDim.x) {
            const long long g = base + node;
            const double *ld = p.ldata + (long long)node * p.Nld;
            NodeState<M> s;
            int err = 0;
            static_for<NC>([&](auto C) {
                constexpr int c = decltype(C)::value;
                s.u[c] = uw[node * NC + c];
            });
            if constexpr (ND > 0) {
                static_for<ND>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    s.v[c] = vbuf[g * ND + c];
                    s.vnew[c] = vnbuf[g * ND + c];
                });
            }
            const int rc = M::pts(s.vnew, s.u, s.v, ld, p.gdata, node, tt);
            if (rc < 0) {
                record_error(p, rep, day, PH_PTS, node, rc);
                myerr = 1;
            } else if (rc > 0 || upd[node]) {
                static_for<NT>([&](auto J) {
                    constexpr int j = decltype(J)::value;
                    s.rate[j] = p.rate_work[(long long)j * p.lanes + g];
                });
                s.a0 = p.a0_work[g];
                err = refresh_rates<M>(s, p, ld, tt);
                if (err) {
                    record_error(p, rep, day, PH_PTS, node, err);
                    myerr = 1;
                }
                static_for<NT>([&](auto J) {
                    constexpr int j = decltype(J)::value;
                    p.rate_work[(long long)j * p.lanes + g] = s.rate[j];
                });
                p.a0_work[g] = s.a0;
                upd[node] = 0;
            }
            if constexpr (ND > 0) {
                static_for<ND>([&](auto C) {
                    constexpr int c = decltype(C)::value;
                    vnbuf[g * ND + c] = s.vnew[c];
                });
            }
            int Ui = U_it, Vi = V_it;
            store_columns<NC, ND>(p, rep, node, tt_next, Ui, Vi, s.u, s.vnew);
        }

        tt = next;
        next += 1.0;
        while (U_it < p.tlen && tt > p.tspan[U_it])
            U_it++;
        while (V_it < p.tlen && tt > p.tspan[V_it])
            V_it++;
        if (__syncthreads_or(myerr) || U_it >= p.tlen)
            break;
    }
    add_transitions(p, fired);
}

/* ------------------------------------------------------------------ */
/* Model registry.                                                     */
/* ------------------------------------------------------------------ */

struct SimInf_hip_model {
    const char *name;
    int Nc, Nt, Nd, Ngdata;
    void (*ssm_nodes)(const KParams &, bool, unsigned, unsigned, hipStream_t);
    void (*aem_nodes)(const KParams &, bool, unsigned, unsigned, hipStream_t);
    hipError_t (*ssm_reps)(const KParams &, int, unsigned, unsigned, size_t,
                           hipStream_t);
};

/* Models without E1 events get a kernel without the event code, which
 * needs fewer registers. */
template <class M>
static void launch_ssm_nodes(const KParams &p, bool events, unsigned grid,
                             unsigned block, hipStream_t st)
{
    if (events)
        hipLaunchKernelGGL((k_ssm_nodes<M, true>), dim3(grid), dim3(block), 0, st, p);
    else
        hipLaunchKernelGGL((k_ssm_nodes<M, false>), dim3(grid), dim3(block), 0, st, p);
}

template <class M>
static void launch_aem_nodes(const KParams &p, bool events, unsigned grid,
                             unsigned block, hipStream_t st)
{
    if (events)
        hipLaunchKernelGGL((k_aem_nodes<M, true>), dim3(grid), dim3(block), 0, st, p);
    else
        hipLaunchKernelGGL((k_aem_nodes<M, false>), dim3(grid), dim3(block), 0, st, p);
}

template <class M>
static hipError_t launch_ssm_reps(const KParams &p, int u_in_lds, unsigned grid,
                                  unsigned block, size_t lds, hipStream_t st)
{
    if (lds > 0) {
        const hipError_t e = hipFuncSetAttribute(
            (const void *)k_ssm_reps<M>,
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess)
            return e;
    }
    hipLaunchKernelGGL(k_ssm_reps<M>, dim3(grid), dim3(block), lds, st, p,
                       u_in_lds);
    return hipGetLastError();
}

#define SIMINF_MODEL_ENTRY(NAME, M, NGDATA)                               \
    {NAME, M::NC, M::NT, M::ND, NGDATA, launch_ssm_nodes<M>,              \
     launch_aem_nodes<M>, launch_ssm_reps<M>}

/* gdata lengths: SISe.R:177-178 (8 values), SISe3.R:188-190 (12), the
 * mparse SIR (beta, gamma); SIR, SEIR and SIS read ldata only. */
static const SimInf_hip_model builtin_models[] = {
    SIMINF_MODEL_ENTRY("SIR", ModelSIR, 0),
    SIMINF_MODEL_ENTRY("SEIR", ModelSEIR, 0),
    SIMINF_MODEL_ENTRY("SIS", ModelSIS, 0),
    SIMINF_MODEL_ENTRY("SISe", ModelSISe, 8),
    SIMINF_MODEL_ENTRY("SISe3", ModelSISe3, 12),
    SIMINF_MODEL_ENTRY("SIR_mparse", ModelSIRmparse, 2),
};

extern "C" const SimInf_hip_model *SimInf_hip_model_builtin(const char *name)
{
    if (!name)
        return nullptr;
    for (const SimInf_hip_model &m : builtin_models)
        if (std::strcmp(m.name, name) == 0)
            return &m;
    return nullptr;
}

extern "C" int SimInf_hip_model_dims(const SimInf_hip_model *model, int *Nc,
                                     int *Nt, int *Nd)
{
    if (!model)
        return SIMINF_ERR_INVALID_MODEL;
    if (Nc) *Nc = model->Nc;
    if (Nt) *Nt = model->Nt;
    if (Nd) *Nd = model->Nd;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Plans.                                                              */
/* ------------------------------------------------------------------ */

namespace {

enum Kernel { K_SSM_NODES, K_AEM_NODES, K_SSM_REPS };

struct DeviceBuffers {
    std::vector<void *> ptrs;
    hipError_t status = hipSuccess;

    template <class T>
    T *alloc(size_t count)
    {
        void *p = nullptr;
        if (status != hipSuccess)
            return nullptr;
        const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
        status = hipMalloc(&p, bytes);
        if (status != hipSuccess)
            return nullptr;
        ptrs.push_back(p);
        return (T *)p;
    }

    template <class T>
    T *upload(const T *host, size_t count)
    {
        T *d = alloc<T>(count);
        if (d && count && host)
            status = hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice);
        return d;
    }

    ~DeviceBuffers()
    {
        for (void *p : ptrs)
            (void)hipFree(p);
    }
};

} // namespace

struct SimInf_hip_plan {
    const SimInf_hip_model *model;
    Kernel kernel;
    KParams p;
    DeviceBuffers buf;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    unsigned grid = 0, block = 0;
    size_t lds = 0;
    int u_in_lds = 0;
    bool has_e1 = false;

    /* Host copies needed after runs. */
    int Nn, Nc, Nd, Nrep, tlen;
    bool dense_U, dense_V;
    size_t nnzU, nnzV;
    int ndays;
    uint64_t host_err_key;     /* errors found while validating events */

    /* Last run. */
    unsigned long long transitions = 0;
    double kernel_ms = 0.0;
};

static const uint64_t NO_ERROR_KEY = ~0ull;

static int key_to_code(uint64_t key)
{
    return key == NO_ERROR_KEY ? 0 : -(int)(key & 63);
}

/* Number of days the day loop runs: it stops once every tspan point is
 * below tt (ssm.c:318-361). */
static int count_days(const double *tspan, int tlen)
{
    if (tlen <= 0)
        return 1;
    const double f0 = std::floor(tspan[0]);
    long long k = 0;
    while (!(f0 + (double)(k + 1) > tspan[tlen - 1]))
        k++;
    return (int)(k + 1);
}

/* Day index on which an event with integer time tm is processed. */
static int event_day(int tm, double tt0)
{
    if ((double)tm <= tt0)
        return 0;
    return (int)((double)tm - std::floor(tt0));
}

extern "C" int SimInf_hip_plan_create(SimInf_hip_plan **out,
                                      const SimInf_solver_args *a,
                                      const SimInf_hip_model *model)
{
    if (!out || !a || !model)
        return SIMINF_ERR_INVALID_MODEL;
    *out = nullptr;

    const int Nn = a->Nn, Nc = a->Nc, Nt = a->Nt, Nd = a->Nd;
    const int Nrep = a->Nrep < 1 ? 1 : a->Nrep;
    if (Nn <= 0 || Nc != model->Nc || Nt != model->Nt || Nd != model->Nd ||
        a->tlen <= 0 || a->Nld < 0 || a->len < 0)
        return SIMINF_ERR_INVALID_MODEL;
    if (a->solver != 0 && a->solver != 1)
        return SIMINF_ERR_UNKNOWN_SOLVER;
    if (a->solver == 1 && Nrep > 1)
        return SIMINF_ERR_AEM_REPLICATED_MODEL;
    if (!a->U && (!a->irU || !a->jcU || !a->prU))
        return SIMINF_ERR_SPARSE_MODEL;
    if (!a->V && (!a->irV || !a->jcV || !a->prV))
        return SIMINF_ERR_SPARSE_MODEL;

    SimInf_hip_plan *P = new SimInf_hip_plan();
    P->model = model;
    P->Nn = Nn;
    P->Nc = Nc;
    P->Nd = Nd;
    P->Nrep = Nrep;
    P->tlen = a->tlen;
    P->dense_U = a->U != nullptr;
    P->dense_V = a->V != nullptr;
    P->host_err_key = NO_ERROR_KEY;

    /* S, G as bit masks and a dense value table. */
    std::vector<uint32_t> Smask(Nt, 0), Gmask(Nt, 0);
    std::vector<int> Sval((size_t)Nt * Nc, 0);
    for (int k = 0; k < Nt; k++) {
        for (int j = a->jcS[k]; j < a->jcS[k + 1]; j++) {
            const int c = a->irS[j];
            if (c < 0 || c >= Nc) { delete P; return SIMINF_ERR_INVALID_MODEL; }
            Smask[k] |= 1u << c;
            Sval[(size_t)k * Nc + c] = a->prS[j];
        }
        for (int j = a->jcG[k]; j < a->jcG[k + 1]; j++) {
            const int r = a->irG[j];
            if (r < 0 || r >= Nt) { delete P; return SIMINF_ERR_INVALID_MODEL; }
            Gmask[k] |= 1u << r;
        }
    }

    /* E as masks and dense weights; the number of columns is the largest
     * select that jcE describes, found from the events. */
    int Nselect = 0;
    for (int k = 0; k < a->len; k++)
        Nselect = std::max(Nselect, a->select[k]);
    std::vector<uint32_t> Emask(std::max(Nselect, 1), 0);
    std::vector<double> Ew((size_t)std::max(Nselect, 1) * Nc, 0.0);
    for (int s = 0; s < Nselect; s++) {
        for (int j = a->jcE[s]; j < a->jcE[s + 1]; j++) {
            const int c = a->irE[j];
            if (c < 0 || c >= Nc) { delete P; return SIMINF_ERR_INVALID_MODEL; }
            Emask[s] |= 1u << c;
            Ew[(size_t)s * Nc + c] = a->prE[j];
        }
    }
    int Nshift = 0;
    for (int k = 0; k < a->len; k++)
        Nshift = std::max(Nshift, a->shift[k]);

    /* Events: zero-based copies, day and position within the day, E1
     * lists per node and the E2 level schedule. */
    const double tt0 = a->tspan[0];
    const int ndays = count_days(a->tspan, a->tlen);
    P->ndays = ndays;
    const int len = a->len;
    std::vector<int> ev_node(len), ev_dest(len), ev_select(len), ev_shift(len),
        ev_rank(len), ev_day(len);
    std::vector<int> e1_count(Nn + 1, 0);
    int prev_time = 0, day_start = 0, prev_day = -1;
    for (int k = 0; k < len; k++) {
        if (k > 0 && a->time[k] < prev_time) { delete P; return SIMINF_ERR_INVALID_MODEL; }
        prev_time = a->time[k];
        if (a->select[k] < 1) { delete P; return SIMINF_ERR_INVALID_MODEL; }
        ev_node[k] = a->node[k] - 1;
        ev_dest[k] = a->dest[k] - 1;
        ev_select[k] = a->select[k] - 1;
        ev_shift[k] = a->shift[k] - 1;
        ev_day[k] = event_day(a->time[k], tt0);
        if (ev_day[k] != prev_day) {
            day_start = k;
            prev_day = ev_day[k];
        }
        ev_rank[k] = k - day_start;
        /* The reference checks node bounds first for every event
         * (SimInf_solver.c:674-678). */
        if (ev_node[k] < 0 || ev_node[k] >= Nn) {
            if (ev_day[k] < ndays) {
                const int ph = a->event[k] == SIMINF_EXTERNAL_TRANSFER_EVENT ? PH_E2 : PH_E1;
                P->host_err_key = std::min<uint64_t>(
                    P->host_err_key,
                    error_key(0, ev_day[k], ph, ev_rank[k], SIMINF_ERR_NODE_OUT_OF_BOUNDS));
            }
            continue;
        }
        if (a->event[k] != SIMINF_EXTERNAL_TRANSFER_EVENT)
            e1_count[ev_node[k] + 1]++;
    }
    std::vector<int> e1_ptr(Nn + 1, 0), e1_idx;
    for (int i = 0; i < Nn; i++)
        e1_ptr[i + 1] = e1_ptr[i] + e1_count[i + 1];
    P->has_e1 = e1_ptr[Nn] > 0;
    e1_idx.resize(std::max(e1_ptr[Nn], 1));
    {
        std::vector<int> fill(e1_ptr.begin(), e1_ptr.end() - 1);
        for (int k = 0; k < len; k++) {
            if (ev_node[k] < 0 || ev_node[k] >= Nn ||
                a->event[k] == SIMINF_EXTERNAL_TRANSFER_EVENT)
                continue;
            e1_idx[fill[ev_node[k]]++] = k;
        }
    }

    /* E2 levels: an event's level is one more than the last level that
     * touched its source or destination on the same day. */
    std::vector<int> e2_day(ndays + 1, 0), e2_lvl(1, 0), e2_idx;
    int n_e2 = 0;
    {
        std::vector<int> last_level(Nn, -1), last_day(Nn, -1);
        std::vector<std::vector<int>> levels;
        int k = 0;
        for (int d = 0; d < ndays; d++) {
            levels.clear();
            while (k < len && ev_day[k] < d)
                k++;
            for (; k < len && ev_day[k] == d; k++) {
                if (a->event[k] != SIMINF_EXTERNAL_TRANSFER_EVENT)
                    continue;
                const int src = ev_node[k], dst = ev_dest[k];
                if (src < 0 || src >= Nn)
                    continue;
                int lvl = 0;
                if (last_day[src] == d)
                    lvl = std::max(lvl, last_level[src] + 1);
                const bool dst_ok = dst >= 0 && dst < Nn;
                if (dst_ok && last_day[dst] == d)
                    lvl = std::max(lvl, last_level[dst] + 1);
                last_day[src] = d;
                last_level[src] = lvl;
                if (dst_ok) {
                    last_day[dst] = d;
                    last_level[dst] = lvl;
                }
                if ((int)levels.size() <= lvl)
                    levels.resize(lvl + 1);
                levels[lvl].push_back(k);
                n_e2++;
            }
            for (auto &lv : levels) {
                e2_idx.insert(e2_idx.end(), lv.begin(), lv.end());
                e2_lvl.push_back((int)e2_idx.size());
            }
            e2_day[d + 1] = (int)e2_lvl.size() - 1;
        }
        if (e2_idx.empty())
            e2_idx.push_back(0);
    }

    /* Kernel choice. */
    if (a->solver == 1) {
        if (n_e2 > 0) {
            /* aem with external transfers is not implemented on the GPU. */
            delete P;
            return SIMINF_ERR_UNKNOWN_SOLVER;
        }
        P->kernel = K_AEM_NODES;
    } else {
        P->kernel = n_e2 > 0 ? K_SSM_REPS : K_SSM_NODES;
    }

    /* Device buffers. */
    DeviceBuffers &B = P->buf;
    KParams &p = P->p;
    memset(&p, 0, sizeof(p));
    const long long lanes = (long long)Nn * Nrep;
    p.Nn = Nn;
    p.Nrep = Nrep;
    p.Nld = a->Nld;
    p.tlen = a->tlen;
    p.len = len;
    p.Nshift = Nshift;
    p.lanes = lanes;
    p.tt0 = tt0;
    p.next0 = std::floor(tt0) + 1.0;
    p.ndays = ndays;

    p.u0 = B.upload(a->u0, (size_t)lanes * Nc);
    p.v0 = B.upload(a->v0, (size_t)lanes * Nd);
    if (a->Nld > 0) {
        p.ldata = B.upload(a->ldata, (size_t)a->Nld * Nn);
    } else {
        const double inf1 = INFINITY;
        p.ldata = B.upload(&inf1, 1);
    }
    {
        /* SimInf_solver_args carries no gdata length; the plugin states
         * how many values its functions read. */
        std::vector<double> gd(std::max(model->Ngdata, 1), 0.0);
        if (a->gdata)
            std::copy(a->gdata, a->gdata + model->Ngdata, gd.begin());
        p.gdata = B.upload(gd.data(), gd.size());
    }
    p.tspan = B.upload(a->tspan, (size_t)a->tlen);
    p.Smask = B.upload(Smask.data(), Smask.size());
    p.Sval = B.upload(Sval.data(), Sval.size());
    p.Gmask = B.upload(Gmask.data(), Gmask.size());
    p.Emask = B.upload(Emask.data(), Emask.size());
    p.Ew = B.upload(Ew.data(), Ew.size());
    p.Nmat = a->N ? B.upload(a->N, (size_t)Nc * std::max(Nshift, 1)) : nullptr;

    p.ev_type = B.upload(a->event, (size_t)len);
    p.ev_time = B.upload(a->time, (size_t)len);
    p.ev_node = B.upload(ev_node.data(), (size_t)len);
    p.ev_dest = B.upload(ev_dest.data(), (size_t)len);
    p.ev_n = B.upload(a->n, (size_t)len);
    p.ev_select = B.upload(ev_select.data(), (size_t)len);
    p.ev_shift = B.upload(ev_shift.data(), (size_t)len);
    p.ev_rank = B.upload(ev_rank.data(), (size_t)len);
    p.ev_day = B.upload(ev_day.data(), (size_t)len);
    p.ev_prop = B.upload(a->proportion, (size_t)len);
    p.e1_ptr = B.upload(e1_ptr.data(), e1_ptr.size());
    p.e1_idx = B.upload(e1_idx.data(), e1_idx.size());
    p.e2_day = B.upload(e2_day.data(), e2_day.size());
    p.e2_lvl = B.upload(e2_lvl.data(), e2_lvl.size());
    p.e2_idx = B.upload(e2_idx.data(), e2_idx.size());

    const size_t ncols = (size_t)Nrep * a->tlen;
    if (a->U) {
        p.U = B.alloc<int>((size_t)Nn * Nc * ncols);
    } else {
        P->nnzU = (size_t)a->jcU[ncols];
        p.irU = B.upload(a->irU, P->nnzU);
        p.jcU = B.upload(a->jcU, ncols + 1);
        p.prU = B.alloc<double>(P->nnzU);
    }
    if (a->V) {
        p.V = B.alloc<double>((size_t)Nn * Nd * ncols);
    } else {
        P->nnzV = (size_t)a->jcV[ncols];
        p.irV = B.upload(a->irV, P->nnzV);
        p.jcV = B.upload(a->jcV, ncols + 1);
        p.prV = B.alloc<double>(P->nnzV);
    }
    p.err_key = B.alloc<unsigned long long>(1);
    p.transitions = B.alloc<unsigned long long>(1);

    if (P->kernel == K_SSM_REPS) {
        p.v_work = B.alloc<double>((size_t)lanes * Nd);
        p.vnew_work = B.alloc<double>((size_t)lanes * Nd);
        p.rate_work = B.alloc<double>((size_t)lanes * Nt);
        p.a0_work = B.alloc<double>((size_t)lanes);
        p.rng_j = B.alloc<unsigned long long>((size_t)lanes);
        p.rng_cache = B.alloc<double>((size_t)lanes);
        /* e1 cursors, then the update flags when u is not in LDS. */
        p.e1_cursor = B.alloc<int>((size_t)lanes * 2);
        int dev = 0, max_lds = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&max_lds,
                                    hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        const size_t need = (size_t)Nn * (Nc + 1) * sizeof(int);
        if (need + 1024 <= (size_t)max_lds) {
            P->u_in_lds = 1;
            P->lds = need;
        } else {
            P->u_in_lds = 0;
            P->lds = 0;
            p.u_work = B.alloc<int>((size_t)lanes * Nc);
        }
        P->block = (unsigned)std::min(1024, ((Nn + 63) / 64) * 64);
        P->grid = (unsigned)Nrep;
    } else {
        P->block = 256;
        P->grid = (unsigned)((lanes + 255) / 256);
    }

    if (B.status != hipSuccess) {
        delete P;
        return SIMINF_ERR_ALLOC_MEMORY_BUFFER;
    }
    if (hipStreamCreate(&P->stream) != hipSuccess ||
        hipEventCreate(&P->ev_start) != hipSuccess ||
        hipEventCreate(&P->ev_stop) != hipSuccess) {
        SimInf_hip_plan_destroy(P);
        return SIMINF_ERR_DEVICE;
    }
    *out = P;
    return 0;
}

extern "C" int SimInf_hip_plan_run(SimInf_hip_plan *P, unsigned long int seed)
{
    if (!P)
        return SIMINF_ERR_INVALID_MODEL;
    KParams p = P->p;
    p.seed = (unsigned long long)seed;

    const unsigned long long init_key = NO_ERROR_KEY, zero = 0;
    if (hipMemcpyAsync(p.err_key, &init_key, sizeof(init_key),
                       hipMemcpyHostToDevice, P->stream) != hipSuccess ||
        hipMemcpyAsync(p.transitions, &zero, sizeof(zero),
                       hipMemcpyHostToDevice, P->stream) != hipSuccess)
        return SIMINF_ERR_DEVICE;

    (void)hipEventRecord(P->ev_start, P->stream);
    switch (P->kernel) {
    case K_SSM_NODES:
        P->model->ssm_nodes(p, P->has_e1, P->grid, P->block, P->stream);
        break;
    case K_AEM_NODES:
        P->model->aem_nodes(p, P->has_e1, P->grid, P->block, P->stream);
        break;
    case K_SSM_REPS:
        if (P->model->ssm_reps(p, P->u_in_lds, P->grid, P->block, P->lds,
                               P->stream) != hipSuccess)
            return SIMINF_ERR_DEVICE;
        break;
    }
    (void)hipEventRecord(P->ev_stop, P->stream);
    if (hipGetLastError() != hipSuccess)
        return SIMINF_ERR_DEVICE;

    unsigned long long key = NO_ERROR_KEY, trans = 0;
    if (hipMemcpyAsync(&key, p.err_key, sizeof(key), hipMemcpyDeviceToHost,
                       P->stream) != hipSuccess ||
        hipMemcpyAsync(&trans, p.transitions, sizeof(trans),
                       hipMemcpyDeviceToHost, P->stream) != hipSuccess ||
        hipStreamSynchronize(P->stream) != hipSuccess)
        return SIMINF_ERR_DEVICE;

    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, P->ev_start, P->ev_stop);
    P->kernel_ms = ms;
    P->transitions = trans;
    return key_to_code(std::min<uint64_t>(key, P->host_err_key));
}

extern "C" int SimInf_hip_plan_download(SimInf_hip_plan *P, int *U, double *V,
                                        double *prU, double *prV)
{
    if (!P)
        return SIMINF_ERR_INVALID_MODEL;
    const size_t ncols = (size_t)P->Nrep * P->tlen;
    hipError_t e = hipSuccess;
    if (U && P->dense_U)
        e = hipMemcpy(U, P->p.U, (size_t)P->Nn * P->Nc * ncols * sizeof(int),
                      hipMemcpyDeviceToHost);
    if (e == hipSuccess && prU && !P->dense_U && P->nnzU)
        e = hipMemcpy(prU, P->p.prU, P->nnzU * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && V && P->dense_V && P->Nd > 0)
        e = hipMemcpy(V, P->p.V, (size_t)P->Nn * P->Nd * ncols * sizeof(double),
                      hipMemcpyDeviceToHost);
    if (e == hipSuccess && prV && !P->dense_V && P->nnzV)
        e = hipMemcpy(prV, P->p.prV, P->nnzV * sizeof(double), hipMemcpyDeviceToHost);
    return e == hipSuccess ? 0 : SIMINF_ERR_DEVICE;
}

extern "C" int SimInf_hip_plan_stats(const SimInf_hip_plan *P,
                                     long long *transitions, long long *node_days,
                                     double *kernel_ms)
{
    if (!P)
        return SIMINF_ERR_INVALID_MODEL;
    if (transitions)
        *transitions = (long long)P->transitions;
    if (node_days)
        *node_days = (long long)P->Nn * P->Nrep * P->ndays;
    if (kernel_ms)
        *kernel_ms = P->kernel_ms;
    return 0;
}

extern "C" void SimInf_hip_plan_destroy(SimInf_hip_plan *P)
{
    if (!P)
        return;
    if (P->stream)
        (void)hipStreamSynchronize(P->stream);
    if (P->ev_start)
        (void)hipEventDestroy(P->ev_start);
    if (P->ev_stop)
        (void)hipEventDestroy(P->ev_stop);
    if (P->stream)
        (void)hipStreamDestroy(P->stream);
    delete P;
}

/* ------------------------------------------------------------------ */
/* Blocking entry points.                                              */
/* ------------------------------------------------------------------ */

static int run_blocking(SimInf_solver_args *a, const SimInf_hip_model *model,
                        int solver)
{
    SimInf_solver_args b = *a;
    b.solver = solver;
    SimInf_hip_plan *P = nullptr;
    int err = SimInf_hip_plan_create(&P, &b, model);
    if (err)
        return err;
    err = SimInf_hip_plan_run(P, a->seed);
    if (!err)
        err = SimInf_hip_plan_download(P, a->U, a->V, a->U ? nullptr : a->prU,
                                       a->V ? nullptr : a->prV);
    SimInf_hip_plan_destroy(P);
    return err;
}

extern "C" int SimInf_hip_run_solver_ssm(SimInf_solver_args *a,
                                         const SimInf_hip_model *model)
{
    if (!a)
        return SIMINF_ERR_INVALID_MODEL;
    if (a->Nrep > 1)
        return SIMINF_ERR_INVALID_MODEL;
    return run_blocking(a, model, 0);
}

extern "C" int SimInf_hip_run_solver_mssm(SimInf_solver_args *a,
                                          const SimInf_hip_model *model)
{
    if (!a)
        return SIMINF_ERR_INVALID_MODEL;
    return run_blocking(a, model, 0);
}

extern "C" int SimInf_hip_run_solver_mssm_crn(SimInf_solver_args *a,
                                              const SimInf_hip_model *model)
{
    if (!a)
        return SIMINF_ERR_INVALID_MODEL;
    return run_blocking(a, model, 0);
}

extern "C" int SimInf_hip_run_solver_aem(SimInf_solver_args *a,
                                         const SimInf_hip_model *model)
{
    if (!a)
        return SIMINF_ERR_INVALID_MODEL;
    if (a->Nrep > 1)
        return SIMINF_ERR_AEM_REPLICATED_MODEL;
    return run_blocking(a, model, 1);
}

extern "C" int SimInf_hip_run(SimInf_solver_args *a,
                              const SimInf_hip_model *model)
{
    if (!a)
        return SIMINF_ERR_INVALID_MODEL;
    if (a->solver == 0) {
        if (a->Nrep > 1)
            return a->crn == 1 ? SimInf_hip_run_solver_mssm_crn(a, model)
                               : SimInf_hip_run_solver_mssm(a, model);
        return SimInf_hip_run_solver_ssm(a, model);
    }
    if (a->solver == 1) {
        if (a->Nrep > 1)
            return SIMINF_ERR_AEM_REPLICATED_MODEL;
        return SimInf_hip_run_solver_aem(a, model);
    }
    return SIMINF_ERR_UNKNOWN_SOLVER;
}

/* Messages of SimInf.c:28-97. */
extern "C" const char *SimInf_hip_error_message(int code)
{
    switch (code) {
    case 0: return "";
    case SIMINF_ERR_NEGATIVE_STATE: return "Negative state detected.";
    case SIMINF_ERR_ALLOC_MEMORY_BUFFER: return "Unable to allocate memory buffer.";
    case SIMINF_UNDEFINED_EVENT: return "Undefined event type.";
    case SIMINF_ERR_V_IS_NOT_FINITE: return "The continuous state 'v' is not finite.";
    case SIMINF_ERR_SAMPLE_SELECT: return "Unable to sample individuals for event.";
    case SIMINF_ERR_INVALID_MODEL: return "Invalid model.";
    case SIMINF_ERR_V_IS_NEGATIVE: return "The continuous state 'v' is negative.";
    case SIMINF_ERR_INVALID_RATE: return "Invalid rate detected (non-finite or < 0.0).";
    case SIMINF_ERR_UNKNOWN_SOLVER: return "Invalid 'solver' value.";
    case SIMINF_ERR_DEST_OUT_OF_BOUNDS: return "'dest' is out of bounds.";
    case SIMINF_ERR_NODE_OUT_OF_BOUNDS: return "'node' is out of bounds.";
    case SIMINF_ERR_EVENTS_N: return "'N' is invalid.";
    case SIMINF_ERR_EVENT_SHIFT: return "'shift' is invalid.";
    case SIMINF_ERR_SHIFT_OUT_OF_BOUNDS: return "'shift' is out of bounds.";
    case SIMINF_ERR_INVALID_PROPORTION: return "Invalid proportion detected (< 0.0 or > 1.0).";
    case SIMINF_ERR_AEM_REPLICATED_MODEL: return "Cannot run the 'aem' solver on a replicated model.";
    case SIMINF_ERR_SPARSE_MODEL: return "Invalid sparse matrix in model.";
    case SIMINF_ERR_INVALID_OPTIONS: return "Invalid options.";
    case SIMINF_ERR_INVALID_SEED: return "Invalid 'seed' value.";
    case SIMINF_ERR_INVALID_CRN: return "Invalid 'crn' value.";
    case SIMINF_ERR_INVALID_RNG_TYPE: return "Invalid 'rng_type' value.";
    case SIMINF_ERR_DEVICE: return "GPU runtime error.";
    default: return "Unknown error code.";
    }
}
