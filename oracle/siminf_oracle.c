/*
 * siminf_oracle.c -- CPU checker for the MI355X solver core.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (siminf_amd/, the HIP
 * library, bench.py's GPU leg) links, loads or calls this file; only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py do,
 * and only as the checker.
 *
 * A plain-C restatement of the reference's solver core, following
 *   src/solvers/SimInf_solver_ssm.c:74-373       (ssm day loop)
 *   src/solvers/SimInf_solver_mssm.c:81-402      (mssm replicate loop)
 *   src/solvers/SimInf_solver_mssm_crn.c:92-416  (crn variant)
 *   src/solvers/SimInf_solver_aem.c:61-459       (aem + calcTimes)
 *   src/misc/binheap.c:15-105                    (indexed binary heap)
 *   src/solvers/SimInf_solver.c:62-324, :658-950 (sampling, events,
 *                                                 sparse output)
 *   R/pfilter.R:315-499, src/misc/SimInf_particle_filter.c:32-66
 *                                                (bootstrap particle filter)
 * with one thread.  Random numbers come from the injected streams of
 * siminf_rand.h instead of GSL:
 *   - SSA draws and E1-event draws of (replicate r, node i) use the node
 *     stream (SIMINF_DOMAIN_NODE, r*Nn + i), in the order the reference
 *     draws them from its thread stream when that thread owns node i;
 *   - each external-transfer (E2) event k of replicate r uses its own
 *     stream (SIMINF_DOMAIN_E2, r*len + k);
 *   - aem channel (r, i, tr) uses (SIMINF_DOMAIN_AEM, (r*Nn + i)*Nt + tr)
 *     like the reference's per-(node, transition) GSL streams
 *     (SimInf_solver_aem.c:555-566).
 * Per day the order is: SSA (or aem) for every node, the E1 events with
 * time <= tt in list order, the E2 events with time <= tt in list order,
 * the post-time-step function, then writeback; for sorted events this is
 * the reference's order (SimInf_solver_ssm.c:144-362).
 *
 * Threads (oracle_set_threads): the node loops of the SSA / aem phase and
 * of the post-time-step phase, and the trajectory copy, run over T
 * contiguous node blocks, one per OpenMP thread, the static partition of
 * SimInf_solver.c:1094-1110; events run on one thread, as the
 * reference's E2 (SimInf_solver_ssm.c:244-250).  Every node keeps its own
 * stream, so results do not depend on T (with several errors in one day
 * the reported one may, as in the reference).  T = 1 is the checker.
 *
 * The stream schedule is this library's, not the reference's: the SSA
 * takes one aligned pair of uniforms per attempt (the reference draws the
 * second only when a transition fires, SimInf_solver_ssm.c:155, :165), E2
 * events draw from their own streams (the reference continues thread 0's
 * stream, SimInf_solver_ssm.c:249), so agreement with the reference is in
 * distribution, and bit-exact agreement is with the GPU path.
 *
 * Parity status: the loops are pinned by the reference's RNG-free golden
 * vectors (tests/SEIR.R, tests/phi.R, tests/sample_select.R), invariants
 * and closed-form distributions in tests/; the GSL samplers it replaces are
 * "parity unpinned" against GSL itself (GSL is not in this image).
 */
#include "../include/siminf_hip.h"
#include "../siminf_amd/csrc/siminf_rand.h"
#include "../siminf_amd/csrc/siminf_models.h"
#include "../siminf_amd/csrc/siminf_pfilter.h"
#include "../siminf_amd/csrc/siminf_abc.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static long long oracle_transitions_count;
static int oracle_threads = 1;

/* Threads of the node loops (1: one thread, the checker). */
void oracle_set_threads(int n)
{
    oracle_threads = n < 1 ? 1 : n;
}

/* Parameters per replicate, as the library's plans take them: ldata
 * Nld x Nn for each replicate in turn, gdata gstride values for each
 * replicate in turn (NULL: the arguments' shared ones). */
static const double *oracle_rep_ldata, *oracle_rep_gdata;
static int oracle_rep_gstride, oracle_rep_lnodes;
static double *oracle_rep_lbuf;

/* ldata_nodes: Nn for a column per node of every replicate, 1 for one
 * column per replicate shared by its nodes. */
void oracle_set_replicate_data(const double *ldata, int ldata_nodes, const double *gdata,
                               int gstride)
{
    oracle_rep_ldata = ldata;
    oracle_rep_lnodes = ldata_nodes;
    oracle_rep_gdata = gdata;
    oracle_rep_gstride = gstride;
}

/* The arguments a replicate runs with: its own parameters if set. */
static SimInf_solver_args oracle_replicate_args(const SimInf_solver_args *a, int rep)
{
    SimInf_solver_args r = *a;
    if (oracle_rep_ldata && oracle_rep_lnodes == 1) {
        /* The replicate's values spread over its nodes. */
        free(oracle_rep_lbuf);
        oracle_rep_lbuf = malloc((size_t)a->Nn * a->Nld * sizeof(double) + 8);
        for (int i = 0; i < a->Nn; i++)
            memcpy(oracle_rep_lbuf + (ptrdiff_t)i * a->Nld,
                   oracle_rep_ldata + (ptrdiff_t)rep * a->Nld, (size_t)a->Nld * sizeof(double));
        r.ldata = oracle_rep_lbuf;
    } else if (oracle_rep_ldata) {
        r.ldata = oracle_rep_ldata + (ptrdiff_t)rep * a->Nn * a->Nld;
    }
    if (oracle_rep_gdata)
        r.gdata = oracle_rep_gdata + (ptrdiff_t)rep * oracle_rep_gstride;
    return r;
}

/* ---------------------------------------------------------------- */
/* Sampling (SimInf_solver.c:62-324).                                */
/* ---------------------------------------------------------------- */

/* Individuals to remove/move for an event; SimInf_solver.c:62-214. */
static int oracle_sample_select(
    const int *irE, const int *jcE, const double *prE, int Nc,
    const int *u_node, int select, int n, double proportion,
    int *individuals, siminf_stream *rng)
{
    const int k0 = jcE[select], k1 = jcE[select + 1];
    const int Nstates = k1 - k0;
    int Nindividuals = 0, Nkinds = 0, i;

    memset(individuals, 0, (size_t)Nc * sizeof(int));

    for (i = k0; i < k1; i++) {
        const int nk = u_node[irE[i]];
        if (nk > 0)
            Nkinds++;
        Nindividuals += nk;
    }

    if (n == 0) {
        if (proportion < 0 || proportion > 1)
            return SIMINF_ERR_INVALID_PROPORTION;
        n = siminf_binomial(rng, proportion, Nindividuals);
    }

    if (Nstates <= 0 || n > Nindividuals || n < 0)
        return SIMINF_ERR_SAMPLE_SELECT;

    if (n == 0)
        return 0;
    if (Nindividuals == n) {
        for (i = k0; i < k1; i++)
            individuals[irE[i]] = u_node[irE[i]];
        return 0;
    }
    if (Nstates == 1) {
        individuals[irE[k0]] = n;
        return 0;
    }
    if (Nkinds == 1) {
        for (i = k0; i < k1; i++) {
            if (u_node[irE[i]] > 0) {
                individuals[irE[i]] = n;
                break;
            }
        }
        return 0;
    }

    /* Equal weights: multivariate hypergeometric by conditional
     * univariate draws over the column's rows in order. */
    {
        int equal = 1;
        for (i = k0 + 1; i < k1; i++) {
            if (prE[i] != prE[i - 1]) {
                equal = 0;
                break;
            }
        }
        if (equal) {
            for (i = k0; i < k1 - 1; i++) {
                const int nk = u_node[irE[i]];
                if (n == 0)
                    break;
                individuals[irE[i]] =
                    siminf_hypergeometric(rng, nk, Nindividuals - nk, n);
                Nindividuals -= nk;
                n -= individuals[irE[i]];
            }
            individuals[irE[i]] = n;
            return 0;
        }
    }

    /* Unequal weights: Wallenius urn, one individual per draw. */
    while (n > 0) {
        double cum = 0.0, rnd;
        for (i = k0; i < k1; i++)
            cum += prE[i] * (u_node[irE[i]] - individuals[irE[i]]);
        rnd = siminf_uniform(rng) * cum;
        i = k0;
        cum = prE[i] * (u_node[irE[i]] - individuals[irE[i]]);
        while (i < k1 && rnd > cum) {
            i++;
            if (i < k1)
                cum += prE[i] * (u_node[irE[i]] - individuals[irE[i]]);
        }
        if (i >= k1)
            i = k1 - 1;
        if (prE[i] * (u_node[irE[i]] - individuals[irE[i]]) == 0.0) {
            while (i > k0 &&
                   prE[i] * (u_node[irE[i]] - individuals[irE[i]]) == 0.0)
                i--;
            if (prE[i] * (u_node[irE[i]] - individuals[irE[i]]) == 0.0)
                return SIMINF_ERR_SAMPLE_SELECT;
        }
        individuals[irE[i]] += 1;
        n--;
    }
    return 0;
}

/* Individuals to enter a node; SimInf_solver.c:243-324. */
static int oracle_sample_select_enter(
    const int *irE, const int *jcE, const double *prE, int Nc,
    const int *u_node, int select, int n, double proportion,
    int *individuals, siminf_stream *rng)
{
    const int k0 = jcE[select], k1 = jcE[select + 1];
    const int Nstates = k1 - k0;
    double w_cum = 0.0;
    int i;

    memset(individuals, 0, (size_t)Nc * sizeof(int));

    if (n == 0) {
        if (proportion < 0 || proportion > 1)
            return SIMINF_ERR_INVALID_PROPORTION;
        for (i = k0; i < k1; i++)
            n += u_node[irE[i]];
        n = siminf_binomial(rng, proportion, n);
    }

    if (Nstates <= 0 || n < 0)
        return SIMINF_ERR_SAMPLE_SELECT;

    if (n == 0)
        return 0;
    if (Nstates == 1) {
        individuals[irE[k0]] = n;
        return 0;
    }

    for (i = k0; i < k1; i++)
        w_cum += prE[i];

    while (n > 0) {
        const double rnd = siminf_uniform(rng) * w_cum;
        double cum;
        i = k0;
        cum = prE[i];
        while (i < k1 && rnd > cum) {
            i++;
            if (i < k1)
                cum += prE[i];
        }
        if (i >= k1)
            i = k1 - 1;
        if (prE[i] == 0.0) {
            while (i > k0 && prE[i] == 0.0)
                i--;
            if (prE[i] == 0.0)
                return SIMINF_ERR_SAMPLE_SELECT;
        }
        individuals[irE[i]] += 1;
        n--;
    }
    return 0;
}

/* ---------------------------------------------------------------- */
/* Event processing (SimInf_solver.c:658-919).                       */
/* ---------------------------------------------------------------- */

typedef struct oracle_ctx {
    SimInf_solver_args *a;
    int Nn, Nc, Nt, Nd, Nld;
    int *individuals;
} oracle_ctx;

/* Apply event k (zero-based fields derived on the fly) to the replicate
 * state u (Nn x Nc).  Returns 0 or an error code. */
static int oracle_apply_event(
    oracle_ctx *c, int k, int *u, int *update_node, siminf_stream *rng)
{
    const SimInf_solver_args *a = c->a;
    const int Nc = c->Nc;
    const int type = a->event[k];
    const int node = a->node[k] - 1;
    const int dest = a->dest[k] - 1;
    const int select = a->select[k] - 1;
    const int shift = a->shift[k] - 1;
    const int n = a->n[k];
    const double proportion = a->proportion[k];
    const int *irE = a->irE, *jcE = a->jcE;
    const double *prE = a->prE;
    int *ind = c->individuals;
    int err, i;

    if (node < 0 || node >= c->Nn)
        return SIMINF_ERR_NODE_OUT_OF_BOUNDS;

    switch (type) {
    case SIMINF_EXIT_EVENT:
        err = oracle_sample_select(irE, jcE, prE, Nc, &u[node * Nc], select,
                                   n, proportion, ind, rng);
        if (err)
            return err;
        for (i = jcE[select]; i < jcE[select + 1]; i++) {
            const int jj = irE[i];
            u[node * Nc + jj] -= ind[jj];
            if (u[node * Nc + jj] < 0)
                return SIMINF_ERR_NEGATIVE_STATE;
        }
        break;

    case SIMINF_ENTER_EVENT:
        err = oracle_sample_select_enter(irE, jcE, prE, Nc, &u[node * Nc],
                                         select, n, proportion, ind, rng);
        if (err)
            return err;
        for (i = jcE[select]; i < jcE[select + 1]; i++) {
            const int jj = irE[i];
            const int kn = node * Nc + jj;
            if (shift < 0) {
                u[kn] += ind[jj];
                if (u[kn] < 0)
                    return SIMINF_ERR_NEGATIVE_STATE;
            } else if (!a->N) {
                return SIMINF_ERR_EVENTS_N;
            } else {
                const int ll = a->N[shift * Nc + jj];
                if (jj + ll < 0 || jj + ll >= Nc)
                    return SIMINF_ERR_SHIFT_OUT_OF_BOUNDS;
                u[kn + ll] += ind[jj];
                if (u[kn + ll] < 0)
                    return SIMINF_ERR_NEGATIVE_STATE;
            }
        }
        break;

    case SIMINF_INTERNAL_TRANSFER_EVENT:
        if (!a->N)
            return SIMINF_ERR_EVENTS_N;
        if (shift < 0)
            return SIMINF_ERR_EVENT_SHIFT;
        err = oracle_sample_select(irE, jcE, prE, Nc, &u[node * Nc], select,
                                   n, proportion, ind, rng);
        if (err)
            return err;
        for (i = jcE[select]; i < jcE[select + 1]; i++) {
            const int jj = irE[i];
            const int kn = node * Nc + jj;
            const int ll = a->N[shift * Nc + jj];
            if (jj + ll < 0 || jj + ll >= Nc)
                return SIMINF_ERR_SHIFT_OUT_OF_BOUNDS;
            u[kn + ll] += ind[jj];
            if (u[kn + ll] < 0)
                return SIMINF_ERR_NEGATIVE_STATE;
            u[kn] -= ind[jj];
            if (u[kn] < 0)
                return SIMINF_ERR_NEGATIVE_STATE;
        }
        break;

    case SIMINF_EXTERNAL_TRANSFER_EVENT:
        if (dest < 0 || dest >= c->Nn)
            return SIMINF_ERR_DEST_OUT_OF_BOUNDS;
        err = oracle_sample_select(irE, jcE, prE, Nc, &u[node * Nc], select,
                                   n, proportion, ind, rng);
        if (err)
            return err;
        for (i = jcE[select]; i < jcE[select + 1]; i++) {
            const int jj = irE[i];
            const int kd = dest * Nc + jj;
            const int kn = node * Nc + jj;
            if (shift < 0) {
                u[kd] += ind[jj];
                if (u[kd] < 0)
                    return SIMINF_ERR_NEGATIVE_STATE;
            } else if (!a->N) {
                return SIMINF_ERR_EVENTS_N;
            } else {
                const int ll = a->N[shift * Nc + jj];
                if (jj + ll < 0 || jj + ll >= Nc)
                    return SIMINF_ERR_SHIFT_OUT_OF_BOUNDS;
                u[kd + ll] += ind[jj];
                if (u[kd + ll] < 0)
                    return SIMINF_ERR_NEGATIVE_STATE;
            }
            u[kn] -= ind[jj];
            if (u[kn] < 0)
                return SIMINF_ERR_NEGATIVE_STATE;
        }
        update_node[dest] = 1;
        break;

    default:
        return SIMINF_UNDEFINED_EVENT;
    }

    update_node[node] = 1;
    return 0;
}

/* Stream identity of the E2 events of a run over part of an event list
 * (a filter interval's events): the full list's length and the part's
 * offset in it; -1 means the run's own list. */
static int g_ev_total = -1, g_ev_offset = 0;

/* Process the events of replicate rep due at tt: first the E1 events in
 * list order (each drawing from its node's stream), then the E2 events
 * in list order (each drawing from its own stream).  *cursor1 / *cursor2
 * are the next unprocessed positions in the E1 / E2 subsequences. */
static int oracle_process_events(
    oracle_ctx *c, int rep, double tt, int *u, int *update_node,
    siminf_stream *node_rng, int *cursor1, int *cursor2)
{
    const SimInf_solver_args *a = c->a;
    int k, err;

    for (k = *cursor1; k < a->len; k++) {
        if (a->event[k] == SIMINF_EXTERNAL_TRANSFER_EVENT)
            continue;
        if (a->time[k] > tt)
            break;
        {
            const int node = a->node[k] - 1;
            siminf_stream *rng = (node >= 0 && node < c->Nn) ?
                &node_rng[node] : &node_rng[0];
            err = oracle_apply_event(c, k, u, update_node, rng);
            if (err) {
                *cursor1 = k;
                return err;
            }
        }
    }
    *cursor1 = k;

    for (k = *cursor2; k < a->len; k++) {
        siminf_stream rng;
        if (a->event[k] != SIMINF_EXTERNAL_TRANSFER_EVENT)
            continue;
        if (a->time[k] > tt)
            break;
        rng = siminf_stream_make(a->seed, SIMINF_DOMAIN_E2,
                                 (uint64_t)rep *
                                     (uint64_t)(g_ev_total >= 0 ? g_ev_total : a->len) +
                                 (uint64_t)(g_ev_offset + k));
        err = oracle_apply_event(c, k, u, update_node, &rng);
        if (err) {
            *cursor2 = k;
            return err;
        }
    }
    *cursor2 = k;
    return 0;
}

/* ---------------------------------------------------------------- */
/* Writeback (SimInf_solver_ssm.c:318-334, mssm.c:323-371,            */
/* SimInf_solver.c:930-950).                                          */
/* ---------------------------------------------------------------- */

static void oracle_store(
    oracle_ctx *c, int rep, double tt, const int *u, const double *v_new,
    int *U_it, int *V_it)
{
    SimInf_solver_args *a = c->a;
    const ptrdiff_t Nn = c->Nn, Nc = c->Nc, Nd = c->Nd, tlen = a->tlen;

    while (*U_it < a->tlen && tt > a->tspan[*U_it]) {
        if (a->U) {
            int *dst = &a->U[(ptrdiff_t)(rep * tlen + *U_it) * Nn * Nc];
            if (oracle_threads > 1) {
                /* Each thread copies its node block. */
                const int T = oracle_threads;
#pragma omp parallel for schedule(static) num_threads(T)
                for (int b = 0; b < T; b++) {
                    const ptrdiff_t n0 = Nn * b / T, n1 = Nn * (b + 1) / T;
                    memcpy(dst + n0 * Nc, u + n0 * Nc, (size_t)((n1 - n0) * Nc) * sizeof(int));
                }
            } else {
                memcpy(dst, u, (size_t)(Nn * Nc) * sizeof(int));
            }
        } else {
            const ptrdiff_t col = rep * tlen + *U_it;
            for (int j = a->jcU[col]; j < a->jcU[col + 1]; j++)
                a->prU[j] = u[a->irU[j]];
        }
        (*U_it)++;
    }

    while (*V_it < a->tlen && tt > a->tspan[*V_it]) {
        if (a->V) {
            if (Nn * Nd > 0)
                memcpy(&a->V[(ptrdiff_t)(rep * tlen + *V_it) * Nn * Nd], v_new,
                       (size_t)(Nn * Nd) * sizeof(double));
        } else {
            const ptrdiff_t col = rep * tlen + *V_it;
            for (int j = a->jcV[col]; j < a->jcV[col + 1]; j++)
                a->prV[j] = v_new[a->irV[j]];
        }
        (*V_it)++;
    }
}

static int oracle_rate_ok(double rate)
{
    return siminf_isfinite(rate) && !(rate < 0.0);
}

/* ---------------------------------------------------------------- */
/* Direct-method day loop for one replicate (ssm.c:74-373 with one     */
/* thread, mssm.c:105-383 for one replicate).                          */
/* ---------------------------------------------------------------- */

/* The SSA of one node over [t_time[node], next_unit_of_time)
 * (ssm.c:144-235); returns 0 or an error code and counts the firings in
 * *fired. */
static int oracle_ssa_node(SimInf_solver_args *a, int node, int Nc, int Nt, int Nd,
                           int Nld, int *u, const double *v, double *t_rate,
                           double *sum_t_rate, double *t_time, siminf_stream *rng,
                           double next_unit_of_time, long long *fired)
{
    int error = 0;
    for (;;) {
        double cum, rnd, tau, u_tau, u_tr, delta = 0.0;
        int tr;

        if (sum_t_rate[node] <= 0.0) {
            t_time[node] = next_unit_of_time;
            break;
        }
        /* One aligned pair of uniforms per attempt: the waiting
         * time from the first, the transition from the second. */
        siminf_uniform2(&rng[node], &u_tau, &u_tr);
        tau = -siminf_log(u_tau) / sum_t_rate[node];
        if ((tau + t_time[node]) >= next_unit_of_time) {
            t_time[node] = next_unit_of_time;
            break;
        }
        t_time[node] += tau;

        rnd = u_tr * sum_t_rate[node];
        tr = 0;
        cum = t_rate[node * Nt];
        while (tr < Nt && rnd > cum) {
            tr++;
            if (tr < Nt)
                cum += t_rate[node * Nt + tr];
        }

        if (tr >= Nt)
            tr = Nt - 1;
        if (t_rate[node * Nt + tr] == 0.0) {
            while (tr > 0 && t_rate[node * Nt + tr] == 0.0)
                tr--;
            if (t_rate[node * Nt + tr] == 0.0) {
                sum_t_rate[node] = 0.0;
                break;
            }
        }

        (*fired)++;

        for (int j = a->jcS[tr]; j < a->jcS[tr + 1]; j++) {
            u[node * Nc + a->irS[j]] += a->prS[j];
            if (u[node * Nc + a->irS[j]] < 0)
                error = SIMINF_ERR_NEGATIVE_STATE;
        }

        for (int j = a->jcG[tr]; j < a->jcG[tr + 1]; j++) {
            const int k = a->irG[j];
            const double old = t_rate[node * Nt + k];
            const double rate = a->tr_fun[k](&u[node * Nc], &v[node * Nd],
                                             &a->ldata[node * Nld],
                                             a->gdata, t_time[node]);
            t_rate[node * Nt + k] = rate;
            delta += rate - old;
            if (!oracle_rate_ok(rate))
                error = SIMINF_ERR_INVALID_RATE;
        }
        sum_t_rate[node] += delta;

        /* The node stops at its first failing step; the
         * reference keeps drawing until the day ends. */
        if (error)
            break;
    }
    return error;
}

/* Post time step and rate refresh of one node (ssm.c:262-302): returns
 * the post-time-step function's error code (< 0), else 0; an invalid
 * rate sets *error and the loop goes on, as in the reference. */
static int oracle_pts_node(SimInf_solver_args *a, int node, int Nc, int Nt, int Nd,
                           int Nld, const int *u, const double *v, double *v_new,
                           double *t_rate, double *sum_t_rate, int *update_node,
                           double tt, int *error)
{
    const int rc = a->pts_fun(&v_new[node * Nd], &u[node * Nc],
                              &v[node * Nd], &a->ldata[node * Nld],
                              a->gdata, node, tt);
    if (rc < 0)
        return rc;
    if (rc > 0 || update_node[node]) {
        double delta = 0.0;
        for (int j = 0; j < Nt; j++) {
            const double old = t_rate[node * Nt + j];
            const double rate = a->tr_fun[j](&u[node * Nc],
                                             &v_new[node * Nd],
                                             &a->ldata[node * Nld],
                                             a->gdata, tt);
            t_rate[node * Nt + j] = rate;
            delta += rate - old;
            if (!oracle_rate_ok(rate))
                *error = SIMINF_ERR_INVALID_RATE;
        }
        sum_t_rate[node] += delta;
        update_node[node] = 0;
    }
    return 0;
}

static int oracle_direct_replicate(oracle_ctx *c, int rep, int is_mssm)
{
    SimInf_solver_args args = oracle_replicate_args(c->a, rep);
    SimInf_solver_args *a = &args;
    const int Nn = c->Nn, Nc = c->Nc, Nt = c->Nt, Nd = c->Nd, Nld = c->Nld;
    int *u = malloc((size_t)Nn * Nc * sizeof(int) + 1);
    double *v = malloc((size_t)Nn * Nd * sizeof(double) + 8);
    double *v_new = malloc((size_t)Nn * Nd * sizeof(double) + 8);
    double *t_rate = malloc((size_t)Nn * Nt * sizeof(double) + 8);
    double *sum_t_rate = malloc((size_t)Nn * sizeof(double) + 8);
    double *t_time = malloc((size_t)Nn * sizeof(double) + 8);
    int *update_node = calloc((size_t)Nn + 1, sizeof(int));
    siminf_stream *rng = malloc((size_t)Nn * sizeof(siminf_stream) + 1);
    int error = 0, U_it = 0, V_it = 0, cursor1 = 0, cursor2 = 0;
    double tt = a->tspan[0];
    double next_unit_of_time = floor(tt) + 1.0;

    memcpy(u, &a->u0[(ptrdiff_t)rep * Nn * Nc], (size_t)Nn * Nc * sizeof(int));
    if (Nd > 0) {
        memcpy(v, &a->v0[(ptrdiff_t)rep * Nn * Nd], (size_t)Nn * Nd * sizeof(double));
        memcpy(v_new, v, (size_t)Nn * Nd * sizeof(double));
    }
    for (int node = 0; node < Nn; node++)
        rng[node] = siminf_stream_make(a->seed, SIMINF_DOMAIN_NODE,
                                       (uint64_t)rep * (uint64_t)Nn +
                                       (uint64_t)node);

    /* Initial rates (ssm.c:96-117). */
    for (int node = 0; node < Nn && (!is_mssm || !error); node++) {
        sum_t_rate[node] = 0.0;
        for (int j = 0; j < Nt; j++) {
            const double rate = a->tr_fun[j](&u[node * Nc], &v[node * Nd],
                                             &a->ldata[node * Nld], a->gdata,
                                             tt);
            t_rate[node * Nt + j] = rate;
            sum_t_rate[node] += rate;
            if (!oracle_rate_ok(rate))
                error = SIMINF_ERR_INVALID_RATE;
        }
        t_time[node] = tt;
    }

    while (!error) {
        /* (1) SSA per node (ssm.c:144-235): one thread, or the node blocks
         * of the threads. */
        if (oracle_threads <= 1) {
            long long fired = 0;
            for (int node = 0; node < Nn && !error; node++)
                error = oracle_ssa_node(a, node, Nc, Nt, Nd, Nld, u, v, t_rate, sum_t_rate,
                                        t_time, rng, next_unit_of_time, &fired);
            oracle_transitions_count += fired;
        } else {
            long long fired = 0;
            const int T = oracle_threads;
#pragma omp parallel num_threads(T) reduction(+:fired)
            {
                int my = 0;
#pragma omp for schedule(static)
                for (int node = 0; node < Nn; node++)
                    if (!my)
                        my = oracle_ssa_node(a, node, Nc, Nt, Nd, Nld, u, v, t_rate,
                                             sum_t_rate, t_time, rng, next_unit_of_time,
                                             &fired);
#pragma omp critical
                if (my && !error)
                    error = my;
            }
            oracle_transitions_count += fired;
        }

        /* (2, 3) Scheduled events. */
        if (!error)
            error = oracle_process_events(c, rep, tt, u, update_node, rng,
                                          &cursor1, &cursor2);

        /* (4) Post time step and rate refresh (ssm.c:262-302). */
        if (oracle_threads <= 1) {
            for (int node = 0; node < Nn && (!is_mssm || !error); node++) {
                const int rc = oracle_pts_node(a, node, Nc, Nt, Nd, Nld, u, v, v_new, t_rate,
                                               sum_t_rate, update_node, tt, &error);
                if (rc < 0) {
                    error = rc;
                    break;
                }
            }
        } else {
            const int T = oracle_threads;
#pragma omp parallel num_threads(T)
            {
                int my = 0;
#pragma omp for schedule(static)
                for (int node = 0; node < Nn; node++) {
                    if (my < 0 && my != SIMINF_ERR_INVALID_RATE)
                        continue;
                    const int rc = oracle_pts_node(a, node, Nc, Nt, Nd, Nld, u, v, v_new,
                                                   t_rate, sum_t_rate, update_node, tt, &my);
                    if (rc < 0)
                        my = rc;
                }
#pragma omp critical
                if (my && !error)
                    error = my;
            }
        }

        /* (5) Advance time. */
        tt = next_unit_of_time;
        next_unit_of_time += 1.0;

        /* (6) Writeback, then swap v and v_new. */
        oracle_store(c, rep, tt, u, v_new, &U_it, &V_it);
        {
            double *tmp = v;
            v = v_new;
            v_new = tmp;
        }

        if (U_it >= a->tlen)
            break;
    }

    free(u); free(v); free(v_new); free(t_rate); free(sum_t_rate);
    free(t_time); free(update_node); free(rng);
    return error;
}

/* ---------------------------------------------------------------- */
/* aem (SimInf_solver_aem.c:61-459, binheap.c:15-105).                */
/* ---------------------------------------------------------------- */

static long long oracle_aem_finite_ties;

static void heap_percolate_down(int n1, double *data, int *INDEX, int *INDEX2,
                                int N)
{
    int child, node = n1;
    const double key = data[node];
    const int j = INDEX[node];
    while ((child = (node << 1) + 1) < N) {
        if (child != N - 1 && data[child + 1] < data[child])
            child++;
        if (data[child] < key) {
            const int rxn = INDEX[child];
            data[node] = data[child];
            INDEX2[rxn] = node;
            INDEX[node] = rxn;
        } else {
            break;
        }
        node = child;
    }
    data[node] = key;
    INDEX[node] = j;
    INDEX2[j] = node;
}

static void heap_percolate_up(int node, double *data, int *INDEX, int *INDEX2)
{
    int parent;
    const double key = data[node];
    const int j = INDEX[node];
    do {
        parent = (node - 1) >> 1;
        if (key < data[parent]) {
            const int rxn = INDEX[parent];
            data[node] = data[parent];
            INDEX2[rxn] = node;
            INDEX[node] = rxn;
        } else {
            break;
        }
        node = parent;
    } while (parent > 0);
    data[node] = key;
    INDEX[node] = j;
    INDEX2[j] = node;
}

static void heap_initialize(double *data, int *INDEX, int *INDEX2, int N)
{
    for (int i = (N - 1) >> 1; i >= 0; i--)
        heap_percolate_down(i, data, INDEX, INDEX2, N);
}

static void heap_update(int node, double *data, int *INDEX, int *INDEX2, int N)
{
    const int parent = (node - 1) >> 1;
    if (node > 0 && data[node] < data[parent])
        heap_percolate_up(node, data, INDEX, INDEX2);
    else
        heap_percolate_down(node, data, INDEX, INDEX2, N);
}

/* calcTimes (SimInf_solver_aem.c:61-86). */
static void oracle_calc_times(double *time, double *infTime, double tt,
                              double old_rate, double new_rate,
                              siminf_stream *rng)
{
    const double oldtime = time[0];
    if (isinf(oldtime)) {
        if (infTime[0] == 0.0)
            time[0] = -siminf_log(siminf_uniform(rng)) / new_rate + tt;
        else if (new_rate > 0.0)
            time[0] = tt + (infTime[0] / new_rate);
    } else if (new_rate >= 2.2250738585072014e-308) {
        if (oldtime == tt)
            time[0] = -siminf_log(siminf_uniform(rng)) / new_rate + tt;
        else
            time[0] = ((old_rate / new_rate) * (oldtime - tt)) + tt;
    } else {
        infTime[0] = (oldtime - tt) * old_rate;
        time[0] = INFINITY;
    }
}

/* Count exact ties among the finite next-reaction times of a node: the
 * heap and a first-index argmin may then pick different channels. */
static void oracle_count_ties(const double *times, int Nt)
{
    for (int i = 0; i < Nt; i++)
        for (int j = i + 1; j < Nt; j++)
            if (!isinf(times[i]) && times[i] == times[j]) {
#pragma omp atomic
                oracle_aem_finite_ties++;
            }
}

/* The all-events method of one node up to next_unit_of_time
 * (aem.c:188-304); returns 0 or an error code and counts the firings in
 * *fired. */
static int oracle_aem_node(SimInf_solver_args *a, int node, int Nc, int Nt, int Nd,
                           int Nld, int *u, const double *v, double *t_rate,
                           double *t_time, siminf_stream *rng_vec, int *reactHeap,
                           int *reactNode, double *reactTimes, double *reactInf,
                           double next_unit_of_time, long long *fired)
{
    int error = 0;
    for (;;) {
        int tr;
        double old_t_rate, rate;

        oracle_count_ties(&reactTimes[Nt * node], Nt);
        t_time[node] = reactTimes[Nt * node];
        if (isinf(t_time[node]) || t_time[node] >= next_unit_of_time) {
            t_time[node] = next_unit_of_time;
            break;
        }

        tr = reactNode[Nt * node] % Nt;
        (*fired)++;

        for (int j = a->jcS[tr]; j < a->jcS[tr + 1]; j++) {
            u[node * Nc + a->irS[j]] += a->prS[j];
            if (u[node * Nc + a->irS[j]] < 0)
                error = SIMINF_ERR_NEGATIVE_STATE;
        }

        for (int ii = a->jcG[tr]; ii < a->jcG[tr + 1]; ii++) {
            const int j = a->irG[ii];
            if (j == tr)
                continue;
            old_t_rate = t_rate[node * Nt + j];
            rate = a->tr_fun[j](&u[node * Nc], &v[node * Nd],
                                &a->ldata[node * Nld], a->gdata,
                                t_time[node]);
            t_rate[node * Nt + j] = rate;
            if (!oracle_rate_ok(rate))
                error = SIMINF_ERR_INVALID_RATE;
            oracle_calc_times(&reactTimes[Nt * node + reactHeap[Nt * node + j]],
                              &reactInf[Nt * node + j], t_time[node],
                              old_t_rate, t_rate[node * Nt + j],
                              &rng_vec[Nt * node + j]);
            heap_update(reactHeap[Nt * node + j], &reactTimes[Nt * node],
                        &reactNode[Nt * node], &reactHeap[Nt * node], Nt);
        }

        old_t_rate = t_rate[node * Nt + tr];
        rate = a->tr_fun[tr](&u[node * Nc], &v[node * Nd],
                             &a->ldata[node * Nld], a->gdata,
                             t_time[node]);
        t_rate[node * Nt + tr] = rate;
        if (!oracle_rate_ok(rate))
            error = SIMINF_ERR_INVALID_RATE;
        oracle_calc_times(&reactTimes[Nt * node + reactHeap[Nt * node + tr]],
                          &reactInf[Nt * node + tr], t_time[node],
                          old_t_rate, t_rate[node * Nt + tr],
                          &rng_vec[Nt * node + tr]);
        heap_update(reactHeap[Nt * node + tr], &reactTimes[Nt * node],
                    &reactNode[Nt * node], &reactHeap[Nt * node], Nt);
        if (error)
            break;
    }
    return error;
}

/* aem's post time step and refresh through calcTimes of one node
 * (aem.c:329-384); as oracle_pts_node. */
static int oracle_aem_pts_node(SimInf_solver_args *a, int node, int Nc, int Nt, int Nd,
                               int Nld, const int *u, const double *v, double *v_new,
                               double *t_rate, double *t_time, siminf_stream *rng_vec,
                               int *reactHeap, int *reactNode, double *reactTimes,
                               double *reactInf, int *update_node, double tt, int *error)
{
    const int rc = a->pts_fun(&v_new[node * Nd], &u[node * Nc],
                              &v[node * Nd], &a->ldata[node * Nld],
                              a->gdata, node, tt);
    if (rc < 0)
        return rc;
    if (rc > 0 || update_node[node]) {
        for (int j = 0; j < Nt; j++) {
            const double old = t_rate[node * Nt + j];
            const double rate = a->tr_fun[j](&u[node * Nc],
                                             &v_new[node * Nd],
                                             &a->ldata[node * Nld],
                                             a->gdata, tt);
            t_rate[node * Nt + j] = rate;
            if (!oracle_rate_ok(rate))
                *error = SIMINF_ERR_INVALID_RATE;
            oracle_calc_times(&reactTimes[Nt * node + reactHeap[Nt * node + j]],
                              &reactInf[Nt * node + j], t_time[node],
                              old, t_rate[node * Nt + j],
                              &rng_vec[Nt * node + j]);
            heap_update(reactHeap[Nt * node + j], &reactTimes[Nt * node],
                        &reactNode[Nt * node], &reactHeap[Nt * node], Nt);
        }
        update_node[node] = 0;
    }
    return 0;
}

static int oracle_aem(oracle_ctx *c)
{
    SimInf_solver_args args = oracle_replicate_args(c->a, 0);
    SimInf_solver_args *a = &args;
    const int Nn = c->Nn, Nc = c->Nc, Nt = c->Nt, Nd = c->Nd, Nld = c->Nld;
    int *u = malloc((size_t)Nn * Nc * sizeof(int) + 1);
    double *v = malloc((size_t)Nn * Nd * sizeof(double) + 8);
    double *v_new = malloc((size_t)Nn * Nd * sizeof(double) + 8);
    double *t_rate = malloc((size_t)Nn * Nt * sizeof(double) + 8);
    double *t_time = malloc((size_t)Nn * sizeof(double) + 8);
    int *update_node = calloc((size_t)Nn + 1, sizeof(int));
    siminf_stream *node_rng = malloc((size_t)Nn * sizeof(siminf_stream) + 1);
    siminf_stream *rng_vec = malloc((size_t)Nn * Nt * sizeof(siminf_stream) + 1);
    int *reactHeap = malloc((size_t)Nn * Nt * sizeof(int) + 1);
    int *reactNode = malloc((size_t)Nn * Nt * sizeof(int) + 1);
    double *reactTimes = malloc((size_t)Nn * Nt * sizeof(double) + 8);
    double *reactInf = calloc((size_t)Nn * Nt + 1, sizeof(double));
    int error = 0, U_it = 0, V_it = 0, cursor1 = 0, cursor2 = 0;
    double tt = a->tspan[0];
    double next_unit_of_time = floor(tt) + 1.0;

    memcpy(u, a->u0, (size_t)Nn * Nc * sizeof(int));
    if (Nd > 0) {
        memcpy(v, a->v0, (size_t)Nn * Nd * sizeof(double));
        memcpy(v_new, v, (size_t)Nn * Nd * sizeof(double));
    }
    for (int node = 0; node < Nn; node++) {
        node_rng[node] = siminf_stream_make(a->seed, SIMINF_DOMAIN_NODE,
                                            (uint64_t)node);
        for (int j = 0; j < Nt; j++)
            rng_vec[node * Nt + j] =
                siminf_stream_make(a->seed, SIMINF_DOMAIN_AEM,
                                   (uint64_t)node * (uint64_t)Nt + (uint64_t)j);
    }

    /* Initial rates and reaction times (aem.c:117-153). */
    for (int node = 0; node < Nn; node++) {
        for (int j = 0; j < Nt; j++) {
            const double rate = a->tr_fun[j](&u[node * Nc], &v[node * Nd],
                                             &a->ldata[node * Nld], a->gdata, tt);
            t_rate[node * Nt + j] = rate;
            if (!oracle_rate_ok(rate))
                error = SIMINF_ERR_INVALID_RATE;
            reactTimes[Nt * node + j] =
                -siminf_log(siminf_uniform(&rng_vec[Nt * node + j])) / rate + tt;
            if (reactTimes[Nt * node + j] <= 0.0)
                reactTimes[Nt * node + j] = INFINITY;
            reactHeap[Nt * node + j] = reactNode[Nt * node + j] = j;
        }
        heap_initialize(&reactTimes[Nt * node], &reactNode[Nt * node],
                        &reactHeap[Nt * node], Nt);
        t_time[node] = tt;
    }

    while (!error) {
        if (oracle_threads <= 1) {
            long long fired = 0;
            for (int node = 0; node < Nn && !error; node++)
                error = oracle_aem_node(a, node, Nc, Nt, Nd, Nld, u, v, t_rate, t_time,
                                        rng_vec, reactHeap, reactNode, reactTimes, reactInf,
                                        next_unit_of_time, &fired);
            oracle_transitions_count += fired;
        } else {
            long long fired = 0;
            const int T = oracle_threads;
#pragma omp parallel num_threads(T) reduction(+:fired)
            {
                int my = 0;
#pragma omp for schedule(static)
                for (int node = 0; node < Nn; node++)
                    if (!my)
                        my = oracle_aem_node(a, node, Nc, Nt, Nd, Nld, u, v, t_rate, t_time,
                                             rng_vec, reactHeap, reactNode, reactTimes,
                                             reactInf, next_unit_of_time, &fired);
#pragma omp critical
                if (my && !error)
                    error = my;
            }
            oracle_transitions_count += fired;
        }

        if (!error)
            error = oracle_process_events(c, 0, tt, u, update_node, node_rng,
                                          &cursor1, &cursor2);

        if (oracle_threads <= 1) {
            for (int node = 0; node < Nn; node++) {
                const int rc = oracle_aem_pts_node(a, node, Nc, Nt, Nd, Nld, u, v, v_new,
                                                   t_rate, t_time, rng_vec, reactHeap,
                                                   reactNode, reactTimes, reactInf,
                                                   update_node, tt, &error);
                if (rc < 0) {
                    error = rc;
                    break;
                }
            }
        } else {
            const int T = oracle_threads;
#pragma omp parallel num_threads(T)
            {
                int my = 0;
#pragma omp for schedule(static)
                for (int node = 0; node < Nn; node++) {
                    if (my < 0 && my != SIMINF_ERR_INVALID_RATE)
                        continue;
                    const int rc = oracle_aem_pts_node(a, node, Nc, Nt, Nd, Nld, u, v, v_new,
                                                       t_rate, t_time, rng_vec, reactHeap,
                                                       reactNode, reactTimes, reactInf,
                                                       update_node, tt, &my);
                    if (rc < 0)
                        my = rc;
                }
#pragma omp critical
                if (my && !error)
                    error = my;
            }
        }

        tt = next_unit_of_time;
        next_unit_of_time += 1.0;

        oracle_store(c, 0, tt, u, v_new, &U_it, &V_it);
        {
            double *tmp = v;
            v = v_new;
            v_new = tmp;
        }
        if (U_it >= a->tlen)
            break;
    }

    free(u); free(v); free(v_new); free(t_rate); free(t_time);
    free(update_node); free(node_rng); free(rng_vec); free(reactHeap);
    free(reactNode); free(reactTimes); free(reactInf);
    return error;
}

/* ---------------------------------------------------------------- */
/* Entry points.                                                     */
/* ---------------------------------------------------------------- */

static int oracle_ctx_init(oracle_ctx *c, SimInf_solver_args *a)
{
    c->a = a;
    c->Nn = a->Nn;
    c->Nc = a->Nc;
    c->Nt = a->Nt;
    c->Nd = a->Nd;
    c->Nld = a->Nld;
    c->individuals = calloc((size_t)a->Nc + 1, sizeof(int));
    return c->individuals ? 0 : SIMINF_ERR_ALLOC_MEMORY_BUFFER;
}

int oracle_run_solver_ssm(SimInf_solver_args *a)
{
    oracle_ctx c;
    int err = oracle_ctx_init(&c, a);
    oracle_transitions_count = 0;
    if (!err)
        err = oracle_direct_replicate(&c, 0, 0);
    free(c.individuals);
    return err;
}

int oracle_run_solver_mssm(SimInf_solver_args *a)
{
    oracle_ctx c;
    int err = oracle_ctx_init(&c, a);
    oracle_transitions_count = 0;
    for (int rep = 0; rep < a->Nrep && !err; rep++)
        err = oracle_direct_replicate(&c, rep, 1);
    free(c.individuals);
    return err;
}

int oracle_run_solver_mssm_crn(SimInf_solver_args *a)
{
    return oracle_run_solver_mssm(a);
}

int oracle_run_solver_aem(SimInf_solver_args *a)
{
    oracle_ctx c;
    int err = oracle_ctx_init(&c, a);
    oracle_transitions_count = 0;
    oracle_aem_finite_ties = 0;
    if (!err)
        err = oracle_aem(&c);
    free(c.individuals);
    return err;
}

/* Dispatch of SimInf.c:383-399. */
int oracle_run(SimInf_solver_args *a)
{
    if (a->solver == 0) {
        if (a->Nrep > 1)
            return a->crn == 1 ? oracle_run_solver_mssm_crn(a)
                               : oracle_run_solver_mssm(a);
        return oracle_run_solver_ssm(a);
    }
    if (a->Nrep > 1)
        return SIMINF_ERR_AEM_REPLICATED_MODEL;
    return oracle_run_solver_aem(a);
}

long long oracle_last_transitions(void)
{
    return oracle_transitions_count;
}

long long oracle_aem_ties(void)
{
    return oracle_aem_finite_ties;
}

/* ---------------------------------------------------------------- */
/* Host plugins of the built-in models (the TRFun/PTSFun tables the   */
/* reference's <Model>_run functions pass to SimInf_run).             */
/* ---------------------------------------------------------------- */

static TRFun SIR_tr[] = {SIR_S_to_I, SIR_I_to_R};
static TRFun SEIR_tr[] = {SEIR_S_to_E, SEIR_E_to_I, SEIR_I_to_R};
static TRFun SIS_tr[] = {SIS_S_to_I, SIS_I_to_S};
static TRFun SISe_tr[] = {SISe_S_to_I, SISe_I_to_S};
static TRFun SISe3_tr[] = {SISe3_S_1_to_I_1, SISe3_I_1_to_S_1,
                           SISe3_S_2_to_I_2, SISe3_I_2_to_S_2,
                           SISe3_S_3_to_I_3, SISe3_I_3_to_S_3};
static TRFun SIR_mparse_tr[] = {SIR_mparse_trFun1, SIR_mparse_trFun2};

/* Fill tr_fun / pts_fun for a built-in model name; 0 if found. */
int oracle_set_model(SimInf_solver_args *a, const char *name)
{
    if (strcmp(name, "SIR") == 0) {
        a->tr_fun = SIR_tr;
        a->pts_fun = siminf_pts_noop;
    } else if (strcmp(name, "SEIR") == 0) {
        a->tr_fun = SEIR_tr;
        a->pts_fun = siminf_pts_noop;
    } else if (strcmp(name, "SIS") == 0) {
        a->tr_fun = SIS_tr;
        a->pts_fun = siminf_pts_noop;
    } else if (strcmp(name, "SISe") == 0) {
        a->tr_fun = SISe_tr;
        a->pts_fun = SISe_pts;
    } else if (strcmp(name, "SISe3") == 0) {
        a->tr_fun = SISe3_tr;
        a->pts_fun = SISe3_pts;
    } else if (strcmp(name, "SISe_sp") == 0) {
        a->tr_fun = SISe_tr;
        a->pts_fun = SISe_sp_pts;
    } else if (strcmp(name, "SISe3_sp") == 0) {
        a->tr_fun = SISe3_tr;
        a->pts_fun = SISe3_sp_pts;
    } else if (strcmp(name, "SIR_mparse") == 0) {
        a->tr_fun = SIR_mparse_tr;
        a->pts_fun = siminf_pts_noop;
    } else {
        return -1;
    }
    return 0;
}

/* ---------------------------------------------------------------- */
/* Primitive probes for the tests.                                   */
/* ---------------------------------------------------------------- */

void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    siminf_philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
}

void oracle_uniforms(unsigned long long seed, unsigned domain,
                     unsigned long long id, int count, double *out)
{
    siminf_stream s = siminf_stream_make(seed, domain, id);
    for (int i = 0; i < count; i++)
        out[i] = siminf_uniform(&s);
}

void oracle_log(const double *x, double *y, long long count)
{
    for (long long i = 0; i < count; i++)
        y[i] = siminf_log(x[i]);
}

void oracle_sqrt(const double *x, double *y, long long count)
{
    for (long long i = 0; i < count; i++)
        y[i] = siminf_sqrt(x[i]);
}

void oracle_binomial(unsigned long long seed, double p, int n, int count,
                     int *out)
{
    siminf_stream s = siminf_stream_make(seed, SIMINF_DOMAIN_E2, 7);
    for (int i = 0; i < count; i++)
        out[i] = siminf_binomial(&s, p, n);
}

void oracle_hypergeometric(unsigned long long seed, int n1, int n2, int t,
                           int count, int *out)
{
    siminf_stream s = siminf_stream_make(seed, SIMINF_DOMAIN_E2, 11);
    for (int i = 0; i < count; i++)
        out[i] = siminf_hypergeometric(&s, n1, n2, t);
}

/* Streams whose first npre uniforms were computed ahead must give the
 * same variates and leave the stream at the same position as streams
 * that compute every uniform themselves.  Runs count trials over a
 * spread of urns (small and large, complement branch, urns that run out
 * of good items) followed by a binomial and three plain uniforms, and
 * returns the number of trials that differ. */
int oracle_prefetch_mismatches(unsigned long long seed, int npre, int count)
{
    double pre[64];
    int bad = 0;
    uint32_t x = 12345u;
    if (npre > 64 || npre % 2)
        return -1;
    for (int i = 0; i < count; i++) {
        siminf_stream a = siminf_stream_make(seed, SIMINF_DOMAIN_E2, (uint64_t)i);
        siminf_stream b = a;
        int n1, n2, t, ka, kb, ba, bb;
        x = x * 1664525u + 1013904223u;
        n1 = (int)(x >> 8) % 40;
        x = x * 1664525u + 1013904223u;
        n2 = (int)(x >> 8) % 40;
        x = x * 1664525u + 1013904223u;
        t = (int)(x >> 8) % (n1 + n2 + 2);
        for (int q = 0; q < npre / 2; q++)
            siminf_uniform_pair(&b, (uint64_t)q, &pre[2 * q], &pre[2 * q + 1]);
        b.pre = pre;
        b.npre = (uint64_t)npre;
        ka = siminf_hypergeometric(&a, n1, n2, t);
        kb = siminf_hypergeometric(&b, n1, n2, t);
        ba = siminf_binomial(&a, 0.3, n1 + 3);
        bb = siminf_binomial(&b, 0.3, n1 + 3);
        if (ka != kb || ba != bb || a.j != b.j) {
            bad++;
            continue;
        }
        for (int r = 0; r < 3; r++)
            if (siminf_uniform(&a) != siminf_uniform(&b))
                bad++;
    }
    return bad;
}

/* sample_select / sample_select_enter on one node's state, repeated
 * count times from one stream; out is count x Nc. */
int oracle_sample(int enter, const int *irE, const int *jcE, const double *prE,
                  int Nc, const int *u_node, int select, int n,
                  double proportion, unsigned long long seed, int count,
                  int *out)
{
    siminf_stream s = siminf_stream_make(seed, SIMINF_DOMAIN_E2, 3);
    for (int i = 0; i < count; i++) {
        const int err = enter ?
            oracle_sample_select_enter(irE, jcE, prE, Nc, u_node, select, n,
                                       proportion, &out[i * Nc], &s) :
            oracle_sample_select(irE, jcE, prE, Nc, u_node, select, n,
                                 proportion, &out[i * Nc], &s);
        if (err)
            return err;
    }
    return 0;
}

/* ---------------------------------------------------------------- */
/* Bootstrap particle filter (R/pfilter.R:315-499).                   */
/* ---------------------------------------------------------------- */

/* The device filter's computation on the host: per interval one mssm run
 * of all particles from their ancestors' states over the interval's
 * events (pfilter.R:373-430), the weights (the sum of the observations'
 * log-probabilities over the observed nodes, added in the device's
 * order: node k into partial sum k mod 64, then the partial sums in
 * order), log-sum-exp, ESS and systematic resampling (pfilter.R:455-474).  a describes one particle, with tr_fun/pts_fun set.  Arguments as
 * SimInf_hip_pfilter. */
int oracle_pfilter(const SimInf_solver_args *a, int P, int n_int,
                   const double *tspan2, int obs_c, int obs_dist, double obs_p1,
                   double obs_p2, int obs_nodes,
                   const int *obs, unsigned long long seed, double *loglik,
                   double *ess, int *ancestors)
{
    const int Nn = a->Nn, Nc = a->Nc, Nd = a->Nd;
    const size_t lanes = (size_t)Nn * P;
    int *u_cur = malloc(lanes * Nc * sizeof(int));
    double *v_cur = malloc(lanes * (Nd > 0 ? Nd : 1) * sizeof(double));
    int *U = malloc(lanes * Nc * 2 * sizeof(int));
    double *V = malloc(lanes * (Nd > 0 ? Nd : 1) * 2 * sizeof(double));
    double *w = malloc((size_t)P * sizeof(double));
    int *anc = malloc((size_t)P * sizeof(int));
    int err = 0;
    *loglik = 0.0;
    if (!u_cur || !v_cur || !U || !V || !w || !anc) {
        err = SIMINF_ERR_ALLOC_MEMORY_BUFFER;
        goto done;
    }
    for (int p = 0; p < P; p++) {
        memcpy(&u_cur[(size_t)p * Nn * Nc], a->u0, (size_t)Nn * Nc * sizeof(int));
        if (Nd > 0)
            memcpy(&v_cur[(size_t)p * Nn * Nd], a->v0, (size_t)Nn * Nd * sizeof(double));
    }
    /* The events of each interval (SimInf_split_events,
     * SimInf_particle_filter.c:80-118): those due after the previous
     * interval's end and at or before this one's, in list order. */
    int ev_first = 0;
    for (int i = 0; i < n_int && !err; i++) {
        SimInf_solver_args b = *a;
        const double tb = tspan2[2 * i];
        const int tlen = isnan(tb) ? 1 : 2;
        const int t_end = (int)tspan2[2 * i + 1];
        int ev_n = 0;
        while (ev_first + ev_n < a->len && a->time[ev_first + ev_n] <= t_end)
            ev_n++;
        b.len = ev_n;
        if (a->len > 0) {
            b.event = a->event + ev_first;
            b.time = a->time + ev_first;
            b.node = a->node + ev_first;
            b.dest = a->dest + ev_first;
            b.n = a->n + ev_first;
            b.proportion = a->proportion + ev_first;
            b.select = a->select + ev_first;
            b.shift = a->shift + ev_first;
        }
        g_ev_total = a->len;
        g_ev_offset = ev_first;
        ev_first += ev_n;
        b.Nrep = P;
        b.solver = 0;
        b.tspan = isnan(tb) ? &tspan2[2 * i + 1] : &tspan2[2 * i];
        b.tlen = tlen;
        b.u0 = u_cur;
        b.v0 = v_cur;
        b.U = U;
        b.V = V;
        b.seed = siminf_interval_seed(seed, i);
        err = oracle_run(&b);
        g_ev_total = -1;
        g_ev_offset = 0;
        if (err)
            break;
        for (int p = 0; p < P; p++) {
            const size_t col = ((size_t)p * tlen + (tlen - 1)) * Nn;
            if (obs_nodes == 1) {
                long long tot = 0;
                for (int k = 0; k < Nn; k++)
                    tot += U[(col + k) * Nc + obs_c];
                w[p] = obs[(size_t)i] < 0 ? 0.0
                                          : siminf_obs_logw(obs_dist, obs[(size_t)i], tot,
                                                            obs_p1, obs_p2);
            } else {
                double partial[64], sum = 0.0;
                for (int l = 0; l < 64; l++)
                    partial[l] = 0.0;
                for (int k = 0; k < Nn; k++) {
                    const int x = obs[(size_t)i * Nn + k];
                    if (x >= 0)
                        partial[k % 64] = partial[k % 64] +
                            siminf_obs_logw(obs_dist, x, U[(col + k) * Nc + obs_c],
                                            obs_p1, obs_p2);
                }
                for (int l = 0; l < 64; l++)
                    sum = sum + partial[l];
                w[p] = sum;
            }
        }
        {
            double max_w = -INFINITY, sum = 0.0, s2 = 0.0;
            for (int p = 0; p < P; p++) {
                if (!(w[p] > -INFINITY && w[p] < INFINITY)) {
                    err = SIMINF_ERR_INVALID_OBS_PROCESS;
                    break;
                }
                max_w = w[p] > max_w ? w[p] : max_w;
            }
            if (err)
                break;
            for (int p = 0; p < P; p++)
                w[p] = siminf_exp(w[p] - max_w);
            for (int p = 0; p < P; p++)
                sum = sum + w[p];
            *loglik = ((*loglik + max_w) + siminf_log(sum)) - siminf_log((double)P);
            for (int p = 0; p < P; p++)
                w[p] = w[p] / sum;
            for (int p = 0; p < P; p++)
                s2 = s2 + w[p] * w[p];
            ess[i] = 1.0 / s2;
            {
                const siminf_stream st = siminf_stream_make(seed, SIMINF_DOMAIN_PF,
                                                            (uint64_t)i);
                const int r = siminf_systematic_resampling(w, P, siminf_uniform_at(&st, 0),
                                                           anc);
                if (r) {
                    err = r == -1 ? SIMINF_ERR_INVALID_WEIGHT : SIMINF_ERR_WEIGHT_SUM;
                    break;
                }
            }
        }
        if (ancestors)
            memcpy(&ancestors[(size_t)i * P], anc, (size_t)P * sizeof(int));
        for (int p = 0; p < P; p++) {
            const size_t src = ((size_t)anc[p] * tlen + (tlen - 1)) * Nn;
            memcpy(&u_cur[(size_t)p * Nn * Nc], &U[src * Nc], (size_t)Nn * Nc * sizeof(int));
            if (Nd > 0)
                memcpy(&v_cur[(size_t)p * Nn * Nd], &V[src * Nd],
                       (size_t)Nn * Nd * sizeof(double));
        }
    }
done:
    free(u_cur);
    free(v_cur);
    free(U);
    free(V);
    free(w);
    free(anc);
    return err;
}

/* siminf_exp, siminf_lfact and siminf_log_dbinom over arrays, for the
 * accuracy tests. */
void oracle_exp(const double *x, double *y, long long n)
{
    for (long long i = 0; i < n; i++)
        y[i] = siminf_exp(x[i]);
}

void oracle_lfact(const int *n, double *y, long long count)
{
    for (long long i = 0; i < count; i++)
        y[i] = siminf_lfact(n[i]);
}

void oracle_log_dpois(const int *x, const double *lambda, double *y, long long count)
{
    for (long long i = 0; i < count; i++)
        y[i] = siminf_log_dpois(x[i], lambda[i]);
}

void oracle_log_dbinom(const int *x, const int *n, double p, double *y, long long count)
{
    for (long long i = 0; i < count; i++)
        y[i] = siminf_log_dbinom(x[i], n[i], p);
}

/* Systematic resampling with a given uniform; returns the error code of
 * siminf_systematic_resampling. */
int oracle_resample(const double *w, int n, double u01, int *idx)
{
    return siminf_systematic_resampling(w, n, u01, idx);
}

/* ABC-SMC weights (SimInf_abc_weights, SimInf_abc.c:283-406), in the
 * reference's loop order: per current particle i, the prior terms, then
 * the kernel sum over the previous particles j in order; the maximum
 * starts at 0; normalisation in particle order.  The proposal density is
 * evaluated on points transformed by L^-1 (siminf_abc.h).  Arguments and
 * error codes as SimInf_hip_abc_weights. */
int oracle_abc_weights(const char *distribution, const double *p1, const double *p2,
                       int n, int d, const double *x, const double *xx, const double *w,
                       const double *sigma, double *ww)
{
    double L[SIMINF_ABC_MAX_PARS * SIMINF_ABC_MAX_PARS], lg[SIMINF_ABC_MAX_PARS];
    double lognorm = 0.0, max_ww = 0.0, sum;
    double *zc = NULL, *zp = NULL;
    int err = 0;
    if (!w) {
        for (int i = 0; i < n; i++)
            ww[i] = 1.0 / (double)n;
        return 0;
    }
    for (int k = 0; k < d; k++) {
        const int c = distribution[k];
        if (c != SIMINF_PRIOR_GAMMA && c != SIMINF_PRIOR_LOGNORMAL &&
            c != SIMINF_PRIOR_NORMAL && c != SIMINF_PRIOR_UNIFORM)
            return SIMINF_ERR_UNKNOWN_DISTRIBUTION;
        lg[k] = c == SIMINF_PRIOR_GAMMA ? lgamma(p1[k]) : 0.0;
    }
    if (siminf_cholesky(sigma, d, L, &lognorm))
        return SIMINF_ERR_ABC_WEIGHTS;
    zc = malloc((size_t)n * d * sizeof(double));
    zp = malloc((size_t)n * d * sizeof(double));
    if (!zc || !zp) {
        err = SIMINF_ERR_ALLOC_MEMORY_BUFFER;
        goto done;
    }
    for (int i = 0; i < n; i++) {
        siminf_forward_subst(xx + i, n, L, d, zc + i, n);
        siminf_forward_subst(x + i, n, L, d, zp + i, n);
    }
    for (int i = 0; i < n; i++) {
        double lp = 0.0;
        for (int k = 0; k < d; k++)
            lp = lp + siminf_prior_logdens(distribution[k], xx[(size_t)k * n + i],
                                           x[(size_t)k * n + i], p1[k], p2[k], lg[k]);
        if (!(lp > -INFINITY && lp < INFINITY)) {
            err = SIMINF_ERR_INVALID_WEIGHT;
            goto done;
        }
        sum = 0.0;
        for (int j = 0; j < n; j++)
            sum = sum + w[j] * siminf_exp(siminf_mvn_logpdf_z(zc + i, n, zp + j, n, d,
                                                              lognorm));
        if (!(sum > 0.0 && sum < INFINITY)) {
            err = SIMINF_ERR_ABC_WEIGHTS;
            goto done;
        }
        ww[i] = lp - siminf_log(sum);
        if (ww[i] > max_ww)
            max_ww = ww[i];
    }
    sum = 0.0;
    for (int i = 0; i < n; i++) {
        ww[i] = siminf_exp(ww[i] - max_ww);
        sum = sum + ww[i];
    }
    for (int i = 0; i < n; i++)
        ww[i] = ww[i] / sum;
done:
    free(zc);
    free(zp);
    return err;
}
