/*
 * siminf_hip.hip -- kernels, device plans and the C ABI of the MI355X
 * solver core (declarations in include/siminf_hip.h).
 *
 * Kernels
 *   k_classify<M>   one lane per (replicate, node): initial rates, nodes
 *                   frozen from the start to the fill kernel, the others
 *                   to the active list.
 *   k_ssm_nodes_free<M>
 *                   ssm/mssm when no external-transfer (E2) event couples
 *                   the nodes, all days in one launch, each lane stepping
 *                   through its own days (SimInf_solver_ssm.c:74-373,
 *                   SimInf_solver_mssm.c:81-402).
 *   k_aem_nodes_free<M>
 *                   the same for the all-events method, a quad of lanes
 *                   per node where the model allows
 *                   (SimInf_solver_aem.c:93-459).
 *   k_fill_frozen   the columns of frozen nodes.
 *   k_coupled<M>    one lane per node, barrier-separated days, for models
 *                   with E2 events: one replicate over a cooperative grid
 *                   or one workgroup per replicate; each day's E2 events
 *                   run level by level on an LDS gather of the herds they
 *                   touch; the levels are conflict-free sets that keep
 *                   each node's events in list order, so no atomics are
 *                   needed (SimInf_solver.c:824-902 run serially in the
 *                   reference, SimInf_solver_ssm.c:244-250).
 */
#include "siminf_kernels.h"
#include "siminf_pfilter.h"
#include "siminf_abc.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

using namespace siminf;

/* ------------------------------------------------------------------ */
/* Built-in models as compile-time plugins.                            */
/* ------------------------------------------------------------------ */


/* gdata lengths: SISe.R:177-178 (8 values), SISe3.R:188-190 (12), the
 * mparse SIR (beta, gamma); SIR, SEIR and SIS read ldata only.  S and G
 * are the constructors' matrices (SIR.R:127-133, SEIR.R:136-143,
 * SIS.R:122-128, SISe.R:146-151, SISe3.R:162-178).  Each record is
 * defined in the model's own translation unit (builtin_*.hip). */
const SimInf_hip_model *siminf_builtin_SIR();
const SimInf_hip_model *siminf_builtin_SEIR();
const SimInf_hip_model *siminf_builtin_SIS();
const SimInf_hip_model *siminf_builtin_SISe();
const SimInf_hip_model *siminf_builtin_SISe3();
const SimInf_hip_model *siminf_builtin_SISe_sp();
const SimInf_hip_model *siminf_builtin_SISe3_sp();
const SimInf_hip_model *siminf_builtin_SIR_mparse();

static const SimInf_hip_model *(*const builtin_models[])() = {
    siminf_builtin_SIR, siminf_builtin_SEIR, siminf_builtin_SIS,
    siminf_builtin_SISe, siminf_builtin_SISe3, siminf_builtin_SISe_sp,
    siminf_builtin_SISe3_sp, siminf_builtin_SIR_mparse,
};

extern "C" const SimInf_hip_model *SimInf_hip_model_builtin(const char *name)
{
    if (!name)
        return nullptr;
    for (auto get : builtin_models) {
        const SimInf_hip_model *m = get();
        if (std::strcmp(m->name, name) == 0)
            return m;
    }
    return nullptr;
}

extern "C" int SimInf_hip_model_dims(const SimInf_hip_model *model, int *Nc,
                                     int *Nt, int *Nd)
{
    if (!model || model->abi != SIMINF_HIP_PLUGIN_ABI ||
        model->kparams_sig != SIMINF_KPARAMS_SIG)
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

enum Kernel { K_SSM_NODES, K_AEM_NODES, K_COUPLED_GRID, K_COUPLED_REPS };

/* Freed pool memory the process keeps for reuse (per device). */
constexpr uint64_t SIMINF_POOL_KEEP = 16ull << 30;

/* Stream-ordered allocation where the device supports memory pools:
 * hipFree waits for the whole device, which would serialise plans that
 * run on their own streams at the same time (concurrent filter
 * requests); hipFreeAsync orders the free on the plan's stream only. */
static bool pool_allocation(int dev)
{
    static std::mutex mu;
    static signed char ready[64];
    if (dev < 0 || dev >= 64)
        return false;
    std::lock_guard<std::mutex> lock(mu);
    if (ready[dev] == 0) {
        int ok = 0;
        hipMemPool_t pool = nullptr;
        uint64_t keep = SIMINF_POOL_KEEP;
        ready[dev] = -1;
        if (hipDeviceGetAttribute(&ok, hipDeviceAttributeMemoryPoolsSupported, dev) ==
                hipSuccess && ok &&
            hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess &&
            hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) ==
                hipSuccess)
            ready[dev] = 1;
    }
    return ready[dev] == 1;
}

struct DeviceBuffers {
    std::vector<void *> ptrs;
    hipStream_t stream = nullptr;   /* copies (and pooled allocations) go here */
    bool pooled = false;
    hipError_t status = hipSuccess;

    template <class T>
    T *alloc(size_t count)
    {
        void *p = nullptr;
        if (status != hipSuccess)
            return nullptr;
        const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
        status = pooled ? hipMallocAsync(&p, bytes, stream) : hipMalloc(&p, bytes);
        if (status != hipSuccess)
            return nullptr;
        ptrs.push_back(p);
        return (T *)p;
    }

    /* A copy on the buffers' stream, complete on return (the host side may
     * be a temporary). */
    void copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
    {
        if (status != hipSuccess || bytes == 0)
            return;
        status = hipMemcpyAsync(dst, src, bytes, kind, stream);
        if (status == hipSuccess)
            status = hipStreamSynchronize(stream);
    }

    template <class T>
    T *upload(const T *host, size_t count)
    {
        T *d = alloc<T>(count);
        if (d && count && host)
            copy(d, host, count * sizeof(T), hipMemcpyHostToDevice);
        return d;
    }

    /* Frees one buffer of these. */
    void free_one(void *p)
    {
        for (size_t k = 0; k < ptrs.size(); k++) {
            if (ptrs[k] == p) {
                (void)(pooled ? hipFreeAsync(p, stream) : hipFree(p));
                ptrs.erase(ptrs.begin() + (long)k);
                return;
            }
        }
    }

    /* Frees everything; the stream must still exist. */
    void release()
    {
        for (void *p : ptrs)
            (void)(pooled ? hipFreeAsync(p, stream) : hipFree(p));
        if (pooled && !ptrs.empty())
            (void)hipStreamSynchronize(stream);
        ptrs.clear();
    }

    ~DeviceBuffers() { release(); }
};

} // namespace

/* Herds a grid-wide coupled plan can give a wavefront of their own. */
#define SIMINF_SOLO_MAX 16
/* The fewest herds a grid-wide coupled plan packs on one wavefront. */
#define SIMINF_PACK_MIN 8

struct SimInf_hip_plan;
static int plan_set_herd_map(SimInf_hip_plan *P, const int *nodes, int n, int pack);

struct SimInf_hip_plan {
    const SimInf_hip_model *model;
    Kernel kernel;
    KParams p;
    DeviceBuffers buf;
    hipStream_t stream = nullptr;
    hipStream_t fill_stream = nullptr;   /* the fill kernel, beside the active pass */
    hipEvent_t ev_classified = nullptr, ev_fill0 = nullptr, ev_fill1 = nullptr,
               ev_active1 = nullptr;
    /* The fill stream's fork from and join to the plan's stream. */
    hipEvent_t dep_fork = nullptr, dep_join = nullptr;
    int simds = 1024;                     /* SIMDs of the device (4 per CU) */
    hipEvent_t ev_start = nullptr, ev_mid = nullptr, ev_stop = nullptr;
    unsigned grid = 0, block = 0;
    unsigned grid_active = 0;             /* node kernels' active pass */
    size_t lds = 0;
    bool has_e1 = false;
    bool canon = false;

    /* Host copies needed after runs. */
    int Nn, Nc, Nd, Nrep, tlen;
    bool dense_U, dense_V;
    /* Fused prevalence (SimInf_hip_plan_fuse_prevalence): the sums and
     * difference array, and whether V was dense before U and V were freed. */
    long long *fp_sum = nullptr, *fp_delta = nullptr;
    bool fp_had_V = false;
    size_t nnzU, nnzV;
    int ndays;
    uint64_t host_err_key;     /* errors found while validating events */
    /* A run's result words on the device (error key, transition count,
     * fill node-columns: p.err_key, p.transitions, p.frozen_cols point
     * into them) and their pinned host copy, fetched with one copy. */
    unsigned long long *result_words = nullptr;
    unsigned long long *host_words = nullptr;
    /* The last run's first error key, and each day's first event in the
     * event list (-1 for days without events), to name its event. */
    uint64_t last_key = ~0ull;
    std::vector<int> ev_day_first;

    /* Lanes per wavefront of the active pass as a power of two, or -1
     * for the rule of about two wavefronts per SIMD. */
    int lpw_fixed = -1;
    /* Columns per fill block (the trajectory-flush schedule), 0 for the
     * default; tuned by the autotuners. */
    int fill_block = 0;
    /* Grid-wide coupled plans: herds given a wavefront of their own, the
     * other herds per wavefront, and the workgroups a grid launch may
     * hold. */
    int solo = 0, pack = SIMINF_GRID_BLOCK, grid_max = 0;
    /* Grid-wide coupled plans: the workgroups one cooperative launch may
     * hold, and the replicates a launch runs under the current map. */
    int max_blocks = 0, reps_per_launch = 1;

    /* Parameters per replicate (SimInf_hip_plan_set_replicate_data):
     * device buffers allocated on first use, and a count of the calls
     * that changed where the kernels read them (a batch recaptures its
     * graph when it changes). */
    double *rep_ldata = nullptr, *rep_gdata = nullptr;
    unsigned param_version = 0;
    int Ngdata = 0, Nld = 0;

    /* Last run. */
    bool has_run = false;       /* a run completed without error */
    unsigned long long transitions = 0;
    double kernel_ms = 0.0;     /* all kernels of the run */
    double solver_ms = 0.0;     /* the solver kernel(s) */
    double fill_ms = 0.0;       /* the frozen-column fill */
    unsigned long long fill_cols = 0;
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

/* LDS of the coupled kernel, about `budget` bytes at most: buffers for
 * one day's E2 work sized for the largest day or, when that does not fit,
 * the budget split 45 / 50 / 5 % over herds, event records with their
 * precomputed uniforms, and level bounds (larger days run from global
 * memory).  Sets the offsets in p and returns the byte size. */
static size_t coupled_lds_layout(KParams &p, int Nc, int Nselect, int Nshift,
                                 bool has_N, int max_t, int max_ev, int max_lv,
                                 size_t budget)
{
    auto align16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
    /* E and N first; when they take more than a quarter of the budget
     * every day runs from global memory. */
    size_t tab = 0;
    p.lds_Ew = (int)tab;
    tab = align16(tab + (size_t)std::max(Nselect, 1) * Nc * sizeof(double));
    p.lds_Emask = (int)tab;
    tab = align16(tab + (size_t)std::max(Nselect, 1) * sizeof(uint32_t));
    p.lds_N = (int)tab;
    if (has_N)
        tab = align16(tab + (size_t)std::max(Nshift, 1) * Nc * sizeof(int));
    if (tab > budget / 4) {
        p.e2_cap_t = p.e2_cap_ev = p.e2_cap_lv = 0;
        p.lds_u = p.lds_upd = p.lds_tn = p.lds_rec = p.lds_prop = p.lds_lvl = p.lds_pre = 0;
        return 16;
    }
    budget -= tab;
    const size_t per_t = (size_t)(Nc + 2) * sizeof(int);
    const size_t per_ev = 2 * sizeof(int4) + (1 + SIMINF_E2_PRE) * sizeof(double);
    auto need = [&](size_t t, size_t ev, size_t lv) {
        return align16((t * Nc + 4) * sizeof(int)) + 2 * align16(t * sizeof(int)) +
               ev * 2 * sizeof(int4) + align16(ev * sizeof(double)) +
               align16((lv + 1) * sizeof(int)) +
               ev * SIMINF_E2_PRE * sizeof(double);
    };
    size_t cap_t = (size_t)max_t, cap_ev = (size_t)max_ev, cap_lv = (size_t)max_lv;
    if (need(cap_t, cap_ev, cap_lv) > budget) {
        const size_t rest = budget > 1024 ? budget - 256 : 768;
        const size_t lv_slots = rest * 5 / 100 / sizeof(int);
        cap_t = std::min(cap_t, rest * 45 / 100 / per_t);
        cap_ev = std::min(cap_ev, rest * 50 / 100 / per_ev);
        cap_lv = std::min(cap_lv, lv_slots > 1 ? lv_slots - 1 : 0);
    }
    p.e2_cap_t = (int)cap_t;
    p.e2_cap_ev = (int)cap_ev;
    p.e2_cap_lv = (int)cap_lv;
    size_t off = tab;
    /* The herds' counts, and four more ints that the one-wavefront level
     * loop's lanes without an event read and write. */
    p.lds_u = (int)off;
    off = align16(off + (cap_t * Nc + 4) * sizeof(int));
    p.lds_upd = (int)off;
    off = align16(off + cap_t * sizeof(int));
    p.lds_tn = (int)off;
    off = align16(off + cap_t * sizeof(int));
    p.lds_rec = (int)off;
    off = align16(off + cap_ev * 2 * sizeof(int4));
    p.lds_prop = (int)off;
    off = align16(off + cap_ev * sizeof(double));
    p.lds_lvl = (int)off;
    off = align16(off + (cap_lv + 1) * sizeof(int));
    p.lds_pre = (int)off;
    off = align16(off + cap_ev * SIMINF_E2_PRE * sizeof(double));
    return std::max<size_t>(off, 16);
}

/* filter_plan: u0/v0 hold one replicate (Nn columns), copied on the
 * device to all a->Nrep replicates, and the caller runs every lane in one
 * pass (no frozen-node buffers). */
/* A neighbour-reading model's lists in ldata (Nld rows, ncol columns):
 * every column's (index, value) pairs from row `row` must end with an
 * index below zero inside the column and name nodes 0..Nn-1 (the index
 * converted to int as the plugin converts it), or the post-time-step
 * function would read outside the state arrays. */
static bool neighbor_lists_ok(const double *ldata, int Nld, long long ncol, int row, int Nn)
{
    for (long long j = 0; j < ncol; j++) {
        const double *col = ldata + j * (long long)Nld;
        int r = row;
        for (;; r += 2) {
            if (r >= Nld || !(col[r] > -2147483648.0 && col[r] < 2147483648.0))
                return false;
            const int idx = (int)col[r];
            if (idx < 0)
                break;
            if (idx >= Nn || r + 1 >= Nld)
                return false;
        }
    }
    return true;
}

static int plan_create_impl(SimInf_hip_plan **out, const SimInf_solver_args *a,
                            const SimInf_hip_model *model, bool filter_plan)
{
    if (!out || !a || !model)
        return SIMINF_ERR_INVALID_MODEL;
    *out = nullptr;

    if (model->abi != SIMINF_HIP_PLUGIN_ABI || model->kparams_sig != SIMINF_KPARAMS_SIG)
        return SIMINF_ERR_INVALID_MODEL;
    const int Nn = a->Nn, Nc = a->Nc, Nt = a->Nt, Nd = a->Nd;
    const int Nrep = a->Nrep < 1 ? 1 : a->Nrep;
    if (Nn <= 0 || Nc != model->Nc || Nt != model->Nt || Nd != model->Nd ||
        a->tlen <= 0 || a->Nld < 0 || a->len < 0)
        return SIMINF_ERR_INVALID_MODEL;
    /* S, G, E and the summary masks are 32-bit sets of compartments,
     * transitions and continuous states. */
    if (Nc > SIMINF_MAX_SET || Nt > SIMINF_MAX_SET || Nd > SIMINF_MAX_SET)
        return SIMINF_ERR_INVALID_MODEL;
    if (a->solver != 0 && a->solver != 1)
        return SIMINF_ERR_UNKNOWN_SOLVER;
    if (a->solver == 1 && Nrep > 1)
        return SIMINF_ERR_AEM_REPLICATED_MODEL;
    if (!a->U && (!a->irU || !a->jcU || !a->prU))
        return SIMINF_ERR_SPARSE_MODEL;
    if (!a->V && (!a->irV || !a->jcV || !a->prV))
        return SIMINF_ERR_SPARSE_MODEL;
    if (model->neighbors > 0 &&
        (!a->ldata || !neighbor_lists_ok(a->ldata, a->Nld, Nn, model->neighbors - 1, Nn)))
        return SIMINF_ERR_INVALID_MODEL;

    SimInf_hip_plan *P = new SimInf_hip_plan();
    P->model = model;
    P->Nn = Nn;
    P->Nc = Nc;
    P->Nd = Nd;
    P->Nrep = Nrep;
    P->tlen = a->tlen;
    P->dense_U = a->U != nullptr;
    P->dense_V = a->V != nullptr;
    P->Ngdata = model->Ngdata;
    P->Nld = a->Nld;
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
        if (ev_day[k] >= 0) {
            if ((size_t)ev_day[k] >= P->ev_day_first.size())
                P->ev_day_first.resize((size_t)ev_day[k] + 1, -1);
            if (P->ev_day_first[ev_day[k]] < 0)
                P->ev_day_first[ev_day[k]] = k;
        }
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
    const int n_e1 = e1_ptr[Nn];
    std::vector<int4> e1_rec((size_t)std::max(n_e1, 1) * 2, int4{0, 0, 0, 0});
    std::vector<double> e1_prop(std::max(n_e1, 1), 0.0);
    for (int j = 0; j < n_e1; j++) {
        const int k = e1_idx[j];
        e1_rec[2 * (size_t)j] = int4{a->time[k], a->event[k], a->n[k], ev_select[k]};
        e1_rec[2 * (size_t)j + 1] = int4{ev_shift[k], k, 0, 0};
        e1_prop[j] = a->proportion[k];
    }

    /* E2 levels: an event's level is one more than the last level that
     * touched its source or destination on the same day. */
    std::vector<int> e2_day(ndays + 1, 0), e2_lvl(1, 0), e2_idx;
    std::vector<int> e2_tn_ptr(ndays + 1, 0), e2_tn, e2_lsrc(std::max(len, 1), -1),
        e2_ldst(std::max(len, 1), -1);
    int n_e2 = 0, max_touched = 0, max_day_events = 0, max_day_levels = 0;
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
            /* Herds touched today, and each event's slots among them. */
            const int t0 = (int)e2_tn.size();
            for (auto &lv : levels) {
                for (int ev : lv) {
                    for (int side = 0; side < 2; side++) {
                        const int nd = side ? ev_dest[ev] : ev_node[ev];
                        int &slot = side ? e2_ldst[ev] : e2_lsrc[ev];
                        if (nd < 0 || nd >= Nn)
                            continue;
                        if (last_day[nd] != d + ndays + 1) {
                            last_day[nd] = d + ndays + 1;
                            last_level[nd] = (int)e2_tn.size() - t0;
                            e2_tn.push_back(nd);
                        }
                        slot = last_level[nd];
                    }
                }
            }
            e2_tn_ptr[d + 1] = (int)e2_tn.size();
            max_touched = std::max(max_touched, (int)e2_tn.size() - t0);
            const int i_start = (int)e2_idx.size();
            /* A level wider than a wavefront is split into levels of at
             * most 64 events (its events are independent). */
            int day_levels = 0;
            for (auto &lv : levels) {
                for (size_t j = 0; j < lv.size(); j += SIMINF_GRID_BLOCK) {
                    const size_t e = std::min(lv.size(), j + SIMINF_GRID_BLOCK);
                    e2_idx.insert(e2_idx.end(), lv.begin() + j, lv.begin() + e);
                    e2_lvl.push_back((int)e2_idx.size());
                    day_levels++;
                }
            }
            e2_day[d + 1] = (int)e2_lvl.size() - 1;
            max_day_events = std::max(max_day_events, (int)e2_idx.size() - i_start);
            max_day_levels = std::max(max_day_levels, day_levels);
        }
        if (e2_tn.empty())
            e2_tn.push_back(0);
    }
    std::vector<int4> e2_rec(std::max<size_t>(e2_idx.size(), 1) * 2, int4{0, 0, 0, 0});
    std::vector<double> e2_prop(std::max<size_t>(e2_idx.size(), 1), 0.0);
    /* Events the fast path takes (SIMINF_E2_FAST in the shift field):
     * a count, no shift, another herd as destination, and an E column of
     * one row or of two rows with equal weights. */
    auto e2_fast = [&](int k) {
        const int sel = ev_select[k];
        if (a->n[k] <= 0 || ev_shift[k] >= 0 || ev_dest[k] < 0 || ev_dest[k] >= Nn ||
            ev_dest[k] == ev_node[k] || sel < 0 || sel >= Nselect)
            return false;
        const uint32_t m = Emask[sel];
        const int rows = __builtin_popcount(m);
        if (rows == 1)
            return true;
        if (rows != 2)
            return false;
        const int r0 = __builtin_ctz(m), r1 = __builtin_ctz(m & (m - 1u));
        return Ew[(size_t)sel * Nc + r0] == Ew[(size_t)sel * Nc + r1];
    };
    /* A fast event carries its E column's rows in place of the select:
     * first row | second row << 8 | (two rows) << 16. */
    auto e2_rows = [&](int k) {
        const uint32_t m = Emask[ev_select[k]];
        const int r0 = __builtin_ctz(m);
        const uint32_t rest = m & (m - 1u);
        return rest ? (r0 | (__builtin_ctz(rest) << 8) | (1 << 16)) : (r0 | (r0 << 8));
    };
    for (size_t i = 0; i < e2_idx.size(); i++) {
        const int k = e2_idx[i];
        const bool fast = e2_fast(k);
        e2_rec[2 * i] = int4{k, a->n[k], fast ? e2_rows(k) : ev_select[k],
                             fast ? SIMINF_E2_FAST : ev_shift[k]};
        e2_rec[2 * i + 1] = int4{e2_lsrc[k], e2_ldst[k], ev_node[k], ev_dest[k]};
        e2_prop[i] = a->proportion[k];
    }

    /* The fast kernels fold the model's own S and G into the code; any
     * other S or G (a caller may pass modified matrices) runs the
     * variant that reads them from the arguments. */
    {
        bool same = true;
        for (int k = 0; k < Nt && same; k++) {
            for (int c = 0; c < Nc; c++) {
                const bool stored = (Smask[k] >> c) & 1u;
                const int v = stored ? Sval[(size_t)k * Nc + c] : 0;
                if (stored != (model->S[k * Nc + c] != 0) || v != model->S[k * Nc + c])
                    same = false;
            }
            for (int j = 0; j < Nt; j++)
                if ((((Gmask[k] >> j) & 1u) != 0) != (model->G[k * Nt + j] != 0))
                    same = false;
        }
        P->canon = same;
    }

    /* Kernel choice.  External transfers couple the nodes, and so does a
     * post-time-step function that reads other nodes' state. */
    const bool coupled = n_e2 > 0 || model->neighbors;
    if (a->solver == 1) {
        /* aem runs one replicate; with coupled nodes it runs in the
         * grid-wide coupled kernel. */
        P->kernel = coupled ? K_COUPLED_GRID : K_AEM_NODES;
    } else if (!coupled) {
        P->kernel = K_SSM_NODES;
    } else {
        P->kernel = (Nrep == 1 || Nn > SIMINF_REP_BLOCK) ? K_COUPLED_GRID
                                                         : K_COUPLED_REPS;
    }

    /* Streams first: every copy of the plan, and its allocations where the
     * device has memory pools, are ordered on its own stream, which does
     * not wait for the legacy default stream. */
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&P->fill_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&P->ev_classified) != hipSuccess ||
        hipEventCreate(&P->ev_fill0) != hipSuccess ||
        hipEventCreate(&P->ev_fill1) != hipSuccess ||
        hipEventCreate(&P->ev_active1) != hipSuccess ||
        hipEventCreate(&P->ev_start) != hipSuccess ||
        hipEventCreate(&P->ev_mid) != hipSuccess ||
        hipEventCreate(&P->ev_stop) != hipSuccess ||
        hipEventCreateWithFlags(&P->dep_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&P->dep_join, hipEventDisableTiming) != hipSuccess) {
        SimInf_hip_plan_destroy(P);
        return SIMINF_ERR_DEVICE;
    }

    /* Device buffers. */
    DeviceBuffers &B = P->buf;
    B.stream = P->stream;
    B.pooled = pool_allocation(dev);
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
    p.coupled_aem = a->solver == 1 && P->kernel == K_COUPLED_GRID;
    p.tt0 = tt0;
    p.next0 = std::floor(tt0) + 1.0;
    p.ndays = ndays;
    p.day0 = 0;
    p.ev_floor = INT_MIN;
    p.e2_floor = -1;

    if (!filter_plan) {
        p.u0 = B.upload(a->u0, (size_t)lanes * Nc);
        p.v0 = B.upload(a->v0, (size_t)lanes * Nd);
    } else {
        /* Upload one replicate, then double the filled part. */
        int *u0 = B.alloc<int>((size_t)lanes * Nc);
        double *v0 = B.alloc<double>((size_t)lanes * std::max(Nd, 1));
        const size_t nu = (size_t)Nn * Nc, nv = (size_t)Nn * Nd;
        B.copy(u0, a->u0, nu * sizeof(int), hipMemcpyHostToDevice);
        B.copy(v0, a->v0, nv * sizeof(double), hipMemcpyHostToDevice);
        for (long long have = 1; have < Nrep && B.status == hipSuccess; have *= 2) {
            const long long n = std::min<long long>(have, Nrep - have);
            B.copy(u0 + have * nu, u0, n * nu * sizeof(int), hipMemcpyDeviceToDevice);
            B.copy(v0 + have * nv, v0, n * nv * sizeof(double), hipMemcpyDeviceToDevice);
        }
        p.u0 = u0;
        p.v0 = v0;
    }
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

    p.ev_rank = B.upload(ev_rank.data(), (size_t)len);
    p.e1_ptr = B.upload(e1_ptr.data(), e1_ptr.size());
    p.e1_rec = B.upload(e1_rec.data(), e1_rec.size());
    p.e1_prop = B.upload(e1_prop.data(), e1_prop.size());
    p.e2_day = B.upload(e2_day.data(), e2_day.size());
    p.e2_lvl = B.upload(e2_lvl.data(), e2_lvl.size());
    p.e2_rec = B.upload(e2_rec.data(), e2_rec.size());
    p.e2_prop = B.upload(e2_prop.data(), e2_prop.size());
    p.e2_tn_ptr = B.upload(e2_tn_ptr.data(), e2_tn_ptr.size());
    p.e2_tn = B.upload(e2_tn.data(), e2_tn.size());
    {
        /* Each node's transfer days (the touched lists, by node). */
        std::vector<int> hptr((size_t)Nn + 1, 0), hday;
        for (size_t d = 0; d + 1 < e2_tn_ptr.size(); d++)
            for (int i = e2_tn_ptr[d]; i < e2_tn_ptr[d + 1]; i++)
                hptr[(size_t)e2_tn[i] + 1]++;
        for (int i = 0; i < Nn; i++)
            hptr[(size_t)i + 1] += hptr[i];
        hday.resize(std::max<size_t>((size_t)hptr[Nn], 1));
        std::vector<int> fill(hptr.begin(), hptr.end() - 1);
        for (size_t d = 0; d + 1 < e2_tn_ptr.size(); d++)
            for (int i = e2_tn_ptr[d]; i < e2_tn_ptr[d + 1]; i++)
                hday[(size_t)fill[e2_tn[i]]++] = (int)d;
        p.e2_hptr = B.upload(hptr.data(), hptr.size());
        p.e2_hday = B.upload(hday.data(), hday.size());
    }
    p.Nselect = Nselect;

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
    P->result_words = B.alloc<unsigned long long>(3);
    p.err_key = P->result_words;
    p.transitions = P->result_words + 1;
    if (hipHostMalloc((void **)&P->host_words, 3 * sizeof(unsigned long long),
                      hipHostMallocDefault) != hipSuccess) {
        P->host_words = nullptr;
        SimInf_hip_plan_destroy(P);
        return SIMINF_ERR_ALLOC_MEMORY_BUFFER;
    }

    if (P->kernel == K_COUPLED_GRID || P->kernel == K_COUPLED_REPS) {
        p.u_work = B.alloc<int>((size_t)lanes * Nc);
        p.upd_work = B.alloc<int>((size_t)lanes);
        if (model->neighbors)
            p.v_work = B.alloc<double>((size_t)lanes * std::max(Nd, 1));
        /* LDS budget: at most 128 KiB, and small enough that the
         * workgroups a grid-wide launch needs fit on the CUs together
         * (every workgroup gets the same allocation). */
        int lds_max = 64 * 1024, lds_cu = 64 * 1024, cus = 1;
        if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                  dev) != hipSuccess)
            lds_max = 64 * 1024;
        if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                  dev) != hipSuccess)
            lds_cu = lds_max;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess || cus < 1)
            cus = 1;
        size_t budget = std::min<size_t>((size_t)lds_max, 128 * 1024);
        if (P->kernel == K_COUPLED_GRID) {
            /* The herds' workgroups and the one that runs E2, per
             * replicate; replicates share a launch, up to 16 one-wavefront
             * workgroups per CU. */
            const long long blocks1 = (Nn + SIMINF_GRID_BLOCK - 1) / SIMINF_GRID_BLOCK + 1;
            const long long blocks = std::min<long long>(blocks1 * Nrep,
                                                         std::max<long long>(blocks1, 16LL * cus));
            const long long per_cu = (blocks + cus - 1) / cus;
            const size_t share = (size_t)lds_cu / (size_t)per_cu;
            budget = std::min<size_t>(budget, share > 2048 ? share - 1024 : 1024);
        }
        P->lds = coupled_lds_layout(p, Nc, Nselect, Nshift, a->N != nullptr,
                                    max_touched, max_day_events, max_day_levels,
                                    budget);
        /* Days whose E2 work fits the LDS buffers, with every event on the
         * fast path and under 128 levels, run on one wavefront; their
         * records' second words become the byte offsets of the four counts
         * an event reads in the staged herd state. */
        {
            std::vector<int> day_fast(e2_day.size(), 0);
            bool any = false;
            for (size_t d = 0; d + 1 < e2_day.size(); d++) {
                const int l0 = e2_day[d], nlv = e2_day[d + 1] - l0;
                if (nlv == 0 || nlv >= 128)
                    continue;
                const int i0 = e2_lvl[l0], nev = e2_lvl[l0 + nlv] - i0;
                const int nt = e2_tn_ptr[d + 1] - e2_tn_ptr[d];
                if (!(nt <= p.e2_cap_t && nev <= p.e2_cap_ev && nlv <= p.e2_cap_lv))
                    continue;
                bool fast = true;
                for (int i = i0; i < i0 + nev && fast; i++)
                    fast = e2_rec[2 * i].w == SIMINF_E2_FAST;
                if (!fast)
                    continue;
                for (int i = i0; i < i0 + nev; i++) {
                    const int rows = e2_rec[2 * i].z, ra = rows & 0xFF, rb = (rows >> 8) & 0xFF;
                    const int sl = e2_rec[2 * i + 1].x, dl = e2_rec[2 * i + 1].y;
                    e2_rec[2 * i + 1] = int4{(sl * Nc + ra) * 4, (sl * Nc + rb) * 4,
                                             (dl * Nc + ra) * 4, (dl * Nc + rb) * 4};
                }
                day_fast[d] = 1;
                any = true;
            }
            if (any)
                B.copy((void *)p.e2_rec, e2_rec.data(), e2_rec.size() * sizeof(int4),
                       hipMemcpyHostToDevice);
            p.e2_day_fast = B.upload(day_fast.data(), day_fast.size());
        }
        if (P->kernel == K_COUPLED_GRID) {
            P->block = SIMINF_GRID_BLOCK;
            P->grid = (unsigned)((Nn + SIMINF_GRID_BLOCK - 1) / SIMINF_GRID_BLOCK + 1);
            int max_blocks = 0;
            if (model->coupled(p, P->canon, true, P->grid, P->block, P->lds,
                               nullptr, &max_blocks) != hipSuccess ||
                (int)P->grid > max_blocks) {
                /* Too many nodes for one co-resident grid. */
                SimInf_hip_plan_destroy(P);
                return SIMINF_ERR_ALLOC_MEMORY_BUFFER;
            }
            /* Room for SIMINF_SOLO_MAX herds on wavefronts of their own
             * and the others packed as few as SIMINF_PACK_MIN to a
             * wavefront, within what one co-resident grid holds. */
            P->grid_max = std::min<int>(
                max_blocks, (Nn + SIMINF_PACK_MIN - 1) / SIMINF_PACK_MIN + SIMINF_SOLO_MAX + 1);
            P->max_blocks = max_blocks;
            p.lane_node = B.alloc<int>((size_t)P->grid_max * SIMINF_GRID_BLOCK);
            p.node_fired = B.alloc<unsigned long long>((size_t)Nn);
            /* Barrier words for the most replicates one launch can hold,
             * and each replicate's error key. */
            p.gsync = B.alloc<unsigned>(SIMINF_GSYNC_WORDS * (size_t)std::min(Nrep, max_blocks));
            if (Nrep > 1)
                p.rep_err = B.alloc<unsigned long long>((size_t)Nrep);
        } else {
            P->block = (unsigned)(((Nn + 63) / 64) * 64);
            P->grid = (unsigned)Nrep;
        }
    } else {
        P->block = 256;
        P->grid = (unsigned)((lanes + 255) / 256);
        if (a->U && (Nd == 0 || a->V) && !filter_plan) {
            p.frozen_from = B.alloc<int>((size_t)lanes);
            p.frozen_u = B.alloc<int>((size_t)lanes * Nc);
            p.frozen_v = B.alloc<double>((size_t)lanes * std::max(Nd, 1));
            p.frozen_cols = P->result_words + 2;
            p.active = B.alloc<int>((size_t)lanes);
            p.active_count = B.alloc<int>(1);
        }
    }

    if (B.status != hipSuccess) {
        SimInf_hip_plan_destroy(P);
        return SIMINF_ERR_ALLOC_MEMORY_BUFFER;
    }
    if (P->kernel == K_COUPLED_GRID) {
        const int err = plan_set_herd_map(P, nullptr, 0, SIMINF_GRID_BLOCK);
        if (err) {
            SimInf_hip_plan_destroy(P);
            return err;
        }
    }
    {
        int cus = 256;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) ==
                hipSuccess && cus > 0)
            P->simds = 4 * cus;
        p.simds = P->simds;
        /* The active pass's fixed grid: eight wavefronts per SIMD (what
         * the node kernels' register use lets the device hold, so that a
         * packing that needs up to that many gives every wave slot a
         * wavefront of its own), never more wavefronts than lanes. */
        const long long waves = std::min<long long>(lanes, 8LL * P->simds);
        P->grid_active = (unsigned)((waves * 64 + P->block - 1) / P->block);
    }
    *out = P;
    return 0;
}

extern "C" int SimInf_hip_plan_create(SimInf_hip_plan **out,
                                      const SimInf_solver_args *a,
                                      const SimInf_hip_model *model)
{
    return plan_create_impl(out, a, model, false);
}

/* A grid-wide coupled plan's replicates, reps_per_launch to a
 * cooperative launch (P->grid workgroups each). */
static hipError_t launch_grid_reps(SimInf_hip_plan *P, KParams q, hipStream_t st)
{
    const int R = std::max(1, P->reps_per_launch);
    q.rep_blocks = (int)P->grid;
    for (int r0 = 0; r0 < P->Nrep; r0 += R) {
        const int n = std::min(R, P->Nrep - r0);
        q.rep0 = r0;
        const hipError_t e = P->model->coupled(q, P->canon, true, P->grid * (unsigned)n,
                                               P->block, P->lds, st, nullptr);
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

/* Clears a run's result words and counters on the device (a kernel, not
 * memcpys or memsets, so that a run's launch sequence is kernels only and
 * can be captured in a graph). */
/* Zero of a fused-prevalence array (a kernel, so that runs can sit in a
 * graph). */
__global__ void k_clear_i64(long long *a, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        a[i] = 0;
}

/* Level-1 prevalence from the fused sums: per replicate the frozen
 * nodes' counts are summed over the columns of the difference array and
 * added to each column's sums, then cases / population as
 * k_prevalence_cols divides. */
__global__ void k_fused_finalize(const long long *sum, const long long *delta, int nrep,
                                 int tlen, double *out)
{
    const int rep = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (rep >= nrep)
        return;
    long long cx = 0, cy = 0;
    for (int it = 0; it < tlen; it++) {
        const long long *d = delta + 2 * ((long long)rep * (tlen + 1) + it);
        const long long *a = sum + 2 * ((long long)rep * tlen + it);
        cx += d[0];
        cy += d[1];
        out[(long long)rep * tlen + it] = (double)(a[0] + cx) / (double)(a[1] + cy);
    }
}

__global__ void k_run_init(unsigned long long *err_key, unsigned long long *transitions,
                           unsigned long long *frozen_cols, int *active_count,
                           unsigned *gsync, unsigned long long *rep_err, int nrep)
{
    if (threadIdx.x == 0) {
        *err_key = ~0ull;
        *transitions = 0;
        if (frozen_cols)
            *frozen_cols = 0;
        if (active_count)
            *active_count = 0;
    }
    if (gsync && threadIdx.x < SIMINF_GSYNC_WORDS)
        gsync[threadIdx.x] = 0u;
    if (rep_err)
        for (int i = (int)threadIdx.x; i < nrep; i += (int)blockDim.x)
            rep_err[i] = ~0ull;
}

/* Enqueue one run of a plan on its streams in three phases: (0) the
 * run's initialisation and, for node plans with a fill kernel, the
 * classify pass; (1) the solver: the active pass or the single kernel;
 * (2) the fill kernel on the second stream, joined back.  A batch runs
 * each phase for every plan before the next phase, so that the small
 * kernels of all plans reach the device before any fill kernel, whose
 * grid holds the dispatcher until it drains.  A single run takes the
 * phases in turn.  filled: whether the fill kernel ran (set in phase 0).
 * With capture the timing events are recorded as graph nodes. */
static int plan_enqueue_phase(SimInf_hip_plan *P, const KParams &base, bool capture,
                              int phase, bool *filled, hipEvent_t after_fill = nullptr)
{
    KParams p = base;
    hipStream_t st = P->stream;
    if (capture && P->kernel == K_COUPLED_GRID)
        return SIMINF_ERR_INVALID_MODEL;
    auto mark = [&](hipEvent_t ev, hipStream_t s) {
        (void)hipEventRecordWithFlags(ev, s, capture ? hipEventRecordExternal
                                                     : hipEventRecordDefault);
    };
    const bool nodes = P->kernel == K_SSM_NODES || P->kernel == K_AEM_NODES;
    auto launch = [&](const KParams &q, unsigned grid) {
        if (P->kernel == K_SSM_NODES)
            P->model->ssm_nodes(q, P->has_e1, P->canon, grid, P->block, st);
        else
            P->model->aem_nodes(q, P->has_e1, P->canon, grid, P->block, st);
    };
    const bool classify = nodes && p.frozen_from;
    if (phase == 0) {
        /* A fused-prevalence run has the classify and active passes but no
         * fill: frozen nodes go into the sums. */
        *filled = classify && !p.fp_sum;
        mark(P->ev_start, st);
        hipLaunchKernelGGL(k_run_init, dim3(1), dim3(256), 0, st, p.err_key, p.transitions,
                           p.frozen_cols, p.active_count, p.gsync, p.rep_err, P->Nrep);
        if (p.fp_sum) {
            const long long n1 = 2LL * P->Nrep * P->tlen, n2 = 2LL * P->Nrep * (P->tlen + 1);
            hipLaunchKernelGGL(k_clear_i64, dim3((unsigned)std::min<long long>((n1 + 255) / 256, 4096)),
                               dim3(256), 0, st, p.fp_sum, n1);
            hipLaunchKernelGGL(k_clear_i64, dim3((unsigned)std::min<long long>((n2 + 255) / 256, 4096)),
                               dim3(256), 0, st, p.fp_delta, n2);
        }
        if (classify) {
            /* Classify every lane; nodes frozen from the start go to the
             * fill kernel on the second stream while the active pass runs
             * the others, packed so that they spread over the SIMDs (the
             * pass reads how many there are on the device). */
            p.mode = NODES_CLASSIFY;
            launch(p, P->grid);
            mark(P->ev_classified, st);
            if (*filled && hipEventRecord(P->dep_fork, st) != hipSuccess)
                return SIMINF_ERR_DEVICE;
        }
    } else if (phase == 1) {
        switch (P->kernel) {
        case K_SSM_NODES:
        case K_AEM_NODES:
            if (!p.frozen_from) {
                p.mode = NODES_ALL;
                launch(p, P->grid);
                mark(P->ev_mid, st);
            } else {
                p.mode = NODES_ACTIVE;
                p.lpw_shift = P->lpw_fixed;
                mark(P->ev_mid, st);
                launch(p, P->grid_active);
                mark(P->ev_active1, st);
            }
            break;
        case K_COUPLED_GRID:
            /* Cooperative launches of reps_per_launch replicates each. */
            if (launch_grid_reps(P, p, st) != hipSuccess)
                return SIMINF_ERR_DEVICE;
            mark(P->ev_mid, st);
            break;
        case K_COUPLED_REPS:
            if (P->model->coupled(p, P->canon, false, P->grid, P->block, P->lds, st,
                                  nullptr) != hipSuccess)
                return SIMINF_ERR_DEVICE;
            mark(P->ev_mid, st);
            break;
        }
    } else {
        if (classify && !p.fp_sum) {
            hipStream_t fs = P->fill_stream;
            if (hipStreamWaitEvent(fs, P->dep_fork, 0) != hipSuccess ||
                (after_fill && hipStreamWaitEvent(fs, after_fill, 0) != hipSuccess))
                return SIMINF_ERR_DEVICE;
            mark(P->ev_fill0, fs);
            p.fill_block = P->fill_block;
            const dim3 fg(fill_grid(p)), fb(256);
            const bool v = P->Nd > 0;
#define SIMINF_FILL(NCV)                                                          \
    do {                                                                          \
        if (v)                                                                    \
            hipLaunchKernelGGL((k_fill_frozen<NCV, true>), fg, fb, 0, fs, p, P->Nc, \
                               P->Nd);                                            \
        else                                                                      \
            hipLaunchKernelGGL((k_fill_frozen<NCV, false>), fg, fb, 0, fs, p,      \
                               P->Nc, P->Nd);                                     \
    } while (0)
            switch (P->Nc) {
            case 2: SIMINF_FILL(2); break;
            case 3: SIMINF_FILL(3); break;
            case 4: SIMINF_FILL(4); break;
            case 6: SIMINF_FILL(6); break;
            default: SIMINF_FILL(0); break;
            }
#undef SIMINF_FILL
            mark(P->ev_fill1, fs);
            if (hipEventRecord(P->dep_join, fs) != hipSuccess ||
                hipStreamWaitEvent(st, P->dep_join, 0) != hipSuccess)
                return SIMINF_ERR_DEVICE;
        }
        mark(P->ev_stop, st);
        if (hipGetLastError() != hipSuccess)
            return SIMINF_ERR_DEVICE;
    }
    return 0;
}

/* All three phases of one run. */
static int plan_enqueue(SimInf_hip_plan *P, const KParams &base, bool capture, bool *filled)
{
    for (int ph = 0; ph < 3; ph++) {
        const int err = plan_enqueue_phase(P, base, capture, ph, filled);
        if (err)
            return err;
    }
    return 0;
}

/* Read a finished run's error key and counts (and, when timed, its
 * kernel times); returns the run's error code. */
/* Enqueue the copy of a run's result words to the plan's pinned host
 * copy on stream st (after the run's work on it). */
static bool plan_fetch(SimInf_hip_plan *P, hipStream_t st)
{
    return hipMemcpyAsync(P->host_words, P->result_words, 3 * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, st) == hipSuccess;
}

/* fetched: the result words were already copied (plan_fetch) and the copy
 * waited for. */
static int plan_collect(SimInf_hip_plan *P, bool filled, bool fetched = false)
{
    const KParams &p = P->p;
    if (!fetched && (!plan_fetch(P, P->stream) || hipStreamSynchronize(P->stream) != hipSuccess))
        return SIMINF_ERR_DEVICE;
    const unsigned long long key = P->host_words[0], trans = P->host_words[1];
    const unsigned long long fill_cols = p.frozen_cols ? P->host_words[2] : 0;
    {
        /* Device times: the whole run; the solver (classify and active
         * passes, or the one kernel); the fill kernel on its own
         * stream. */
        float ms = 0.0f, ms_solver = 0.0f, ms_active = 0.0f, ms_fill = 0.0f;
        (void)hipEventElapsedTime(&ms, P->ev_start, P->ev_stop);
        const bool classified =
            (P->kernel == K_SSM_NODES || P->kernel == K_AEM_NODES) && p.frozen_from;
        if (classified) {
            (void)hipEventElapsedTime(&ms_solver, P->ev_start, P->ev_classified);
            (void)hipEventElapsedTime(&ms_active, P->ev_mid, P->ev_active1);
            ms_solver += ms_active;
            if (filled)
                (void)hipEventElapsedTime(&ms_fill, P->ev_fill0, P->ev_fill1);
        } else {
            (void)hipEventElapsedTime(&ms_solver, P->ev_start, P->ev_mid);
        }
        P->kernel_ms = ms;
        P->solver_ms = ms_solver;
        P->fill_ms = ms_fill;
    }
    P->fill_cols = fill_cols;
    P->transitions = trans;
    P->last_key = std::min<uint64_t>(key, P->host_err_key);
    const int code = key_to_code(P->last_key);
    P->has_run = code == 0;
    return code;
}

extern "C" int SimInf_hip_plan_run(SimInf_hip_plan *P, unsigned long int seed)
{
    if (!P)
        return SIMINF_ERR_INVALID_MODEL;
    P->has_run = false;
    KParams p = P->p;
    p.seed = (unsigned long long)seed;
    p.seed_dev = nullptr;
    bool filled = false;
    const int err = plan_enqueue(P, p, false, &filled);
    if (err)
        return err;
    return plan_collect(P, filled);
}

extern "C" int SimInf_hip_plan_download(SimInf_hip_plan *P, int *U, double *V,
                                        double *prU, double *prV)
{
    if (!P)
        return SIMINF_ERR_INVALID_MODEL;
    if (P->fp_sum && (U || V || prU || prV))
        return SIMINF_ERR_INVALID_SUMMARY;   /* a fused plan keeps no trajectory */
    const size_t ncols = (size_t)P->Nrep * P->tlen;
    hipError_t e = hipSuccess;
    hipStream_t st = P->stream;
    if (U && P->dense_U)
        e = hipMemcpyAsync(U, P->p.U, (size_t)P->Nn * P->Nc * ncols * sizeof(int),
                           hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && prU && !P->dense_U && P->nnzU)
        e = hipMemcpyAsync(prU, P->p.prU, P->nnzU * sizeof(double), hipMemcpyDeviceToHost,
                           st);
    if (e == hipSuccess && V && P->dense_V && P->Nd > 0)
        e = hipMemcpyAsync(V, P->p.V, (size_t)P->Nn * P->Nd * ncols * sizeof(double),
                           hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && prV && !P->dense_V && P->nnzV)
        e = hipMemcpyAsync(prV, P->p.prV, P->nnzV * sizeof(double), hipMemcpyDeviceToHost,
                           st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
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

extern "C" int SimInf_hip_plan_kernel_times(const SimInf_hip_plan *P,
                                            double *solver_ms, double *fill_ms,
                                            long long *fill_columns)
{
    if (!P)
        return SIMINF_ERR_INVALID_MODEL;
    if (solver_ms)
        *solver_ms = P->solver_ms;
    if (fill_ms)
        *fill_ms = P->fill_ms;
    if (fill_columns)
        *fill_columns = (long long)P->fill_cols;
    return 0;
}

/* The lane -> node map of a grid-wide coupled plan: the n given herds
 * each alone on a wavefront (its other lanes only help it draw its
 * attempts), then the other herds in node order, pack to a wavefront (the
 * wavefront's remaining lanes help), then the workgroup that runs E2. */
static int plan_set_herd_map(SimInf_hip_plan *P, const int *nodes, int n, int pack)
{
    if (P->kernel != K_COUPLED_GRID)
        return n == 0 && pack == SIMINF_GRID_BLOCK ? 0 : SIMINF_ERR_INVALID_MODEL;
    const int Nn = P->Nn;
    if (n < 0 || n > SIMINF_SOLO_MAX || n > Nn || (n > 0 && !nodes) ||
        pack < SIMINF_PACK_MIN || pack > SIMINF_GRID_BLOCK)
        return SIMINF_ERR_INVALID_MODEL;
    std::vector<char> solo((size_t)Nn, 0);
    for (int i = 0; i < n; i++) {
        if (nodes[i] < 0 || nodes[i] >= Nn || solo[nodes[i]])
            return SIMINF_ERR_INVALID_MODEL;
        solo[nodes[i]] = 1;
    }
    const int grid = n + (Nn - n + pack - 1) / pack + 1;
    if (grid > P->grid_max)
        return SIMINF_ERR_ALLOC_MEMORY_BUFFER;
    std::vector<int> map((size_t)grid * SIMINF_GRID_BLOCK, -1);
    for (int i = 0; i < n; i++)
        map[(size_t)i * SIMINF_GRID_BLOCK] = nodes[i];
    size_t k = 0;
    for (int node = 0; node < Nn; node++) {
        if (solo[node])
            continue;
        map[(size_t)(n + k / pack) * SIMINF_GRID_BLOCK + k % pack] = node;
        k++;
    }
    if (hipMemcpyAsync((void *)P->p.lane_node, map.data(), map.size() * sizeof(int),
                       hipMemcpyHostToDevice, P->stream) != hipSuccess ||
        hipStreamSynchronize(P->stream) != hipSuccess)
        return SIMINF_ERR_DEVICE;
    P->grid = (unsigned)grid;
    P->solo = n;
    P->pack = pack;
    P->reps_per_launch = std::max(1, std::min(P->Nrep, P->max_blocks / grid));
    return 0;
}

extern "C" int SimInf_hip_plan_set_herd_map(SimInf_hip_plan *P, const int *nodes, int n,
                                            int herds_per_wavefront)
{
    if (!P)
        return SIMINF_ERR_INVALID_MODEL;
    return plan_set_herd_map(P, nodes, n, herds_per_wavefront);
}

extern "C" int SimInf_hip_plan_get_herd_map(const SimInf_hip_plan *P, int *n_solo,
                                            int *herds_per_wavefront)
{
    if (!P)
        return SIMINF_ERR_INVALID_MODEL;
    if (n_solo)
        *n_solo = P->solo;
    if (herds_per_wavefront)
        *herds_per_wavefront = P->pack;
    return 0;
}

extern "C" int SimInf_hip_plan_set_replicate_data(SimInf_hip_plan *P, const double *ldata,
                                                  int ldata_nodes, const double *gdata)
{
    if (!P || (ldata && ldata_nodes != P->Nn && ldata_nodes != 1))
        return SIMINF_ERR_INVALID_MODEL;
    const size_t nl = (size_t)P->Nld * (ldata ? ldata_nodes : 0) * P->Nrep;
    const size_t ng = (size_t)P->Ngdata * P->Nrep;
    if (ldata && nl == 0)
        return SIMINF_ERR_INVALID_MODEL;
    if (gdata && ng == 0)
        return SIMINF_ERR_INVALID_MODEL;
    if (ldata && P->model->neighbors > 0 &&
        !neighbor_lists_ok(ldata, P->Nld, (long long)ldata_nodes * P->Nrep,
                           P->model->neighbors - 1, P->Nn))
        return SIMINF_ERR_INVALID_MODEL;
    DeviceBuffers &B = P->buf;
    if (ldata) {
        const int mode = ldata_nodes == P->Nn ? 1 : 2;
        if (!P->rep_ldata || mode != P->p.ld_rep) {
            /* Room for the per-node layout, which also holds the other. */
            if (!P->rep_ldata)
                P->rep_ldata = B.alloc<double>((size_t)P->Nld * P->Nn * P->Nrep);
            P->param_version++;
        }
        B.copy(P->rep_ldata, ldata, nl * sizeof(double), hipMemcpyHostToDevice);
        P->p.ldata = P->rep_ldata;
        P->p.ld_rep = mode;
    }
    if (gdata) {
        if (!P->rep_gdata) {
            P->rep_gdata = B.alloc<double>(ng);
            P->param_version++;
        }
        B.copy(P->rep_gdata, gdata, ng * sizeof(double), hipMemcpyHostToDevice);
        P->p.gdata = P->rep_gdata;
        P->p.gd_stride = P->Ngdata;
    }
    return B.status == hipSuccess ? 0 : SIMINF_ERR_DEVICE;
}

/* Columns per fill block that the autotuners try. */
static const int fill_block_candidates[] = {4, 8, 16, 32, 64};

extern "C" int SimInf_hip_plan_last_error(const SimInf_hip_plan *P, int *code, int *replicate,
                                          int *day, int *phase, long long *where)
{
    if (!P)
        return SIMINF_ERR_INVALID_MODEL;
    const uint64_t k = P->last_key;
    const int c = key_to_code(k);
    const int rep = (int)(k >> 48), d = (int)((k >> 30) & 0x3FFFF);
    const int ph = (int)((k >> 27) & 7);
    long long pos = (long long)((k >> 6) & 0x1FFFFF);
    if (c && (ph == PH_E1 || ph == PH_E2))
        pos = d < (int)P->ev_day_first.size() && P->ev_day_first[d] >= 0
                  ? P->ev_day_first[d] + pos : -1;
    if (code)
        *code = c;
    if (replicate)
        *replicate = c ? rep : -1;
    if (day)
        *day = c ? d : -1;
    if (phase)
        *phase = c ? ph : -1;
    if (where)
        *where = c ? pos : -1;
    return 0;
}

extern "C" int SimInf_hip_plan_set_fill_block(SimInf_hip_plan *P, int cols)
{
    if (!P || cols < 0 || cols > 4096)
        return SIMINF_ERR_INVALID_MODEL;
    if (cols != P->fill_block) {
        P->fill_block = cols;
        P->param_version++;      /* a captured batch graph holds the old one */
    }
    return 0;
}

extern "C" int SimInf_hip_plan_get_fill_block(const SimInf_hip_plan *P, int *cols)
{
    if (!P || !cols)
        return SIMINF_ERR_INVALID_MODEL;
    *cols = P->fill_block > 0 ? P->fill_block : SIMINF_FILL_COLS;
    return 0;
}

extern "C" int SimInf_hip_plan_set_lanes_per_wave(SimInf_hip_plan *P, int shift)
{
    if (!P || shift < -1 || shift > 6)
        return SIMINF_ERR_INVALID_MODEL;
    if (shift != P->lpw_fixed) {
        P->lpw_fixed = shift;
        P->param_version++;
    }
    return 0;
}

extern "C" int SimInf_hip_plan_autotune(SimInf_hip_plan *P, unsigned long int seed, int reps,
                                        int *best_shift, double *ms)
{
    if (!P || reps < 1)
        return SIMINF_ERR_INVALID_MODEL;
    if (P->kernel == K_COUPLED_GRID) {
        /* Herds ranked by their transitions in a first run; then, with
         * the 4 busiest on wavefronts of their own, the others 64, 32, 16
         * or 8 to a wavefront; then, with the best of those, 0, 2, 8 or
         * 16 herds on their own. */
        int err = plan_set_herd_map(P, nullptr, 0, SIMINF_GRID_BLOCK);
        if (!err)
            err = SimInf_hip_plan_run(P, seed);
        if (err)
            return err;
        std::vector<unsigned long long> nf((size_t)P->Nn);
        if (hipMemcpyAsync(nf.data(), P->p.node_fired, nf.size() * sizeof(nf[0]),
                           hipMemcpyDeviceToHost, P->stream) != hipSuccess ||
            hipStreamSynchronize(P->stream) != hipSuccess)
            return SIMINF_ERR_DEVICE;
        std::vector<int> order((size_t)P->Nn);
        for (int i = 0; i < P->Nn; i++)
            order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return nf[a] > nf[b]; });
        int best_solo = 0, best_pack = SIMINF_GRID_BLOCK;
        double best_ms = INFINITY;
        auto trial = [&](int c, int solo, int pack) -> int {
            if (ms)
                ms[c] = INFINITY;
            if (solo > P->Nn || plan_set_herd_map(P, order.data(), solo, pack) != 0)
                return 0;
            double sum = 0.0;
            for (int r = 0; r < reps; r++) {
                const int e = SimInf_hip_plan_run(P, seed + (unsigned long)r);
                if (e)
                    return e;
                sum += P->solver_ms;
            }
            if (ms)
                ms[c] = sum / reps;
            if (sum / reps < best_ms) {
                best_ms = sum / reps;
                best_solo = solo;
                best_pack = pack;
            }
            return 0;
        };
        static const int packs[4] = {64, 32, 16, 8}, solos[4] = {0, 2, 8, 16};
        for (int c = 0; c < 4 && !err; c++)
            err = trial(c, std::min(4, P->Nn), packs[c]);
        const int pack = best_pack;
        for (int c = 0; c < 4 && !err; c++)
            err = trial(4 + c, solos[c], pack);
        if (!err)
            err = plan_set_herd_map(P, order.data(), best_solo, best_pack);
        if (err)
            return err;
        if (best_shift)
            *best_shift = best_solo;
        return 0;
    }
    if (P->kernel != K_SSM_NODES && P->kernel != K_AEM_NODES) {
        /* Nothing to tune: one workgroup per replicate. */
        if (best_shift)
            *best_shift = -1;
        return 0;
    }
    int best = -1;
    double best_ms = INFINITY;
    for (int shift = 0; shift <= 6; shift++) {
        P->lpw_fixed = shift;
        double sum = 0.0;
        for (int r = 0; r < reps; r++) {
            const int err = SimInf_hip_plan_run(P, seed + (unsigned long)r);
            if (err) {
                P->lpw_fixed = -1;
                return err;
            }
            sum += P->solver_ms;
        }
        if (ms)
            ms[shift] = sum / reps;
        if (sum / reps < best_ms) {
            best_ms = sum / reps;
            best = shift;
        }
    }
    P->lpw_fixed = best;
    if (best_shift)
        *best_shift = best;
    /* Then the fill kernel's columns per block, by the run's device time
     * (the fill runs beside the active pass). */
    if (P->p.frozen_from) {
        int best_fb = P->fill_block;
        double best_run = INFINITY;
        for (int fb : fill_block_candidates) {
            P->fill_block = fb;
            double sum = 0.0;
            for (int r = 0; r < reps; r++) {
                const int err = SimInf_hip_plan_run(P, seed + (unsigned long)r);
                if (err) {
                    P->fill_block = best_fb;
                    return err;
                }
                sum += P->kernel_ms;
            }
            if (sum / reps < best_run) {
                best_run = sum / reps;
                best_fb = fb;
            }
        }
        P->fill_block = best_fb;
    }
    P->param_version++;
    return 0;
}

extern "C" void SimInf_hip_plan_destroy(SimInf_hip_plan *P)
{
    if (!P)
        return;
    if (P->fill_stream)
        (void)hipStreamSynchronize(P->fill_stream);
    if (P->stream) {
        (void)hipStreamSynchronize(P->stream);
        P->buf.release();
    }
    if (P->ev_start)
        (void)hipEventDestroy(P->ev_start);
    if (P->ev_mid)
        (void)hipEventDestroy(P->ev_mid);
    if (P->ev_classified)
        (void)hipEventDestroy(P->ev_classified);
    if (P->ev_fill0)
        (void)hipEventDestroy(P->ev_fill0);
    if (P->ev_fill1)
        (void)hipEventDestroy(P->ev_fill1);
    if (P->ev_active1)
        (void)hipEventDestroy(P->ev_active1);
    if (P->dep_fork)
        (void)hipEventDestroy(P->dep_fork);
    if (P->dep_join)
        (void)hipEventDestroy(P->dep_join);
    if (P->fill_stream)
        (void)hipStreamDestroy(P->fill_stream);
    if (P->ev_stop)
        (void)hipEventDestroy(P->ev_stop);
    if (P->stream)
        (void)hipStreamDestroy(P->stream);
    if (P->host_words)
        (void)hipHostFree(P->host_words);
    delete P;
}

/* ------------------------------------------------------------------ */
/* Batches: several plans in one graph launch.                         */
/* ------------------------------------------------------------------ */

struct SimInf_hip_batch {
    std::vector<SimInf_hip_plan *> plans;
    std::vector<char> filled;
    std::vector<hipEvent_t> fork, join;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    unsigned long long *seeds = nullptr;   /* device, one per plan */
    double ms = 0.0;
    bool use_graph = true;
    /* The plans' parameter versions the graph was captured with. */
    std::vector<unsigned> versions;
};

extern "C" void SimInf_hip_batch_destroy(SimInf_hip_batch *B)
{
    if (!B)
        return;
    if (B->stream)
        (void)hipStreamSynchronize(B->stream);
    if (B->exec)
        (void)hipGraphExecDestroy(B->exec);
    if (B->graph)
        (void)hipGraphDestroy(B->graph);
    for (hipEvent_t e : B->fork)
        if (e)
            (void)hipEventDestroy(e);
    for (hipEvent_t e : B->join)
        if (e)
            (void)hipEventDestroy(e);
    if (B->ev0)
        (void)hipEventDestroy(B->ev0);
    if (B->ev1)
        (void)hipEventDestroy(B->ev1);
    if (B->seeds)
        (void)hipFree(B->seeds);
    if (B->stream)
        (void)hipStreamDestroy(B->stream);
    delete B;
}

/* Capture the batch's graph (replacing a previous one): every plan's run
 * on its own branch, forked from the batch's stream into the plan's
 * streams, the plan's kernels with the seed read from the batch's seed
 * table and its current launch shape, joined back. */
static int batch_capture(SimInf_hip_batch *B)
{
    if (B->exec) {
        (void)hipGraphExecDestroy(B->exec);
        B->exec = nullptr;
    }
    if (B->graph) {
        (void)hipGraphDestroy(B->graph);
        B->graph = nullptr;
    }
    const int n = (int)B->plans.size();
    int err = 0;
    B->versions.assign(n, 0);
    for (int i = 0; i < n; i++)
        B->versions[i] = B->plans[i]->param_version;
    if (hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return SIMINF_ERR_DEVICE;
    for (int i = 0; i < n && !err; i++) {
        if (hipEventRecord(B->fork[i], B->stream) != hipSuccess ||
            hipStreamWaitEvent(B->plans[i]->stream, B->fork[i], 0) != hipSuccess)
            err = SIMINF_ERR_DEVICE;
    }
    for (int ph = 0; ph < 3 && !err; ph++) {
        for (int i = 0; i < n && !err; i++) {
            SimInf_hip_plan *P = B->plans[i];
            KParams p = P->p;
            p.seed_dev = B->seeds + i;
            bool filled = B->filled[i] != 0;
            /* The captured graph leaves the fills side by side (no wait of
             * one plan's fill stream on another's, a form not captured
             * here). */
            err = plan_enqueue_phase(P, p, true, ph, &filled);
            B->filled[i] = filled;
        }
    }
    for (int i = 0; i < n && !err; i++) {
        if (hipEventRecord(B->join[i], B->plans[i]->stream) != hipSuccess ||
            hipStreamWaitEvent(B->stream, B->join[i], 0) != hipSuccess)
            err = SIMINF_ERR_DEVICE;
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(B->stream, &g);
    B->graph = g;
    if (!err && (ec != hipSuccess || !g ||
                 hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0) != hipSuccess))
        err = SIMINF_ERR_DEVICE;
    (void)hipGetLastError();
    return err;
}

extern "C" int SimInf_hip_batch_create(SimInf_hip_batch **out, SimInf_hip_plan *const *plans,
                                       int n)
{
    if (!out || !plans || n < 1)
        return SIMINF_ERR_INVALID_MODEL;
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!plans[i] || plans[i]->kernel == K_COUPLED_GRID)
            return SIMINF_ERR_INVALID_MODEL;
        for (int j = 0; j < i; j++)
            if (plans[j] == plans[i])
                return SIMINF_ERR_INVALID_MODEL;
    }
    SimInf_hip_batch *B = new SimInf_hip_batch();
    B->plans.assign(plans, plans + n);
    B->filled.assign(n, 0);
    B->fork.assign(n, nullptr);
    B->join.assign(n, nullptr);
    bool ok = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&B->ev0) == hipSuccess && hipEventCreate(&B->ev1) == hipSuccess &&
              hipMalloc((void **)&B->seeds, sizeof(unsigned long long) * n) == hipSuccess;
    for (int i = 0; ok && i < n; i++)
        ok = hipEventCreateWithFlags(&B->fork[i], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&B->join[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        SimInf_hip_batch_destroy(B);
        return SIMINF_ERR_DEVICE;
    }
    const int err = batch_capture(B);
    if (err) {
        SimInf_hip_batch_destroy(B);
        return err;
    }
    *out = B;
    return 0;
}

extern "C" int SimInf_hip_batch_set_graph(SimInf_hip_batch *B, int use_graph)
{
    if (!B)
        return SIMINF_ERR_INVALID_MODEL;
    B->use_graph = use_graph != 0;
    return 0;
}

/* A batch run without the graph: every plan's run enqueued on its own
 * streams behind one event of the batch's stream, which then waits for
 * all of them. */
static int batch_run_streams(SimInf_hip_batch *B, const unsigned long *seeds)
{
    const int n = (int)B->plans.size();
    /* The plans go in by their last solver time, longest first: the
     * host's enqueue of each plan delays the next one's start, which is
     * best spent on plans that have time to spare. */
    std::vector<int> order((size_t)n);
    for (int i = 0; i < n; i++)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return B->plans[a]->solver_ms > B->plans[b]->solver_ms;
    });
    if (hipEventRecord(B->ev0, B->stream) != hipSuccess)
        return SIMINF_ERR_DEVICE;
    for (int i : order)
        if (hipStreamWaitEvent(B->plans[i]->stream, B->ev0, 0) != hipSuccess)
            return SIMINF_ERR_DEVICE;
    /* The fill kernels run one after another (each after the previous
     * plan's fill): they are bound by HBM writes, so running them at once
     * takes as long as in turn, and in turn each one's timing is its
     * own. */
    for (int ph = 0; ph < 3; ph++) {
        hipEvent_t prev = nullptr;
        for (int i : order) {
            SimInf_hip_plan *P = B->plans[i];
            KParams p = P->p;
            p.seed = (unsigned long long)seeds[i];
            p.seed_dev = nullptr;
            bool filled = B->filled[i] != 0;
            const int err = plan_enqueue_phase(P, p, false, ph, &filled, ph == 2 ? prev : nullptr);
            if (err)
                return err;
            B->filled[i] = filled;
            if (ph == 2 && filled)
                prev = P->dep_join;
        }
    }
    for (int i : order) {
        if (hipEventRecord(B->join[i], B->plans[i]->stream) != hipSuccess ||
            hipStreamWaitEvent(B->stream, B->join[i], 0) != hipSuccess)
            return SIMINF_ERR_DEVICE;
    }
    if (hipEventRecord(B->ev1, B->stream) != hipSuccess)
        return SIMINF_ERR_DEVICE;
    /* Every plan's result words in one wait. */
    for (SimInf_hip_plan *P : B->plans)
        if (!plan_fetch(P, B->stream))
            return SIMINF_ERR_DEVICE;
    if (hipStreamSynchronize(B->stream) != hipSuccess)
        return SIMINF_ERR_DEVICE;
    return 0;
}

extern "C" int SimInf_hip_batch_run(SimInf_hip_batch *B, const unsigned long *seeds)
{
    if (!B || !seeds)
        return SIMINF_ERR_INVALID_MODEL;
    const int n = (int)B->plans.size();
    if (!B->use_graph) {
        for (SimInf_hip_plan *P : B->plans)
            P->has_run = false;
        const int err = batch_run_streams(B, seeds);
        if (err)
            return err;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, B->ev0, B->ev1);
        B->ms = ms;
        int first = 0;
        for (int i = 0; i < n; i++) {
            const int code = plan_collect(B->plans[i], B->filled[i] != 0, true);
            if (code && !first)
                first = code;
        }
        return first;
    }
    std::vector<unsigned long long> sd(seeds, seeds + n);
    for (SimInf_hip_plan *P : B->plans)
        P->has_run = false;
    /* A plan whose kernels now read their parameters elsewhere: capture
     * again. */
    for (int i = 0; i < n; i++) {
        if (B->plans[i]->param_version != B->versions[i]) {
            const int err = batch_capture(B);
            if (err)
                return err;
            break;
        }
    }
    if (hipMemcpyAsync(B->seeds, sd.data(), sizeof(sd[0]) * n, hipMemcpyHostToDevice,
                       B->stream) != hipSuccess ||
        hipEventRecord(B->ev0, B->stream) != hipSuccess ||
        hipGraphLaunch(B->exec, B->stream) != hipSuccess ||
        hipEventRecord(B->ev1, B->stream) != hipSuccess)
        return SIMINF_ERR_DEVICE;
    for (SimInf_hip_plan *P : B->plans)
        if (!plan_fetch(P, B->stream))
            return SIMINF_ERR_DEVICE;
    if (hipStreamSynchronize(B->stream) != hipSuccess)
        return SIMINF_ERR_DEVICE;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, B->ev0, B->ev1);
    B->ms = ms;
    /* Each plan's results and kernel times (its timing events are
     * event-record nodes of the graph). */
    int first = 0;
    for (int i = 0; i < n; i++) {
        const int code = plan_collect(B->plans[i], B->filled[i] != 0, true);
        if (code && !first)
            first = code;
    }
    return first;
}

/* Autotuner of a batch's launch (no reference counterpart): the plans
 * share the device, so each one's lanes per wavefront and fill block
 * are tuned in the batch rather than alone.  Coordinate descent, two
 * passes: for each plan with an active pass and a fill kernel in turn,
 * every shift (-1 for the default rule, 0 to 6), then every columns-per-
 * fill-block candidate, with the others held, keeping the least median
 * launch time over reps launches (seeds from seed); then the graph
 * against the direct enqueue, the faster kept (the graph recaptured with
 * the kept shapes).  Results do not depend on any of it. */
extern "C" int SimInf_hip_batch_autotune(SimInf_hip_batch *B, unsigned long int seed, int reps,
                                         double *best_ms)
{
    if (!B || reps < 1)
        return SIMINF_ERR_INVALID_MODEL;
    const int n = (int)B->plans.size();
    std::vector<unsigned long> sd((size_t)n);
    unsigned long next = seed;
    /* The median launch time over reps launches: a launch slowed by
     * something outside the batch (a clock dip, the host) does not decide
     * a shape. */
    auto measure = [&](double *median) -> int {
        std::vector<double> t((size_t)reps);
        for (int r = 0; r < reps; r++) {
            for (int i = 0; i < n; i++)
                sd[i] = next++;
            const int err = SimInf_hip_batch_run(B, sd.data());
            if (err)
                return err;
            t[r] = B->ms;
        }
        std::sort(t.begin(), t.end());
        *median = reps % 2 ? t[reps / 2] : 0.5 * (t[reps / 2 - 1] + t[reps / 2]);
        return 0;
    };
    B->use_graph = false;
    double best = 0.0;
    int err = measure(&best);
    for (int pass = 0; pass < 2 && !err; pass++) {
        for (int i = 0; i < n && !err; i++) {
            SimInf_hip_plan *P = B->plans[i];
            if ((P->kernel != K_SSM_NODES && P->kernel != K_AEM_NODES) || !P->p.frozen_from)
                continue;
            const int keep0 = P->lpw_fixed;
            int keep = keep0;
            for (int shift = -1; shift <= 6 && !err; shift++) {
                if (shift == keep0)
                    continue;
                P->lpw_fixed = shift;
                double t = 0.0;
                err = measure(&t);
                if (!err && t < best) {
                    best = t;
                    keep = shift;
                }
            }
            P->lpw_fixed = keep;
            /* The plan's trajectory-flush schedule: columns per fill
             * block. */
            const int fb0 = P->fill_block;
            int keep_fb = fb0;
            for (int fb : fill_block_candidates) {
                if (err || fb == fb0)
                    continue;
                P->fill_block = fb;
                double t = 0.0;
                err = measure(&t);
                if (!err && t < best) {
                    best = t;
                    keep_fb = fb;
                }
            }
            P->fill_block = keep_fb;
        }
    }
    if (!err)
        err = batch_capture(B);
    if (!err) {
        B->use_graph = true;
        double tg = 0.0;
        err = measure(&tg);
        if (!err && tg < best)
            best = tg;
        else
            B->use_graph = false;
    }
    if (best_ms)
        *best_ms = best;
    return err;
}

/* The lanes-per-wavefront shift the batch's autotuner kept for plan i
 * (-1: the default rule) and whether it launches the graph. */
extern "C" int SimInf_hip_batch_get_shape(const SimInf_hip_batch *B, int i, int *shift,
                                          int *use_graph)
{
    if (!B || i < 0 || i >= (int)B->plans.size())
        return SIMINF_ERR_INVALID_MODEL;
    if (shift)
        *shift = B->plans[i]->lpw_fixed;
    if (use_graph)
        *use_graph = B->use_graph ? 1 : 0;
    return 0;
}

extern "C" int SimInf_hip_batch_time(const SimInf_hip_batch *B, double *ms)
{
    if (!B || !ms)
        return SIMINF_ERR_INVALID_MODEL;
    *ms = B->ms;
    return 0;
}

extern "C" int SimInf_hip_release(void)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return SIMINF_ERR_DEVICE;
    if (pool_allocation(dev)) {
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess ||
            hipMemPoolTrimTo(pool, 0) != hipSuccess)
            return SIMINF_ERR_DEVICE;
    }
    return 0;
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
    case SIMINF_ERR_INVALID_OBS_PROCESS: return "Invalid observation process.";
    case SIMINF_ERR_INVALID_WEIGHT: return "Invalid weight detected (non-finite or < 0.0).";
    case SIMINF_ERR_WEIGHT_SUM: return "Non-positive sum of weights detected.";
    case SIMINF_ERR_EMPTY_TRAJECTORY:
        return "Please run the model first, the trajectory is empty.";
    case SIMINF_ERR_INVALID_SUMMARY: return "Invalid summary specification.";
    case SIMINF_ERR_UNKNOWN_DISTRIBUTION: return "Unknown distribution.";
    case SIMINF_ERR_ABC_WEIGHTS: return "Unable to calculate weights.";
    default: return "Unknown error code.";
    }
}

/* ------------------------------------------------------------------ */
/* Bootstrap particle filter (R/pfilter.R:315-499).                     */
/* ------------------------------------------------------------------ */

/* Weight of each particle: the log-probability of the interval's
 * observations given its state at the interval's end.  One workgroup of
 * 64 lanes per particle; lane l adds the nodes l, l+64, ... in order and
 * lane 0 adds the 64 partial sums in lane order, the order the checker
 * uses.  With one observation for all nodes the count observed is the
 * sum over the nodes (exact in any order). */
__global__ __launch_bounds__(64) void k_pf_weights(const int *U, int Nn, int Nc,
                                                   int tlen, int c_obs, int dist,
                                                   double p1, double p2,
                                                   const int *obs, int obs_nodes,
                                                   double *w)
{
    __shared__ double partial[64];
    __shared__ long long total[64];
    const int part = (int)blockIdx.x, lane = (int)threadIdx.x;
    const long long col = ((long long)part * tlen + (tlen - 1)) * Nn;
    if (obs_nodes == 1) {
        long long t = 0;
        for (int node = lane; node < Nn; node += 64)
            t += U[(col + node) * Nc + c_obs];
        total[lane] = t;
        __syncthreads();
        if (lane == 0) {
            long long sum = 0;
            for (int l = 0; l < 64; l++)
                sum += total[l];
            w[part] = obs[0] < 0 ? 0.0 : siminf_obs_logw(dist, obs[0], sum, p1, p2);
        }
        return;
    }
    double acc = 0.0;
    for (int node = lane; node < Nn; node += 64) {
        const int x = obs[node];
        if (x >= 0)
            acc = acc + siminf_obs_logw(dist, x, U[(col + node) * Nc + c_obs], p1, p2);
    }
    partial[lane] = acc;
    __syncthreads();
    if (lane == 0) {
        double sum = 0.0;
        for (int l = 0; l < 64; l++)
            sum = sum + partial[l];
        w[part] = sum;
    }
}

/* Particles whose weights the resampling kernel keeps in LDS (8 bytes
 * each, 56 KiB in all, within the default dynamic LDS limit); more use a
 * global buffer. */
#define SIMINF_PF_LDS 7168

/* Sum of x[0..n) in index order, from 0.0 (the loads run ahead of the
 * dependent adds). */
__device__ __forceinline__ double pf_serial_sum(const double *x, int n)
{
    double s = 0.0;
    int i = 0;
    /* Sixteen loads issued together, then the dependent adds. */
    for (; i + 16 <= n; i += 16) {
        double v[16];
#pragma unroll
        for (int k = 0; k < 16; k++)
            v[k] = x[i + k];
#pragma unroll
        for (int k = 0; k < 16; k++)
            s = s + v[k];
    }
    for (; i < n; i++)
        s = s + x[i];
    return s;
}

/* Log-sum-exp of the weights, the log-likelihood increment, the ESS and
 * systematic resampling (pfilter.R:455-474), in one workgroup, with the
 * reference's floating-point sequence: the sums run in particle order on
 * one lane, over weights staged in LDS.  Resampling
 * (SimInf_particle_filter.c:32-66) walks u_i = u_0 + du + ... + du against
 * the running sums c_j = w_0 + ... + w_j; since both sequences are
 * nondecreasing, the walk's ancestor of particle i is the first j with
 * u_i <= c_j (at most n-1), found here by a binary search per particle
 * over the stored running sums, after one lane has written them and the
 * u_i (the same additions, in the same order, as the walk). */
template <bool LDS>
__global__ __launch_bounds__(256) void k_pf_resample(double *w, double *c_global, int n,
                                                     int interval, unsigned long long seed,
                                                     double *loglik, double *ess,
                                                     int *anc, int *err)
{
    extern __shared__ double c_lds[];
    __shared__ double s_max[256];
    __shared__ int s_bad[256];
    __shared__ double s_sum;
    __shared__ int s_res;
    double *c = LDS ? c_lds : c_global;
    const int tid = (int)threadIdx.x;
    const double NEG_INF = -INFINITY;
    double m = NEG_INF;
    int bad = 0;
    for (int i = tid; i < n; i += 256) {
        const double x = w[i];
        if (!(x > NEG_INF && x < INFINITY))
            bad = 1;
        m = x > m ? x : m;
    }
    s_max[tid] = m;
    s_bad[tid] = bad;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
            s_max[tid] = s_max[tid + h] > s_max[tid] ? s_max[tid + h] : s_max[tid];
            s_bad[tid] |= s_bad[tid + h];
        }
        __syncthreads();
    }
    if (s_bad[0]) {
        /* Keep every particle (the run reports the error). */
        for (int i = tid; i < n; i += 256)
            anc[i] = i;
        if (tid == 0)
            *err = SIMINF_ERR_INVALID_OBS_PROCESS;
        return;
    }
    const double max_w = s_max[0];
    for (int i = tid; i < n; i += 256)
        c[i] = siminf_exp(w[i] - max_w);
    __syncthreads();
    if (tid == 0) {
        const double sum = pf_serial_sum(c, n);
        *loglik = ((*loglik + max_w) + siminf_log(sum)) - siminf_log((double)n);
        s_sum = sum;
    }
    __syncthreads();
    const double sum = s_sum;
    for (int i = tid; i < n; i += 256)
        c[i] = c[i] / sum;
    __syncthreads();
    if (tid == 0) {
        /* ESS, the validity pass and the running sums; then the u_i. */
        double s2 = 0.0, cs = 0.0;
        int r = 0;
        int i = 0;
        for (; i + 16 <= n; i += 16) {
            double v[16];
#pragma unroll
            for (int k = 0; k < 16; k++)
                v[k] = c[i + k];
#pragma unroll
            for (int k = 0; k < 16; k++) {
                const double x = v[k];
                s2 = s2 + x * x;
                if (!(x >= 0.0) || x == INFINITY)
                    r = -1;
                cs = cs + x;
                v[k] = cs;
            }
#pragma unroll
            for (int k = 0; k < 16; k++)
                c[i + k] = v[k];
        }
        for (; i < n; i++) {
            const double x = c[i];
            s2 = s2 + x * x;
            if (!(x >= 0.0) || x == INFINITY)
                r = -1;
            cs = cs + x;
            c[i] = cs;
        }
        ess[interval] = 1.0 / s2;
        if (r == 0 && cs <= 0.0)
            r = -2;
        if (r == 0) {
            const siminf_stream st = siminf_stream_make(seed, SIMINF_DOMAIN_PF,
                                                        (uint64_t)interval);
            const double du = cs / (double)n;
            double u = du * siminf_uniform_at(&st, 0);
            for (int i = 0; i < n; i++) {
                w[i] = u;
                u += du;
            }
        }
        s_res = r;
    }
    __syncthreads();
    const int r = s_res;
    if (r) {
        for (int i = tid; i < n; i += 256)
            anc[i] = i;
        if (tid == 0)
            *err = r == -1 ? SIMINF_ERR_INVALID_WEIGHT : SIMINF_ERR_WEIGHT_SUM;
        return;
    }
    for (int i = tid; i < n; i += 256) {
        const double u = w[i];
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (u > c[mid])
                lo = mid + 1;
            else
                hi = mid;
        }
        anc[i] = lo;
    }
}

/* Each particle continues from its ancestor's state at the interval's
 * end (pfilter.R:398-410). */
__global__ __launch_bounds__(256) void k_pf_gather(const int *U, const double *V,
                                                   int Nn, int Nc, int Nd, int tlen,
                                                   const int *anc, long long lanes,
                                                   int *u_cur, double *v_cur)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= lanes)
        return;
    const long long part = g / Nn;
    const long long node = g - part * Nn;
    const long long src = ((long long)anc[part] * tlen + (tlen - 1)) * Nn + node;
    for (int c = 0; c < Nc; c++)
        u_cur[g * Nc + c] = U[src * Nc + c];
    for (int c = 0; c < Nd; c++)
        v_cur[g * Nd + c] = V[src * Nd + c];
}

extern "C" int SimInf_hip_pfilter(const SimInf_solver_args *a,
                                  const SimInf_hip_model *model, int n_particles,
                                  int n_intervals, const double *tspan2,
                                  int obs_compartment, int obs_dist, double obs_p1,
                                  double obs_p2, int obs_nodes, const int *obs,
                                  unsigned long int seed,
                                  double *loglik, double *ess, int *ancestors)
{
    if (!a || !model || !tspan2 || !obs || !loglik || !ess || n_particles < 1 ||
        n_intervals < 1 || a->Nrep > 1 || a->len < 0 || obs_compartment < 0 ||
        obs_compartment >= a->Nc || (obs_nodes != 1 && obs_nodes != a->Nn) ||
        (obs_dist != SIMINF_OBS_BINOMIAL && obs_dist != SIMINF_OBS_POISSON) ||
        (obs_dist == SIMINF_OBS_BINOMIAL && !(obs_p1 >= 0.0 && obs_p1 <= 1.0)) ||
        (obs_dist == SIMINF_OBS_POISSON && !(std::isfinite(obs_p1) && std::isfinite(obs_p2))))
        return SIMINF_ERR_INVALID_MODEL;
    for (int i = 0; i < n_intervals; i++) {
        const double tb = tspan2[2 * i], te = tspan2[2 * i + 1];
        if (!std::isfinite(te) || (!std::isnan(tb) && !(tb < te)))
            return SIMINF_ERR_INVALID_MODEL;
    }

    /* A plan for n_particles replicates of the particle, two columns,
     * whose days span all intervals: the scheduled events are laid out
     * once, and each interval runs its own days and only the events due
     * after the previous interval's end (pfilter_events and
     * SimInf_split_events, R/pfilter.R:200-220, SimInf_particle_filter.c:
     * 80-118). */
    SimInf_solver_args b = *a;
    const double t_first = std::isnan(tspan2[0]) ? tspan2[1] : tspan2[0];
    const double t_last = tspan2[2 * (n_intervals - 1) + 1];
    double ts0[2] = {t_first, std::max(t_last, t_first + 1.0)};
    int dummy_u = 0;
    double dummy_v = 0.0;
    b.Nrep = n_particles;
    b.solver = 0;
    b.tspan = ts0;
    b.tlen = 2;
    b.U = &dummy_u;
    b.V = &dummy_v;
    b.irU = b.jcU = b.irV = b.jcV = nullptr;
    b.prU = b.prV = nullptr;
    SimInf_hip_plan *P = nullptr;
    int err = plan_create_impl(&P, &b, model, true);
    if (err)
        return err;
    if (P->host_err_key != NO_ERROR_KEY) {
        err = key_to_code(P->host_err_key);
        SimInf_hip_plan_destroy(P);
        return err;
    }

    const int Nn = a->Nn, Nc = a->Nc, Nd = a->Nd;
    const long long lanes = (long long)Nn * n_particles;
    DeviceBuffers &B = P->buf;
    int *u_cur = B.alloc<int>((size_t)lanes * Nc);
    double *v_cur = B.alloc<double>((size_t)lanes * std::max(Nd, 1));
    double *w = B.alloc<double>((size_t)n_particles);
    double *c_global =
        n_particles > SIMINF_PF_LDS ? B.alloc<double>((size_t)n_particles) : nullptr;
    int *anc = B.alloc<int>((size_t)n_particles * n_intervals);
    int *obs_d = B.upload(obs, (size_t)n_intervals * obs_nodes);
    double *ts_d = B.upload(tspan2, (size_t)2 * n_intervals);
    double *ll_d = B.alloc<double>(1);
    double *ess_d = B.alloc<double>((size_t)n_intervals);
    /* Errors per interval (solver keys and filter codes), so that the
     * first interval that failed is the one reported. */
    std::vector<unsigned long long> keys(n_intervals, NO_ERROR_KEY);
    std::vector<int> perrs(n_intervals, 0);
    unsigned long long *keys_d = B.upload(keys.data(), keys.size());
    if (P->p.rep_err) {
        /* Each particle's error key of a grid-wide coupled plan. */
        const std::vector<unsigned long long> none((size_t)n_particles, NO_ERROR_KEY);
        B.copy(P->p.rep_err, none.data(), none.size() * sizeof(none[0]),
               hipMemcpyHostToDevice);
    }
    int *perr_d = B.upload(perrs.data(), perrs.size());
    const double zero = 0.0;
    if (B.status != hipSuccess ||
        hipMemcpyAsync(ll_d, &zero, sizeof(zero), hipMemcpyHostToDevice, P->stream) !=
            hipSuccess) {
        SimInf_hip_plan_destroy(P);
        return SIMINF_ERR_DEVICE;
    }

    const unsigned grid = (unsigned)((lanes + 255) / 256);
    for (int i = 0; i < n_intervals; i++) {
        KParams q = P->p;
        q.mode = NODES_ALL;
        q.frozen_from = nullptr;
        q.u0 = i == 0 ? P->p.u0 : u_cur;
        q.v0 = i == 0 ? P->p.v0 : v_cur;
        const double tb = tspan2[2 * i], te = tspan2[2 * i + 1];
        if (std::isnan(tb)) {
            q.tlen = 1;
            q.tspan = ts_d + 2 * i + 1;
            q.tt0 = te;
        } else {
            q.tlen = 2;
            q.tspan = ts_d + 2 * i;
            q.tt0 = tb;
        }
        q.next0 = std::floor(q.tt0) + 1.0;
        q.seed = siminf_interval_seed((uint64_t)seed, i);
        q.err_key = keys_d + i;
        q.day0 = (int)(std::floor(q.tt0) - std::floor(t_first));
        if (i > 0) {
            const double t_prev = tspan2[2 * (i - 1) + 1];
            q.ev_floor = (int)t_prev;
            q.e2_floor = (int)(t_prev - std::floor(t_first));
        }
        if (P->kernel == K_SSM_NODES) {
            model->ssm_nodes(q, P->has_e1, P->canon, grid, 256, P->stream);
        } else if (P->kernel == K_COUPLED_REPS) {
            model->coupled(q, P->canon, false, P->grid, P->block, P->lds, P->stream, nullptr);
        } else {
            (void)launch_grid_reps(P, q, P->stream);
        }
        hipLaunchKernelGGL(k_pf_weights, dim3(n_particles), dim3(64), 0, P->stream, q.U,
                           Nn, Nc, q.tlen, obs_compartment, obs_dist, obs_p1, obs_p2,
                           obs_d + (size_t)i * obs_nodes, obs_nodes, w);
        if (c_global)
            hipLaunchKernelGGL(k_pf_resample<false>, dim3(1), dim3(256), 0, P->stream, w,
                               c_global, n_particles, i, (unsigned long long)seed, ll_d,
                               ess_d, anc + (size_t)i * n_particles, perr_d + i);
        else
            hipLaunchKernelGGL(k_pf_resample<true>, dim3(1), dim3(256),
                               (size_t)n_particles * sizeof(double), P->stream, w, c_global,
                               n_particles, i, (unsigned long long)seed, ll_d, ess_d,
                               anc + (size_t)i * n_particles, perr_d + i);
        hipLaunchKernelGGL(k_pf_gather, dim3(grid), dim3(256), 0, P->stream, q.U, q.V,
                           Nn, Nc, Nd, q.tlen, anc + (size_t)i * n_particles, lanes,
                           u_cur, v_cur);
    }
    if (hipGetLastError() != hipSuccess) {
        SimInf_hip_plan_destroy(P);
        return SIMINF_ERR_DEVICE;
    }
    hipError_t e = hipMemcpyAsync(keys.data(), keys_d, keys.size() * sizeof(keys[0]),
                                  hipMemcpyDeviceToHost, P->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(perrs.data(), perr_d, perrs.size() * sizeof(int),
                           hipMemcpyDeviceToHost, P->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(loglik, ll_d, sizeof(double), hipMemcpyDeviceToHost, P->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ess, ess_d, sizeof(double) * n_intervals, hipMemcpyDeviceToHost,
                           P->stream);
    if (e == hipSuccess && ancestors)
        e = hipMemcpyAsync(ancestors, anc, sizeof(int) * (size_t)n_particles * n_intervals,
                           hipMemcpyDeviceToHost, P->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(P->stream);
    SimInf_hip_plan_destroy(P);
    if (e != hipSuccess)
        return SIMINF_ERR_DEVICE;
    for (int i = 0; i < n_intervals; i++) {
        if (keys[i] != NO_ERROR_KEY)
            return key_to_code(keys[i]);
        if (perrs[i])
            return perrs[i];
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* On-device summaries of a resident trajectory (R/prevalence.R:24-110, */
/* SimInf_trajectory.c:566-650).                                        */
/* ------------------------------------------------------------------ */

/* A validated condition program, passed by value in the kernel
 * arguments (uniform for all lanes). */
struct CondProgram {
    int n;
    int op[SIMINF_COND_MAX];
    double arg[SIMINF_COND_MAX];
};

/* The condition for one node's state.  Values are doubles (exact for the
 * integer sums and products of compartment counts); comparisons and the
 * logical operators give 1 or 0, and nonzero counts as TRUE, as R's
 * coercion of numbers to logicals.  Each lane's stack is a column of an
 * LDS array. */
__device__ __forceinline__ bool cond_holds(const CondProgram &c, const int *u,
                                           double (*stk)[256], int lane)
{
    if (c.n == 0)
        return true;
    int sp = 0;
    for (int k = 0; k < c.n; k++) {
        const int op = c.op[k];
        if (op == SIMINF_OP_U) {
            stk[sp++][lane] = (double)u[(int)c.arg[k]];
            continue;
        }
        if (op == SIMINF_OP_CONST) {
            stk[sp++][lane] = c.arg[k];
            continue;
        }
        if (op == SIMINF_OP_NEG) {
            stk[sp - 1][lane] = -stk[sp - 1][lane];
            continue;
        }
        if (op == SIMINF_OP_NOT) {
            stk[sp - 1][lane] = stk[sp - 1][lane] == 0.0 ? 1.0 : 0.0;
            continue;
        }
        sp--;
        const double b = stk[sp][lane];
        const double a = stk[sp - 1][lane];
        double r;
        switch (op) {
        case SIMINF_OP_ADD: r = a + b; break;
        case SIMINF_OP_SUB: r = a - b; break;
        case SIMINF_OP_MUL: r = a * b; break;
        case SIMINF_OP_DIV: r = a / b; break;
        case SIMINF_OP_EQ: r = a == b; break;
        case SIMINF_OP_NE: r = a != b; break;
        case SIMINF_OP_LT: r = a < b; break;
        case SIMINF_OP_LE: r = a <= b; break;
        case SIMINF_OP_GT: r = a > b; break;
        case SIMINF_OP_GE: r = a >= b; break;
        case SIMINF_OP_AND: r = (a != 0.0) && (b != 0.0); break;
        default: r = (a != 0.0) || (b != 0.0); break;
        }
        stk[sp - 1][lane] = r;
    }
    return stk[0][lane] != 0.0;
}

/* Cases and population of one node's state. */
__device__ __forceinline__ void node_counts(const int *u, int Nc, unsigned cases_mask,
                                            unsigned pop_mask, long long &x, long long &y)
{
    x = 0;
    y = 0;
    for (int c = 0; c < Nc; c++) {
        const int v = u[c];
        if ((cases_mask >> c) & 1u)
            x += v;
        if ((pop_mask >> c) & 1u)
            y += v;
    }
}

/* Levels 1 and 2: one workgroup per U column; exact integer sums. */
__global__ __launch_bounds__(256) void k_prevalence_cols(const int *U, int Nn, int Nc,
                                                         unsigned cases_mask,
                                                         unsigned pop_mask, CondProgram cond,
                                                         int level, const int *index,
                                                         int n_sel, double *out)
{
    __shared__ double stk[SIMINF_COND_STACK][256];
    __shared__ long long sa[256], sb[256];
    const int lane = (int)threadIdx.x;
    const long long col = blockIdx.x;
    long long a = 0, b = 0;
    for (int k = lane; k < n_sel; k += 256) {
        const int node = index ? index[k] : k;
        const int *u = U + (col * Nn + node) * (long long)Nc;
        long long x, y;
        node_counts(u, Nc, cases_mask, pop_mask, x, y);
        if (!cond_holds(cond, u, stk, lane))
            x = y = 0;
        if (level == 1) {
            a += x;
            b += y;
        } else {
            a += x > 0;
            b += y > 0;
        }
    }
    sa[lane] = a;
    sb[lane] = b;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (lane < h) {
            sa[lane] += sa[lane + h];
            sb[lane] += sb[lane + h];
        }
        __syncthreads();
    }
    if (lane == 0)
        out[col] = (double)sa[0] / (double)sb[0];
}

/* Level 3: one lane per (selected node, U column). */
__global__ __launch_bounds__(256) void k_prevalence_nodes(const int *U, int Nn, int Nc,
                                                          unsigned cases_mask,
                                                          unsigned pop_mask,
                                                          CondProgram cond,
                                                          const int *index, int n_sel,
                                                          long long total, double *out)
{
    __shared__ double stk[SIMINF_COND_STACK][256];
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= total)
        return;
    const long long col = g / n_sel;
    const int k = (int)(g - col * n_sel);
    const int node = index ? index[k] : k;
    const int *u = U + (col * Nn + node) * (long long)Nc;
    long long x, y;
    node_counts(u, Nc, cases_mask, pop_mask, x, y);
    if (!cond_holds(cond, u, stk, (int)threadIdx.x))
        x = y = 0;
    out[g] = (double)x / (double)y;
}

/* Selected rows (compartments of U, or continuous states of V) of
 * selected nodes, one lane per (node, column). */
template <typename T>
__global__ __launch_bounds__(256) void k_trajectory(const T *U, int Nn, int Nc,
                                                    unsigned mask, int ncomp,
                                                    const int *index, int n_sel,
                                                    long long total, T *out)
{
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= total)
        return;
    const long long col = g / n_sel;
    const int k = (int)(g - col * n_sel);
    const int node = index ? index[k] : k;
    const T *u = U + (col * Nn + node) * (long long)Nc;
    T *o = out + g * ncomp;
    int j = 0;
    for (int c = 0; c < Nc; c++)
        if ((mask >> c) & 1u)
            o[j++] = u[c];
}

static int summary_checks(const SimInf_hip_plan *P, const int *index, int n_index)
{
    if (!P)
        return SIMINF_ERR_INVALID_MODEL;
    if (!P->has_run)
        return SIMINF_ERR_EMPTY_TRAJECTORY;
    if (!P->dense_U)
        return SIMINF_ERR_INVALID_SUMMARY;
    if (index) {
        if (n_index < 1)
            return SIMINF_ERR_INVALID_SUMMARY;
        for (int k = 0; k < n_index; k++)
            if (index[k] < 0 || index[k] >= P->Nn || (k > 0 && index[k] <= index[k - 1]))
                return SIMINF_ERR_INVALID_SUMMARY;
    }
    return 0;
}

static unsigned compartment_bits(int Nc)
{
    return Nc >= 32 ? 0xffffffffu : ((1u << Nc) - 1u);
}

extern "C" int SimInf_hip_plan_fuse_prevalence(SimInf_hip_plan *P, unsigned int cases_mask,
                                               unsigned int pop_mask)
{
    if (!P)
        return SIMINF_ERR_INVALID_MODEL;
    const unsigned all = compartment_bits(P->Nc);
    if (!cases_mask || !pop_mask || (cases_mask & ~all) || (pop_mask & ~all))
        return SIMINF_ERR_INVALID_SUMMARY;
    if (!P->fp_sum) {
        if ((P->kernel != K_SSM_NODES && P->kernel != K_AEM_NODES) || !P->dense_U)
            return 1;   /* not fusable: the plan keeps its trajectory */
        const size_t n1 = 2 * (size_t)P->Nrep * P->tlen;
        const size_t n2 = 2 * (size_t)P->Nrep * (P->tlen + 1);
        long long *sum = P->buf.alloc<long long>(n1);
        long long *delta = P->buf.alloc<long long>(n2);
        if (P->buf.status != hipSuccess || hipStreamSynchronize(P->stream) != hipSuccess)
            return SIMINF_ERR_DEVICE;
        P->buf.free_one(P->p.U);
        P->p.U = nullptr;
        P->fp_had_V = P->dense_V;
        if (P->p.V)
            P->buf.free_one(P->p.V);
        P->p.V = nullptr;
        P->dense_U = P->dense_V = false;
        if (hipStreamSynchronize(P->stream) != hipSuccess)
            return SIMINF_ERR_DEVICE;
        P->fp_sum = P->p.fp_sum = sum;
        P->fp_delta = P->p.fp_delta = delta;
    }
    P->p.fp_cases = cases_mask;
    P->p.fp_pop = pop_mask;
    P->has_run = false;
    P->param_version++;   /* a batch holding the plan captures again */
    return 0;
}

/* Level-1 prevalence of a fused plan: the formula it was fused with, every
 * node, no condition. */
static int fused_prevalence(SimInf_hip_plan *P, unsigned cases_mask, unsigned pop_mask,
                            int cond_len, int level, const int *index, double *out)
{
    if (!P->has_run)
        return SIMINF_ERR_EMPTY_TRAJECTORY;
    if (!out || level != 1 || cond_len != 0 || index || cases_mask != P->p.fp_cases ||
        pop_mask != P->p.fp_pop)
        return SIMINF_ERR_INVALID_SUMMARY;
    const size_t nout = (size_t)P->Nrep * P->tlen;
    DeviceBuffers T;
    T.stream = P->stream;
    T.pooled = P->buf.pooled;
    double *out_d = T.alloc<double>(nout);
    if (T.status != hipSuccess)
        return SIMINF_ERR_DEVICE;
    hipLaunchKernelGGL(k_fused_finalize, dim3((unsigned)((P->Nrep + 63) / 64)), dim3(64), 0,
                       P->stream, P->fp_sum, P->fp_delta, P->Nrep, P->tlen, out_d);
    if (hipGetLastError() != hipSuccess)
        return SIMINF_ERR_DEVICE;
    T.copy(out, out_d, nout * sizeof(double), hipMemcpyDeviceToHost);
    return T.status == hipSuccess ? 0 : SIMINF_ERR_DEVICE;
}

extern "C" int SimInf_hip_plan_prevalence(SimInf_hip_plan *P, unsigned int cases_mask,
                                          unsigned int pop_mask, const int *cond_ops,
                                          const double *cond_args, int cond_len, int level,
                                          const int *index, int n_index, double *out)
{
    if (P && P->fp_sum)
        return fused_prevalence(P, cases_mask, pop_mask, cond_len, level, index, out);
    int err = summary_checks(P, index, n_index);
    if (err)
        return err;
    const unsigned all = compartment_bits(P->Nc);
    if (!out || level < 1 || level > 3 || !cases_mask || !pop_mask ||
        (cases_mask & ~all) || (pop_mask & ~all) || cond_len < 0 ||
        cond_len > SIMINF_COND_MAX || (cond_len && (!cond_ops || !cond_args)))
        return SIMINF_ERR_INVALID_SUMMARY;
    CondProgram cond;
    memset(&cond, 0, sizeof(cond));
    cond.n = cond_len;
    int depth = 0;
    for (int k = 0; k < cond_len; k++) {
        const int op = cond_ops[k];
        const double arg = cond_args[k];
        if (op == SIMINF_OP_U) {
            if (!(arg >= 0.0 && arg < (double)P->Nc && arg == std::floor(arg)))
                return SIMINF_ERR_INVALID_SUMMARY;
            depth++;
        } else if (op == SIMINF_OP_CONST) {
            depth++;
        } else if (op == SIMINF_OP_NEG || op == SIMINF_OP_NOT) {
            if (depth < 1)
                return SIMINF_ERR_INVALID_SUMMARY;
        } else if (op >= SIMINF_OP_ADD && op <= SIMINF_OP_OR) {
            if (depth < 2)
                return SIMINF_ERR_INVALID_SUMMARY;
            depth--;
        } else {
            return SIMINF_ERR_INVALID_SUMMARY;
        }
        if (depth > SIMINF_COND_STACK)
            return SIMINF_ERR_INVALID_SUMMARY;
        cond.op[k] = op;
        cond.arg[k] = arg;
    }
    if (cond_len && depth != 1)
        return SIMINF_ERR_INVALID_SUMMARY;

    const int n_sel = index ? n_index : P->Nn;
    const long long ncols = (long long)P->Nrep * P->tlen;
    const size_t nout = level == 3 ? (size_t)n_sel * (size_t)ncols : (size_t)ncols;
    DeviceBuffers T;
    T.stream = P->stream;
    T.pooled = P->buf.pooled;
    const int *idx_d = index ? T.upload(index, (size_t)n_index) : nullptr;
    double *out_d = T.alloc<double>(nout);
    if (T.status != hipSuccess)
        return SIMINF_ERR_DEVICE;
    if (level < 3)
        hipLaunchKernelGGL(k_prevalence_cols, dim3((unsigned)ncols), dim3(256), 0, P->stream,
                           P->p.U, P->Nn, P->Nc, cases_mask, pop_mask, cond, level, idx_d,
                           n_sel, out_d);
    else
        hipLaunchKernelGGL(k_prevalence_nodes, dim3((unsigned)((nout + 255) / 256)), dim3(256),
                           0, P->stream, P->p.U, P->Nn, P->Nc, cases_mask, pop_mask, cond,
                           idx_d, n_sel, (long long)nout, out_d);
    if (hipGetLastError() != hipSuccess)
        return SIMINF_ERR_DEVICE;
    T.copy(out, out_d, nout * sizeof(double), hipMemcpyDeviceToHost);
    return T.status == hipSuccess ? 0 : SIMINF_ERR_DEVICE;
}

/* Gather of the selected rows of the resident dense U (T int, Nr = Nc)
 * or V (T double, Nr = Nd) for trajectory(..., format = "matrix"). */
template <typename T>
static int plan_trajectory(SimInf_hip_plan *P, const T *src, int Nr, unsigned int comp_mask,
                           const int *index, int n_index, T *out)
{
    const int ncomp = __builtin_popcount(comp_mask);
    const int n_sel = index ? n_index : P->Nn;
    const long long total = (long long)n_sel * P->Nrep * P->tlen;
    DeviceBuffers T_;
    T_.stream = P->stream;
    T_.pooled = P->buf.pooled;
    const int *idx_d = index ? T_.upload(index, (size_t)n_index) : nullptr;
    T *out_d = T_.alloc<T>((size_t)total * ncomp);
    if (T_.status != hipSuccess)
        return SIMINF_ERR_DEVICE;
    hipLaunchKernelGGL(k_trajectory<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       P->stream, src, P->Nn, Nr, comp_mask, ncomp, idx_d, n_sel, total,
                       out_d);
    if (hipGetLastError() != hipSuccess)
        return SIMINF_ERR_DEVICE;
    T_.copy(out, out_d, (size_t)total * ncomp * sizeof(T), hipMemcpyDeviceToHost);
    return T_.status == hipSuccess ? 0 : SIMINF_ERR_DEVICE;
}

extern "C" int SimInf_hip_plan_trajectory_v(SimInf_hip_plan *P, unsigned int state_mask,
                                            const int *index, int n_index, double *out)
{
    if (!P)
        return SIMINF_ERR_INVALID_MODEL;
    if (!P->has_run)
        return SIMINF_ERR_EMPTY_TRAJECTORY;
    if (!P->dense_V || P->Nd < 1 || P->Nd > 32 || !out || !state_mask ||
        (state_mask & ~compartment_bits(P->Nd)))
        return SIMINF_ERR_INVALID_SUMMARY;
    if (index) {
        if (n_index < 1)
            return SIMINF_ERR_INVALID_SUMMARY;
        for (int k = 0; k < n_index; k++)
            if (index[k] < 0 || index[k] >= P->Nn || (k > 0 && index[k] <= index[k - 1]))
                return SIMINF_ERR_INVALID_SUMMARY;
    }
    return plan_trajectory<double>(P, P->p.V, P->Nd, state_mask, index, n_index, out);
}

extern "C" int SimInf_hip_plan_trajectory(SimInf_hip_plan *P, unsigned int comp_mask,
                                          const int *index, int n_index, int *out)
{
    int err = summary_checks(P, index, n_index);
    if (err)
        return err;
    if (!out || !comp_mask || (comp_mask & ~compartment_bits(P->Nc)))
        return SIMINF_ERR_INVALID_SUMMARY;
    return plan_trajectory<int>(P, P->p.U, P->Nc, comp_mask, index, n_index, out);
}

/* ------------------------------------------------------------------ */
/* ABC-SMC weights (SimInf_abc_weights, SimInf_abc.c:283-406).          */
/* ------------------------------------------------------------------ */

/* z = L^-1 x for every particle (columns of a particle matrix). */
__global__ __launch_bounds__(256) void k_abc_transform(const double *x, int n, int d,
                                                       const double *L, double *z)
{
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i < n)
        siminf_forward_subst(x + i, n, L, d, z + i, n);
}

/* Unnormalised log-weight of each current particle: the sum of its prior
 * log-densities less the log of sum_j w_j N(xx_i; x_j, sigma), the
 * kernel sum in particle order; flag 1 for a non-finite prior term, 2 for
 * a kernel sum that is not finite and positive. */
__global__ __launch_bounds__(256) void k_abc_logw(const double *xx, const double *x,
                                                  const double *zc, const double *zp,
                                                  const double *w, int n, int d,
                                                  const int *kind, const double *p1,
                                                  const double *p2, const double *lg,
                                                  double lognorm, double *logw, int *flag)
{
    const int i = (int)(blockIdx.x * 256 + threadIdx.x);
    if (i >= n)
        return;
    double lp = 0.0;
    for (int k = 0; k < d; k++)
        lp = lp + siminf_prior_logdens(kind[k], xx[(long long)k * n + i],
                                       x[(long long)k * n + i], p1[k], p2[k], lg[k]);
    if (!(lp > -INFINITY && lp < INFINITY)) {
        logw[i] = 0.0;
        flag[i] = 1;
        return;
    }
    double sum = 0.0;
    for (int j = 0; j < n; j++)
        sum = sum + w[j] * siminf_exp(siminf_mvn_logpdf_z(zc + i, n, zp + j, n, d, lognorm));
    if (!(sum > 0.0 && sum < INFINITY)) {
        logw[i] = 0.0;
        flag[i] = 2;
        return;
    }
    logw[i] = lp - siminf_log(sum);
    flag[i] = 0;
}

/* Normalisation in one workgroup with the reference's sequence: the
 * maximum starts at 0, exp(ww - max) summed in particle order on one
 * lane, then divided; or the error of the first flagged particle. */
__global__ __launch_bounds__(256) void k_abc_normalize(double *ww, const int *flag, int n,
                                                       int *err)
{
    __shared__ double s_max[256];
    __shared__ int s_first[256];
    __shared__ double s_sum;
    const int tid = (int)threadIdx.x;
    double m = 0.0;
    int first = INT_MAX;
    for (int i = tid; i < n; i += 256) {
        if (flag[i] && i < first)
            first = i;
        m = ww[i] > m ? ww[i] : m;
    }
    s_max[tid] = m;
    s_first[tid] = first;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
            s_max[tid] = s_max[tid + h] > s_max[tid] ? s_max[tid + h] : s_max[tid];
            s_first[tid] = s_first[tid + h] < s_first[tid] ? s_first[tid + h] : s_first[tid];
        }
        __syncthreads();
    }
    if (s_first[0] != INT_MAX) {
        if (tid == 0)
            *err = flag[s_first[0]] == 1 ? SIMINF_ERR_INVALID_WEIGHT : SIMINF_ERR_ABC_WEIGHTS;
        return;
    }
    const double mx = s_max[0];
    for (int i = tid; i < n; i += 256)
        ww[i] = siminf_exp(ww[i] - mx);
    __syncthreads();
    if (tid == 0)
        s_sum = pf_serial_sum(ww, n);
    __syncthreads();
    const double sum = s_sum;
    for (int i = tid; i < n; i += 256)
        ww[i] = ww[i] / sum;
}

/* The host part of the ABC weights: argument checks, each prior's kind
 * and log-gamma term, and the Cholesky factor L of sigma with its
 * log-normaliser.  Returns 0, or an error code; *uniform set when w is
 * NULL (every weight 1/n, already written to ww). */
static int abc_prepare(const char *distribution, const double *p1, const double *p2, int n,
                       int d, const double *x, const double *xx, const double *w,
                       const double *sigma, double *ww, std::vector<int> &kind,
                       std::vector<double> &lg, std::vector<double> &L, double *lognorm,
                       bool *uniform)
{
    *uniform = false;
    if (n < 1 || !ww)
        return SIMINF_ERR_INVALID_MODEL;
    if (!w) {
        for (int i = 0; i < n; i++)
            ww[i] = 1.0 / (double)n;
        *uniform = true;
        return 0;
    }
    if (d < 1 || d > SIMINF_ABC_MAX_PARS || !distribution || !p1 || !p2 || !x || !xx ||
        !sigma)
        return SIMINF_ERR_INVALID_MODEL;
    kind.assign(d, 0);
    lg.assign(d, 0.0);
    for (int k = 0; k < d; k++) {
        kind[k] = distribution[k];
        if (kind[k] != SIMINF_PRIOR_GAMMA && kind[k] != SIMINF_PRIOR_LOGNORMAL &&
            kind[k] != SIMINF_PRIOR_NORMAL && kind[k] != SIMINF_PRIOR_UNIFORM)
            return SIMINF_ERR_UNKNOWN_DISTRIBUTION;
        if (kind[k] == SIMINF_PRIOR_GAMMA)
            lg[k] = std::lgamma(p1[k]);
    }
    L.assign((size_t)d * d, 0.0);
    *lognorm = 0.0;
    if (siminf_cholesky(sigma, d, L.data(), lognorm))
        return SIMINF_ERR_ABC_WEIGHTS;
    return 0;
}

/* Device operands of the ABC weight kernels. */
struct AbcOperands {
    const double *x, *xx, *w, *L, *p1, *p2, *lg;
    const int *kind;
    double *zc, *zp, *logw;
    int *flag, *err;
};

/* The weight kernels on stream st over operands already in device memory
 * (err cleared); then the error code and the weights copied to the host,
 * complete on return. */
static int abc_launch(const AbcOperands &o, int n, int d, double lognorm, double *ww,
                      hipStream_t st)
{
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_abc_transform, dim3(grid), dim3(256), 0, st, o.xx, n, d, o.L, o.zc);
    hipLaunchKernelGGL(k_abc_transform, dim3(grid), dim3(256), 0, st, o.x, n, d, o.L, o.zp);
    hipLaunchKernelGGL(k_abc_logw, dim3(grid), dim3(256), 0, st, o.xx, o.x, o.zc, o.zp, o.w,
                       n, d, o.kind, o.p1, o.p2, o.lg, lognorm, o.logw, o.flag);
    hipLaunchKernelGGL(k_abc_normalize, dim3(1), dim3(256), 0, st, o.logw, o.flag, n, o.err);
    if (hipGetLastError() != hipSuccess)
        return SIMINF_ERR_DEVICE;
    int err = 0;
    if (hipMemcpyAsync(&err, o.err, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(ww, o.logw, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return SIMINF_ERR_DEVICE;
    return err;
}

extern "C" int SimInf_hip_abc_weights(const char *distribution, const double *p1,
                                      const double *p2, int n_particles, int n_parameters,
                                      const double *x, const double *xx, const double *w,
                                      const double *sigma, double *ww)
{
    const int n = n_particles, d = n_parameters;
    std::vector<int> kind;
    std::vector<double> lg, L;
    double lognorm = 0.0;
    bool uniform = false;
    const int perr = abc_prepare(distribution, p1, p2, n, d, x, xx, w, sigma, ww, kind, lg,
                                 L, &lognorm, &uniform);
    if (perr || uniform)
        return perr;

    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        return SIMINF_ERR_DEVICE;
    int rc = 0;
    {
        int dev = 0;
        (void)hipGetDevice(&dev);
        DeviceBuffers B;
        B.stream = st;
        B.pooled = pool_allocation(dev);
        const size_t nd = (size_t)n * d;
        AbcOperands o;
        o.x = B.upload(x, nd);
        o.xx = B.upload(xx, nd);
        o.w = B.upload(w, (size_t)n);
        o.L = B.upload(L.data(), L.size());
        o.kind = B.upload(kind.data(), kind.size());
        o.p1 = B.upload(p1, (size_t)d);
        o.p2 = B.upload(p2, (size_t)d);
        o.lg = B.upload(lg.data(), lg.size());
        o.zc = B.alloc<double>(nd);
        o.zp = B.alloc<double>(nd);
        o.logw = B.alloc<double>((size_t)n);
        o.flag = B.alloc<int>((size_t)n);
        const int zero = 0;
        o.err = B.upload(&zero, 1);
        rc = B.status != hipSuccess ? SIMINF_ERR_DEVICE : abc_launch(o, n, d, lognorm, ww, st);
        B.release();
    }
    (void)hipStreamDestroy(st);
    return rc;
}

/* ------------------------------------------------------------------ */
/* ABC weights without allocation.                                     */
/* ------------------------------------------------------------------ */

/* Device memory of SimInf_hip_abc_weights_noalloc: a fixed array, one
 * copy per GPU, held until the process ends.  Host code never passes the
 * array's name to a kernel or a copy (on the host it is only a handle to
 * the symbol): it uses the address hipGetSymbolAddress returns for the
 * current device. */
static __device__ __align__(256) unsigned char scratch[SIMINF_NOALLOC_SCRATCH_BYTES];

/* The scratch array's address on device dev (looked up once per device,
 * with dev current), or nullptr. */
static unsigned char *scratch_on(int dev)
{
    static std::mutex mu;
    static void *ptr[64];
    if (dev < 0 || dev >= 64)
        return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return nullptr;
        ptr[dev] = p;
    }
    return (unsigned char *)ptr[dev];
}

/* Bytes of scratch the weights of n particles and d parameters take:
 * x, xx, zc, zp (n*d doubles each), w, logw (n doubles), L (d*d), p1,
 * p2, lg (d doubles), kind (d ints), flag (n ints), err (one int), each
 * at a 256-byte boundary. */
static size_t abc_scratch_bytes(long long n, long long d)
{
    auto a = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t nd = (size_t)(n * d);
    return 4 * a(nd * 8) + 2 * a((size_t)n * 8) + a((size_t)(d * d) * 8) + 3 * a((size_t)d * 8) +
           a((size_t)d * 4) + a((size_t)n * 4) + a(4);
}

extern "C" int SimInf_hip_abc_weights_noalloc(const char *distribution, const double *p1,
                                              const double *p2, int n_particles,
                                              int n_parameters, const double *x,
                                              const double *xx, const double *w,
                                              const double *sigma, double *ww)
{
    const int n = n_particles, d = n_parameters;
    std::vector<int> kind;
    std::vector<double> lg, L;
    double lognorm = 0.0;
    bool uniform = false;
    const int perr = abc_prepare(distribution, p1, p2, n, d, x, xx, w, sigma, ww, kind, lg,
                                 L, &lognorm, &uniform);
    if (perr || uniform)
        return perr;
    if (abc_scratch_bytes(n, d) > (size_t)SIMINF_NOALLOC_SCRATCH_BYTES)
        return SIMINF_ERR_ALLOC_MEMORY_BUFFER;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return SIMINF_ERR_DEVICE;
    unsigned char *base = scratch_on(dev);
    if (!base)
        return SIMINF_ERR_DEVICE;

    /* Carve the operands out of the scratch array. */
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char *q = base + off;
        off += (bytes + 255) & ~(size_t)255;
        return q;
    };
    const size_t nd = (size_t)n * d;
    double *x_d = (double *)take(nd * 8), *xx_d = (double *)take(nd * 8);
    double *w_d = (double *)take((size_t)n * 8), *L_d = (double *)take((size_t)d * d * 8);
    double *p1_d = (double *)take((size_t)d * 8), *p2_d = (double *)take((size_t)d * 8);
    double *lg_d = (double *)take((size_t)d * 8);
    int *kind_d = (int *)take((size_t)d * 4);
    AbcOperands o;
    o.x = x_d;
    o.xx = xx_d;
    o.w = w_d;
    o.L = L_d;
    o.p1 = p1_d;
    o.p2 = p2_d;
    o.lg = lg_d;
    o.kind = kind_d;
    o.zc = (double *)take(nd * 8);
    o.zp = (double *)take(nd * 8);
    o.logw = (double *)take((size_t)n * 8);
    o.flag = (int *)take((size_t)n * 4);
    o.err = (int *)take(4);

    /* Inputs into scratch on the legacy default stream; err starts at 0
     * (the previous call's value stays in the array otherwise). */
    hipStream_t st = nullptr;
    const int zero = 0;
    struct Up {
        void *dst;
        const void *src;
        size_t bytes;
    } ups[] = {{x_d, x, nd * 8},       {xx_d, xx, nd * 8},          {w_d, w, (size_t)n * 8},
               {L_d, L.data(), L.size() * 8}, {p1_d, p1, (size_t)d * 8},
               {p2_d, p2, (size_t)d * 8},     {lg_d, lg.data(), (size_t)d * 8},
               {kind_d, kind.data(), (size_t)d * 4}, {o.err, &zero, 4}};
    for (const Up &u : ups)
        if (hipMemcpyAsync(u.dst, u.src, u.bytes, hipMemcpyHostToDevice, st) != hipSuccess)
            return SIMINF_ERR_DEVICE;
    return abc_launch(o, n, d, lognorm, ww, st);
}
