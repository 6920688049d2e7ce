/*
 * siminf_hip.h -- C ABI of the MI355X (gfx950) solver core.
 *
 * The library replaces the reference's solver entry points
 *
 *     int SimInf_run_solver_ssm(SimInf_solver_args *args);
 *     int SimInf_run_solver_mssm(SimInf_solver_args *args);
 *     int SimInf_run_solver_mssm_crn(SimInf_solver_args *args);
 *     int SimInf_run_solver_aem(SimInf_solver_args *args);
 *                                    (src/SimInf_internal.h:406-420)
 *
 * and the dispatch that SimInf_run performs on them (src/SimInf.c:383-399).
 * SimInf_solver_args below has the reference's field order and meaning
 * (src/SimInf_internal.h:44-200), so the struct the reference's SimInf_run
 * fills can be passed as is.  Host function pointers cannot run on the
 * GPU, so every entry point takes a device plugin (SimInf_hip_model) that
 * holds the same TRFun/PTSFun bodies compiled for gfx950; tr_fun and
 * pts_fun in the struct are not called.
 *
 * Return values are 0 or a negative SimInf error code (SimInf.h:66-89);
 * SimInf_hip_error_message gives the reference's message for a code
 * (src/SimInf.c:28-97).
 *
 * All pointers in SimInf_solver_args are host pointers.  The entry points
 * are blocking: they stage inputs in device memory, run, and copy U/V (or
 * the sparse prU/prV) back before returning.  The SimInf_hip_plan_* calls
 * keep inputs and outputs resident in device memory across runs.
 */
#ifndef SIMINF_HIP_H
#define SIMINF_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Event types (SimInf.h:56-59). */
enum {
    SIMINF_EXIT_EVENT = 0,
    SIMINF_ENTER_EVENT = 1,
    SIMINF_INTERNAL_TRANSFER_EVENT = 2,
    SIMINF_EXTERNAL_TRANSFER_EVENT = 3
};

/* Error codes (SimInf.h:66-89). */
typedef enum {
    SIMINF_ERR_NEGATIVE_STATE       = -1,
    SIMINF_ERR_ALLOC_MEMORY_BUFFER  = -2,
    SIMINF_UNDEFINED_EVENT          = -3,
    SIMINF_INVALID_EDGE_PROBABILITY = -4,
    SIMINF_ERR_V_IS_NOT_FINITE      = -7,
    SIMINF_ERR_SAMPLE_SELECT        = -8,
    SIMINF_ERR_INVALID_MODEL        = -9,
    SIMINF_ERR_V_IS_NEGATIVE        = -10,
    SIMINF_ERR_INVALID_RATE         = -11,
    SIMINF_ERR_UNKNOWN_SOLVER       = -12,
    SIMINF_ERR_DEST_OUT_OF_BOUNDS   = -13,
    SIMINF_ERR_NODE_OUT_OF_BOUNDS   = -14,
    SIMINF_ERR_EVENTS_N             = -15,
    SIMINF_ERR_EVENT_SHIFT          = -16,
    SIMINF_ERR_SHIFT_OUT_OF_BOUNDS  = -17,
    SIMINF_ERR_INVALID_PROPORTION   = -18,
    SIMINF_ERR_AEM_REPLICATED_MODEL = -19,
    SIMINF_ERR_SPARSE_MODEL         = -20,
    SIMINF_ERR_INVALID_OPTIONS      = -21,
    SIMINF_ERR_INVALID_SEED         = -22,
    SIMINF_ERR_INVALID_CRN          = -23,
    SIMINF_ERR_INVALID_RNG_TYPE     = -24,
    /* Not in the reference: the GPU runtime failed (allocation, launch). */
    SIMINF_ERR_DEVICE               = -100,
    /* Particle filter (no code in the reference's C API: R-level errors
     * of pfilter.R:455-458 and SimInf_particle_filter.c:42-48). */
    SIMINF_ERR_INVALID_OBS_PROCESS  = -101,
    SIMINF_ERR_INVALID_WEIGHT       = -102,
    SIMINF_ERR_WEIGHT_SUM           = -103,
    /* Summaries of a plan's trajectory (R-level errors of prevalence()
     * and trajectory(), R/prevalence.R:233-244). */
    SIMINF_ERR_EMPTY_TRAJECTORY     = -104,
    SIMINF_ERR_INVALID_SUMMARY      = -105,
    /* ABC-SMC weights (SimInf_abc_error, SimInf_abc.c:30-52). */
    SIMINF_ERR_UNKNOWN_DISTRIBUTION = -106,
    SIMINF_ERR_ABC_WEIGHTS          = -107
} SimInf_error_code;

/* The most compartments (Nc), transitions (Nt) and continuous states
 * (Nd) a model may have: the library holds sets of them as 32-bit masks.
 * Plans of larger models are refused with SIMINF_ERR_INVALID_MODEL (the
 * reference has no such limit). */
#define SIMINF_MAX_SET 32

/* Plugin signatures (SimInf.h:92-107). */
typedef double (*TRFun)(const int *u, const double *v, const double *ldata,
                        const double *gdata, double t);
typedef int (*PTSFun)(double *v_new, const int *u, const double *v,
                      const double *ldata, const double *gdata, int node,
                      double t);

/* Solver arguments; field order and meaning of SimInf_internal.h:44-200. */
typedef struct SimInf_solver_args {
    const int *u0;          /* int Nc x (Nn*Nrep), column per node      */
    const double *v0;       /* double Nd x (Nn*Nrep)                    */
    const int *irG;         /* G (Nt x Nt) CSC row indices              */
    const int *jcG;         /* G column pointers                        */
    const int *irS;         /* S (Nc x Nt) CSC row indices              */
    const int *jcS;         /* S column pointers                        */
    const int *prS;         /* S values (integers)                      */
    const double *tspan;    /* output times                             */
    int tlen;
    int *U;                 /* dense (Nn*Nc) x (Nrep*tlen) or NULL      */
    const int *irU;         /* sparse U pattern when U == NULL          */
    const int *jcU;
    double *prU;
    double *V;              /* dense (Nn*Nd) x (Nrep*tlen) or NULL      */
    const int *irV;
    const int *jcV;
    double *prV;
    const double *ldata;    /* double Nld x Nn (shared by replicates)   */
    const double *gdata;
    int Nn;
    int Nc;
    int Nt;
    int Nd;
    int Nld;
    int Nrep;
    const int *irE;         /* E (Nc x Nselect) CSC                     */
    const int *jcE;
    double *prE;
    const int *N;           /* shift matrix Nc x Nshift or NULL         */
    int len;                /* number of scheduled events               */
    const int *event;
    const int *time;
    const int *node;        /* one-based                                */
    const int *dest;        /* one-based                                */
    const int *n;
    const double *proportion;
    const int *select;      /* one-based                                */
    const int *shift;       /* one-based                                */
    int Nthread;
    int solver;             /* 0 = ssm, 1 = aem                         */
    int crn;
    int rng_type;
    unsigned long int seed;
    TRFun *tr_fun;
    PTSFun pts_fun;
} SimInf_solver_args;

/* ------------------------------------------------------------------ */
/* Device plugins.                                                     */
/* ------------------------------------------------------------------ */

/* Opaque handle of a model compiled for gfx950. */
typedef struct SimInf_hip_model SimInf_hip_model;

/* Built-in device plugins, by the reference's model names: "SIR",
 * "SEIR", "SIS", "SISe", "SISe3", "SISe_sp", "SISe3_sp" (src/models/SIR.c
 * etc.; the spatial models' post-time-step functions read the
 * neighbours listed in ldata, SISe_sp.c:90-138) and "SIR_mparse" (the
 * code mparse emits for S -> beta*S*I/N -> I, I -> gamma*I -> R,
 * N <- S+I+R).  NULL if the name is unknown. */
const SimInf_hip_model *SimInf_hip_model_builtin(const char *name);

/* Dimensions the plugin was compiled for (Nc, Nt, Nd). */
int SimInf_hip_model_dims(const SimInf_hip_model *model,
                          int *Nc, int *Nt, int *Nd);

/* ------------------------------------------------------------------ */
/* Blocking entry points (SimInf_internal.h:406-420, SimInf.c:383-399). */
/* ------------------------------------------------------------------ */

/* ssm: one replicate (Nrep == 1); replaces SimInf_run_solver_ssm. */
int SimInf_hip_run_solver_ssm(SimInf_solver_args *args,
                              const SimInf_hip_model *model);

/* mssm: Nrep >= 1 replicates; replaces SimInf_run_solver_mssm. */
int SimInf_hip_run_solver_mssm(SimInf_solver_args *args,
                               const SimInf_hip_model *model);

/* mssm with common random numbers; replaces SimInf_run_solver_mssm_crn.
 * Every (replicate, node) has its own stream in this library, which is
 * the property crn asks for, so it runs the same code as mssm. */
int SimInf_hip_run_solver_mssm_crn(SimInf_solver_args *args,
                                   const SimInf_hip_model *model);

/* aem: one replicate; replaces SimInf_run_solver_aem. */
int SimInf_hip_run_solver_aem(SimInf_solver_args *args,
                              const SimInf_hip_model *model);

/* Dispatch of SimInf.c:383-399: ssm -> mssm_crn / mssm / ssm by Nrep and
 * crn; aem -> aem, or SIMINF_ERR_AEM_REPLICATED_MODEL if Nrep > 1. */
int SimInf_hip_run(SimInf_solver_args *args, const SimInf_hip_model *model);

/* Reference error message for a code (SimInf.c:28-97). */
const char *SimInf_hip_error_message(int code);

/* ------------------------------------------------------------------ */
/* Device-resident plans (no counterpart in the reference: they keep   */
/* the inputs and trajectories in HBM between runs, for callers that   */
/* run one model many times).                                          */
/* ------------------------------------------------------------------ */

typedef struct SimInf_hip_plan SimInf_hip_plan;

/* Validate args, build the event schedules and stage every input in
 * device memory.  args->solver/crn select the solver as SimInf_hip_run
 * does.  The host arrays may be freed after the call. */
int SimInf_hip_plan_create(SimInf_hip_plan **plan,
                           const SimInf_solver_args *args,
                           const SimInf_hip_model *model);

/* Run the plan with a seed; blocks until the kernels have finished.  The
 * trajectory stays in device memory. */
int SimInf_hip_plan_run(SimInf_hip_plan *plan, unsigned long int seed);

/* Copy the last run's trajectory to host buffers shaped like
 * args->U / args->V (or prU / prV for sparse output).  NULL skips. */
int SimInf_hip_plan_download(SimInf_hip_plan *plan, int *U, double *V,
                             double *prU, double *prV);

/* Statistics of the last run: fired transitions (SSA events and aem
 * firings), node-days simulated and the device time of the solver
 * kernels in milliseconds (HIP events on the plan's stream). */
int SimInf_hip_plan_stats(const SimInf_hip_plan *plan,
                          long long *transitions, long long *node_days,
                          double *kernel_ms);

/* Device times of the last run split by kernel (no reference
 * counterpart): the solver kernel(s), and the kernel that writes the
 * trajectory columns of nodes whose state stopped changing (0 when the
 * plan has none), with the number of node-columns it wrote. */
int SimInf_hip_plan_kernel_times(const SimInf_hip_plan *plan, double *solver_ms,
                                 double *fill_ms, long long *fill_columns);

/* Parameters per replicate (no reference counterpart; the reference's
 * replicates share ldata and gdata): ldata, if not NULL, holds Nld x
 * ldata_nodes values for each of the plan's Nrep replicates in turn:
 * ldata_nodes Nn gives every node of every replicate its own column
 * (column replicate * Nn + node), 1 gives all nodes of a replicate the
 * same Nld values; gdata, if not NULL, holds the model's gdata values for
 * each replicate in turn.  Later runs read them; a batch of proposals
 * (ABC-SMC, PMCMC) then runs as one plan of Nrep replicates.  The first
 * call allocates device memory, later ones copy into it. */
int SimInf_hip_plan_set_replicate_data(SimInf_hip_plan *plan, const double *ldata,
                                       int ldata_nodes, const double *gdata);

/* Lanes per wavefront of the node kernels' active pass, 2^shift
 * (shift 0..6), or -1 for the default rule (about two wavefronts per
 * SIMD).  Results do not depend on it: every lane keeps its own stream
 * (no reference counterpart). */
int SimInf_hip_plan_set_lanes_per_wave(SimInf_hip_plan *plan, int shift);

/* The trajectory-flush schedule of a plan whose node kernels leave the
 * columns of nodes whose state stopped changing to a fill kernel: each
 * of its workgroups writes `cols` columns of 256 nodes (1..4096; 0 for
 * the default, 8).  Results do not depend on it (no reference
 * counterpart). */
int SimInf_hip_plan_set_fill_block(SimInf_hip_plan *plan, int cols);

/* Where the plan's last run failed (no reference counterpart; the
 * reference prints the failing node's or event's state to stderr,
 * SimInf_solver.c:550-645, :1191-1249): *code the error code (0 if the
 * run succeeded; the other outputs are then -1), *replicate, *day (the
 * run's day, 0 for the day of tspan[0]), *phase (0 initial rates, 1 the
 * SSA or aem, 2 scheduled events within a node, 3 external transfers,
 * 4 the post-time-step function) and *where: the node (zero-based) for
 * phases 0, 1 and 4, the event's index in the model's event list for
 * phases 2 and 3. */
int SimInf_hip_plan_last_error(const SimInf_hip_plan *plan, int *code, int *replicate,
                               int *day, int *phase, long long *where);

/* The columns per fill block the plan uses. */
int SimInf_hip_plan_get_fill_block(const SimInf_hip_plan *plan, int *cols);

/* For a plan whose coupled kernel spans the grid (one replicate with
 * external transfers, or more herds than one workgroup holds): give each
 * of the n herds `nodes` (zero-based, distinct, n <= 16) a wavefront of
 * its own, and pack the other herds herds_per_wavefront (8..64) to a
 * wavefront; a wavefront's lanes without a herd help its herds draw their
 * SSA attempts.  (nodes, 0, 64) is the default map.  Results do not
 * depend on it (no reference counterpart).  Returns "Invalid model." for
 * other plans or bad arguments. */
int SimInf_hip_plan_set_herd_map(SimInf_hip_plan *plan, const int *nodes, int n,
                                 int herds_per_wavefront);

/* The current map's solo herd count and herds per wavefront (0 and 64
 * for plans without a map). */
int SimInf_hip_plan_get_herd_map(const SimInf_hip_plan *plan, int *n_solo,
                                 int *herds_per_wavefront);

/* Autotuner of the launch shape (no reference counterpart): runs the plan
 * reps times (seeds seed, seed+1, ...) for each candidate, keeps the one
 * with the least mean solver time and returns it in *best_shift; ms, if
 * not NULL, gets the candidates' mean times (at most 8 entries).  Node
 * kernels: the lanes-per-wavefront shifts 0..6 (7 entries).  Grid-wide
 * coupled kernels (8 entries): herds ranked by their transitions in a
 * first run; with the 4 busiest on wavefronts of their own, the others
 * 64, 32, 16 or 8 to a wavefront; then, with the best of those, 0, 2, 8
 * or 16 herds on their own (an infeasible candidate's time is +Inf);
 * *best_shift gets the kept solo count (SimInf_hip_plan_get_herd_map
 * gives both).  Other plans: -1, nothing run.  Node kernels with a fill
 * kernel then try 4, 8, 16, 32 and 64 columns per fill block and keep
 * the least mean device time of the run (not reported in ms).  The last
 * run's trajectory is the last candidate's. */
int SimInf_hip_plan_autotune(SimInf_hip_plan *plan, unsigned long int seed, int reps,
                             int *best_shift, double *ms);

void SimInf_hip_plan_destroy(SimInf_hip_plan *plan);

/* Wait for all device work and return the memory that destroyed plans
 * left in the stream-ordered pool to the device (no reference
 * counterpart; for a caller about to exit or to hand the device to other
 * code).  Call it after every plan and batch has been destroyed. */
int SimInf_hip_release(void);

/* Batches (no reference counterpart): the runs of several plans -- any
 * models, solvers, node counts and replicate counts -- captured once in a
 * graph whose branches run concurrently, then launched as one graph per
 * SimInf_hip_batch_run.  Plans whose coupled kernel spans the grid (one
 * replicate with external transfers) need a cooperative launch and are
 * refused (SIMINF_ERR_INVALID_MODEL), as is a plan given twice.  A plan
 * must outlive the batches that hold it, and runs of a plan inside and
 * outside a batch must not overlap. */
typedef struct SimInf_hip_batch SimInf_hip_batch;

int SimInf_hip_batch_create(SimInf_hip_batch **batch, SimInf_hip_plan *const *plans,
                            int n);

/* One launch of the batch, plan i run with seeds[i]; blocks until it has
 * finished.  Each plan's results (SimInf_hip_plan_stats, _download, the
 * summaries) are those of a SimInf_hip_plan_run with that seed, bit for
 * bit.  Returns the first plan's error code in batch order, or 0. */
int SimInf_hip_batch_run(SimInf_hip_batch *batch, const unsigned long int *seeds);

/* use_graph 0: SimInf_hip_batch_run enqueues the plans' kernels directly
 * on their own streams (the same kernels and results) instead of
 * launching the graph; 1 (the default): the graph. */
int SimInf_hip_batch_set_graph(SimInf_hip_batch *batch, int use_graph);

/* Autotuner of the batch's launch shape (no reference counterpart): the
 * plans share the device, so each plan's lanes per wavefront (see
 * SimInf_hip_plan_set_lanes_per_wave) and columns per fill block (see
 * SimInf_hip_plan_set_fill_block) are chosen by coordinate descent on
 * the batch's median launch time (two passes, every candidate of one
 * plan with the others held, reps launches each, seeds from seed), then
 * the graph or the direct enqueue, whichever is faster, is kept.  best_ms (optional)
 * gets the kept shape's mean launch time.  Results do not depend on it. */
int SimInf_hip_batch_autotune(SimInf_hip_batch *batch, unsigned long int seed, int reps,
                              double *best_ms);

/* Plan i's kept lanes-per-wavefront shift (-1: the default rule) and 1 if
 * the batch launches its graph, 0 if it enqueues directly. */
int SimInf_hip_batch_get_shape(const SimInf_hip_batch *batch, int i, int *shift,
                               int *use_graph);

/* Device time of the last launch in milliseconds (HIP events around the
 * graph launch, or around the direct enqueue). */
int SimInf_hip_batch_time(const SimInf_hip_batch *batch, double *ms);

void SimInf_hip_batch_destroy(SimInf_hip_batch *batch);

/* Bootstrap particle filter with the state of every particle kept in HBM
 * between observation intervals (pfilter_internal, R/pfilter.R:315-499;
 * systematic resampling, SimInf_particle_filter.c:32-66).
 *
 * args describes one particle (Nrep 1); the filter runs n_particles
 * replicates of it.  Interval i runs the scheduled events due after the
 * end of interval i-1 and at or before its own end (pfilter_events,
 * R/pfilter.R:200-220), the first interval also those before it.  Interval i runs from tspan2[2i] to
 * tspan2[2i+1], or over the one time point tspan2[2i+1] when tspan2[2i]
 * is NaN (pfilter_tspan, R/pfilter.R:177-193).  The observation process
 * (the binomial and poisson forms of pfilter's formula, R/pfilter.R:279-295)
 * relates the observed count x to the count X of compartment
 * obs_compartment (zero-based), either per node (obs_nodes == Nn) or of
 * the sum over the nodes (obs_nodes == 1): obs_dist 0 is
 * x ~ Binomial(size = X, prob = obs_p1), obs_dist 1 is
 * x ~ Poisson(lambda = obs_p1 * X + obs_p2).  obs is n_intervals x
 * obs_nodes, and a negative count is missing.  The weight of a particle is
 * the sum of the log-probabilities of its observations.
 * Outputs: loglik (the log-likelihood estimate), ess (n_intervals
 * effective sample sizes), ancestors (n_intervals x n_particles
 * zero-based indices, or NULL).  Returns 0 or an error code. */
int SimInf_hip_pfilter(const SimInf_solver_args *args,
                       const SimInf_hip_model *model, int n_particles,
                       int n_intervals, const double *tspan2,
                       int obs_compartment, int obs_dist, double obs_p1,
                       double obs_p2, int obs_nodes, const int *obs,
                       unsigned long int seed, double *loglik,
                       double *ess, int *ancestors);

/* On-device summaries of a plan's resident dense U after a successful
 * SimInf_hip_plan_run; only the summary returns to the host.
 *
 * Prevalence (calculate_prevalence, R/prevalence.R:70-110): per node and
 * U column, cases = the sum of the compartments in cases_mask (bit c =
 * compartment c) and population = the sum of those in pop_mask, both set
 * to zero where the condition is false.  level 1: sum(cases) /
 * sum(population) per column; level 2: #(cases > 0) / #(population > 0)
 * per column; level 3: cases / population per node and column.  index
 * holds n_index zero-based nodes (sorted, unique), or is NULL for all
 * nodes.  out: Nrep*tlen values (levels 1, 2), or n_index x (Nrep*tlen)
 * column-major (level 3).  The condition is a postfix program of cond_len
 * (op, arg) pairs, at most SIMINF_COND_MAX, stack depth at most
 * SIMINF_COND_STACK, ending with one value (nonzero = TRUE); cond_len 0
 * means no condition. */
#define SIMINF_COND_MAX 64
#define SIMINF_COND_STACK 16
enum {
    SIMINF_OP_U = 0,        /* push u[arg] */
    SIMINF_OP_CONST = 1,    /* push arg */
    SIMINF_OP_ADD, SIMINF_OP_SUB, SIMINF_OP_MUL, SIMINF_OP_DIV, SIMINF_OP_NEG,
    SIMINF_OP_EQ, SIMINF_OP_NE, SIMINF_OP_LT, SIMINF_OP_LE, SIMINF_OP_GT,
    SIMINF_OP_GE, SIMINF_OP_AND, SIMINF_OP_OR, SIMINF_OP_NOT
};
int SimInf_hip_plan_prevalence(SimInf_hip_plan *plan, unsigned int cases_mask,
                               unsigned int pop_mask, const int *cond_ops,
                               const double *cond_args, int cond_len, int level,
                               const int *index, int n_index, double *out);

/* Fused level-1 prevalence (no counterpart in the reference; for callers
 * that need only prevalence(model, cases ~ population, level = 1) of every
 * run, such as ABC distances, R/abc.R:585-611): later runs of the plan
 * keep, instead of U, exact integer sums of cases_mask and pop_mask per
 * replicate and U column, and SimInf_hip_plan_prevalence with level 1,
 * the same masks, no condition and index NULL returns
 * sum(cases) / sum(population) from them -- the values the unfused plan
 * gives.  Frees the plan's dense U and V: download, trajectories and other
 * summaries then return SIMINF_ERR_INVALID_SUMMARY.  Calling it again sets
 * other masks.  Returns 0, 1 when the plan cannot fuse (not a node-kernel
 * plan, or U not dense: the plan is unchanged) or an error code. */
int SimInf_hip_plan_fuse_prevalence(SimInf_hip_plan *plan, unsigned int cases_mask,
                                    unsigned int pop_mask);

/* Trajectory of the compartments in comp_mask (in model order) for the
 * nodes in index (zero-based, or NULL for all) from the resident U, in
 * the layout of trajectory(..., format = "matrix") (trajectory_as_is,
 * R/trajectory.R:73-85, called at R/trajectory.R:220-225): rows
 * node-major, then compartment; columns the Nrep*tlen U columns;
 * column-major int32. */
int SimInf_hip_plan_trajectory(SimInf_hip_plan *plan, unsigned int comp_mask,
                               const int *index, int n_index, int *out);

/* The same for continuous states: the states in state_mask (bit i is
 * row i of v0) from the resident dense V, as trajectory(..., format =
 * "matrix") returns them when only V states are selected
 * (R/trajectory.R:227-228); column-major double.  Errors as above, and
 * SIMINF_ERR_INVALID_SUMMARY when V was not kept dense or Nd is 0. */
int SimInf_hip_plan_trajectory_v(SimInf_hip_plan *plan, unsigned int state_mask,
                                 const int *index, int n_index, double *out);

/* ABC-SMC weights of the current generation (SimInf_abc_weights,
 * SimInf_abc.c:283-406), on the device.  distribution holds one letter
 * per parameter: 'g' gamma (p1 shape, p2 rate), 'l' lognormal (p2
 * sdlog), 'n' normal (p2 sd), 'u' uniform (p1, p2 bounds); the lognormal
 * and normal priors are centred on the previous generation's particle of
 * the same row, as in the reference.  x (previous, rows ordered by the
 * current particles' ancestors) and xx (current) are n_particles x
 * n_parameters column-major; w the previous weights (NULL: every weight
 * 1/n_particles); sigma the proposal covariance (column-major).  ww gets
 * the normalised weights.  Errors: SIMINF_ERR_UNKNOWN_DISTRIBUTION,
 * SIMINF_ERR_INVALID_WEIGHT (a prior density of 0 or a non-finite one),
 * SIMINF_ERR_ABC_WEIGHTS (a kernel sum that is not finite and positive,
 * or sigma not positive definite), at the first particle in order. */
int SimInf_hip_abc_weights(const char *distribution, const double *p1,
                           const double *p2, int n_particles, int n_parameters,
                           const double *x, const double *xx, const double *w,
                           const double *sigma, double *ww);

/* Bytes of device memory SimInf_hip_abc_weights_noalloc may use. */
#define SIMINF_NOALLOC_SCRATCH_BYTES (32u << 20)

/* SimInf_hip_abc_weights without allocation: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code
 * next to graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): it calls no hipMalloc, hipFree or allocating library
 * routine, and stages its inputs, intermediates and results in a fixed
 * device array of SIMINF_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends (found with hipGetSymbolAddress on the first call on each
 * device).  Same arguments, results (bit for bit) and errors as
 * SimInf_hip_abc_weights, on the current device's legacy default stream,
 * blocking; SIMINF_ERR_ALLOC_MEMORY_BUFFER when n_particles and
 * n_parameters need more than the array (x, xx and two transformed copies
 * take 32 bytes per particle and parameter, plus 20 bytes per particle;
 * about 250,000 particles of 3 parameters).  Calls that overlap on one
 * GPU share the array and race; the caller serialises them. */
int SimInf_hip_abc_weights_noalloc(const char *distribution, const double *p1,
                                   const double *p2, int n_particles, int n_parameters,
                                   const double *x, const double *xx, const double *w,
                                   const double *sigma, double *ww);

#ifdef __cplusplus
}
#endif

#endif /* SIMINF_HIP_H */
