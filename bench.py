"""Benchmark of the MI355X solver core.

Default workload: SEIR on 1,000,000 nodes (sizes ~ Poisson(500), 0.5 % of
nodes seeded with 5 infected) over 365 days, run under both the ssm and
the aem solver in one batched launch: the two plans' kernels (classify,
frozen-node fill, active pass of each) are captured once in a HIP graph
and every step is one graph launch (SimInf_hip_batch_run) with new seeds,
inputs and trajectories resident in HBM.  Before the timed steps each
plan's autotuner picks the active pass's lanes per wavefront (results do
not depend on it).

Prints one JSON line: node-days/s of both solvers together (value),
transitions/s, p50/p99 step latency, the roofline of the kernel that moves
the step's bytes (the frozen-node fill; its time from event-record nodes
inside the graph, on the stream it runs on) and of the SSA active passes
beside it, the PCIe-inclusive time of the blocking entry point, a CPU
baseline (the C checker with the reference's OpenMP node-block partition,
on all host cores and on one), and, as extra keys, the SISe3 workload with
scheduled events (1600 herds, 1460 days) and a particle-filter request of
2000 particles with binomial observations.

--workload pfilter_stream measures the particle-filter serving loop
instead: a stream of filter requests (SIR on log-uniform node counts,
2000 particles, 30 weekly observations) with Poisson arrivals through
siminf_amd.serving.PFilterServer, reported as particle-simulations/s
with p50/p99 request latency, beside the plain path (one run() per
interval, resampling on the host) on a sample of the same distribution.
Other --workload names time one plan of one solver.
"""
import argparse
import ctypes
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(model):
    """U/V writeback plus one read of u0, v0, ldata and the events, plus a
    read-modify-write of source and destination per external transfer."""
    Nn, Nc, Nd, Nld, Nrep = model.Nn, model.Nc, model.Nd, model.Nld, model.replicates
    tlen = len(model.tspan)
    ev = model.events["event"]
    len_e2 = int((ev == 3).sum())
    return (4 * Nc * Nn * Nrep * tlen + 8 * Nd * Nn * Nrep * tlen
            + 4 * Nc * Nn * Nrep + 8 * Nd * Nn * Nrep + 8 * Nld * Nn
            + 36 * len(ev) + 2 * (2 * 4 * Nc) * len_e2)


def pmc_traffic(workload, dominant):
    """HBM bytes per launch of the dominant kernel from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs recorded in profiles/pmc_traffic.json
    (each counter scaled by its reading for a copy kernel of known size);
    None when no run was recorded for this workload and kernel."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    entry = data.get(workload)
    if entry is None:
        return None
    if dominant == "k_fill_frozen":
        entry = entry.get("fill_kernel")
        if entry is None:
            return None
    return entry.get("traffic_bytes")


def make_workload(name):
    from siminf_amd import workloads as W
    if name.startswith("sise3_herds"):
        solver = "aem" if name.endswith("aem") else "ssm"
        return W.sise3_herds(), {"workload": "SISe3 1600 herds x 1460 days, events",
                                 "solver": solver}
    if name.startswith("sise3_sp_herds"):
        solver = "aem" if name.endswith("aem") else "ssm"
        return W.sise3_sp_herds(), {"workload": "SISe3_sp 1600 herds x 1460 days, events, "
                                                "local spread of phi", "solver": solver}
    if name == "mparse_sir":
        return W.mparse_sir_readme(), {"workload": "mparse SIR 1000 nodes x 150 days",
                                       "solver": "ssm"}
    if name.startswith("seir_million"):
        solver = "aem" if name.endswith("aem") else "ssm"
        return W.seir_million(), {"workload": "SEIR 1e6 nodes x 365 days",
                                  "solver": solver}
    raise SystemExit("unknown workload %s" % name)


def host_cores():
    """Cores this process may use (its CPU affinity), capped by
    OMP_NUM_THREADS when set."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(n, 1)


def cpu_baseline(model, solvers, threads, budget_s=20.0, slice_nodes=None):
    """The C checker (the reference's loops; node loops over `threads`
    OpenMP threads in contiguous node blocks, the reference's partition)
    on the same workload, or on its first slice_nodes nodes: runs of each
    solver, repeated while they fit in about budget_s seconds.  Times
    cover the solver call only."""
    from oracle import bridge
    m = model if slice_nodes is None else _node_slice(model, slice_nodes)
    node_days = m.Nn * m.replicates * _ndays(m)
    secs, trans, runs = 0.0, 0, 0
    t_start = time.perf_counter()
    while True:
        for solver in solvers:
            r = bridge.run(m, solver=solver, seed=1000 + runs, threads=threads)
            if r.err:
                raise SystemExit("checker failed with %d" % r.err)
            secs += r.seconds
            trans += r.transitions
            runs += 1
        if time.perf_counter() - t_start > budget_s or runs >= 4 * len(solvers):
            break
    what = ("the full workload" if slice_nodes is None
            else "the first %d nodes of the workload" % m.Nn)
    return {"value": node_days * runs / secs, "unit": "node-days/s", "cores": threads,
            "kind": "port",
            "transitions_per_s": trans / secs,
            "sample": "%d run(s) (%s) of %s, %.2f s of solver time" % (
                runs, "+".join(solvers), what, secs)}


def _node_slice(model, n):
    """The first n nodes of a model without events (same parameters)."""
    from siminf_amd.model import SimInfModel
    return SimInfModel(model.plugin, model.G, model.S, model.tspan,
                       model.u0[:, :n].copy(), ldata=model.ldata[:, :n].copy(),
                       gdata=model.gdata, v0=model.v0[:, :n].copy() if model.Nd else None,
                       E=model.E, compartments=model.compartments)


def _plan_times(plan):
    lib = plan.lib
    sm, fm, fc = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
    lib.SimInf_hip_plan_kernel_times(plan._handle, ctypes.byref(sm), ctypes.byref(fm),
                                     ctypes.byref(fc))
    tr, nd, ms = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
    lib.SimInf_hip_plan_stats(plan._handle, ctypes.byref(tr), ctypes.byref(nd),
                              ctypes.byref(ms))
    return {"solver_ms": sm.value, "fill_ms": fm.value, "fill_cols": fc.value,
            "transitions": tr.value, "node_days": nd.value, "run_ms": ms.value}


def bench_batch(args):
    """SEIR 1e6 nodes x 365 days, ssm and aem in one graph launch per
    step."""
    import siminf_amd as si
    from siminf_amd import workloads as W
    model = W.seir_million()
    solvers = ("ssm", "aem")
    plans = [si.Plan(model, solver=s) for s in solvers]
    for p in plans:
        for i in range(args.warmup):
            p.run(100 + i)
    batch = si.Batch(plans)
    tuned = {}
    if args.shape:
        # A launch shape given on the command line (as a profiled run
        # repeats the bench's autotuned one): lanes-per-wavefront shifts,
        # columns per fill block ("fill": {solver: cols}) and "graph" or
        # "streams".
        spec = json.loads(args.shape)
        fills = spec.get("fill", {})
        for s, p in zip(solvers, plans):
            p.set_lanes_per_wave(int(spec[s]))
            p.set_fill_block(int(fills.get(s, 0)))
        batch.close()
        batch = si.Batch(plans)
        batch.set_graph(spec.get("launch", "streams") == "graph")
        tuned = {"lanes_per_wave_shift": {s: int(spec[s]) for s in solvers},
                 "fill_block": {s: p.fill_block() for s, p in zip(solvers, plans)},
                 "launch": spec.get("launch", "streams"), "given": True}
    elif not args.no_autotune:
        # The launch shape is tuned for the plans together: each plan's
        # lanes per wavefront and columns per fill block, and the graph
        # against the direct enqueue.
        ms, shifts, graph = batch.autotune(seed=200, reps=7)
        tuned = {"lanes_per_wave_shift": dict(zip(solvers, shifts)),
                 "fill_block": {s: p.fill_block() for s, p in zip(solvers, plans)},
                 "launch": "graph" if graph else "streams", "tuned_ms": ms}
    for i in range(args.warmup):
        batch.run([300 + 2 * i, 301 + 2 * i])
    lat, launch_ms, per = [], [], {s: [] for s in solvers}
    node_days = trans = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        launch_ms.append(batch.run([1 + 2 * i, 2 + 2 * i]))
        lat.append(time.perf_counter() - s0)
        for s, p in zip(solvers, plans):
            t = _plan_times(p)
            per[s].append(t)
            node_days += t["node_days"]
            trans += t["transitions"]
    wall = time.perf_counter() - t0

    # The kernel that moves the step's bytes: each plan's frozen-node fill
    # (Nc ints per node-column it writes), timed by the event-record nodes
    # around it inside the graph.  The active passes (latency-bound SSA
    # chains) are reported beside it with the bytes the plan writes for
    # its active nodes.
    per_col = 4 * model.Nc + 8 * model.Nd
    fill_ms = [t["fill_ms"] for s in solvers for t in per[s]]
    fill_bytes = [t["fill_cols"] * per_col for s in solvers for t in per[s]]
    fill_gbs = statistics.mean(fill_bytes) / (statistics.mean(fill_ms) / 1e3) / 1e9
    total_bytes = algorithmic_bytes(model)
    solver_rows = {}
    for s in solvers:
        sm = statistics.mean(t["solver_ms"] for t in per[s])
        fb = statistics.mean(t["fill_cols"] for t in per[s]) * per_col
        solver_rows[s] = {"ms": sm, "fill_ms": statistics.mean(t["fill_ms"] for t in per[s]),
                          "transitions_per_run": statistics.mean(
                              t["transitions"] for t in per[s]),
                          "achieved_gbs": (total_bytes - fb) / (sm / 1e3) / 1e9}
    traffic = pmc_traffic("seir_batch", "k_fill_frozen")
    out = {
        "metric": "node-days/sec (SSA transitions/sec and p50/p99 simulation latency alongside)",
        "value": node_days / wall,
        "unit": "node-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 state, f64 rates/time",
        "data": "synthetic (seeded generator of the configuration's shape)",
        "config": {"workload": "SEIR 1e6 nodes x 365 days, ssm+aem in one batched "
                               "launch",
                   "solvers": list(solvers), "nodes": model.Nn, "days": _ndays(model),
                   "seeded_nodes": int((model.u0[2] > 0).sum())},
        "transitions_per_s": trans / wall,
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3)},
        "launch_ms_mean": statistics.mean(launch_ms),
        "autotuned": tuned,
        "roofline": {"bound": "hbm", "achieved": fill_gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": fill_gbs / HBM_PEAK_GBS,
                     "traffic": traffic, "kernel": "k_fill_frozen<4,false>",
                     "algorithmic_bytes_per_launch": statistics.mean(fill_bytes),
                     "kernel_ms_mean": statistics.mean(fill_ms),
                     "traffic_source": "profiles/pmc_traffic.json (rocprofv3 FETCH_SIZE + "
                                       "WRITE_SIZE of the fill kernels in this launch, "
                                       "scaled by a copy kernel)" if traffic else None},
        "active_passes": solver_rows,
    }
    batch.close()
    for p in plans:
        p.close()
    if not args.no_extras:
        for key, fn in (("pcie_inclusive", lambda: pcie_inclusive(model)),
                        ("sise3_herds", lambda: bench_sise3(args)),
                        ("sise3_sp_herds", lambda: bench_sise3(args, spatial=True)),
                        ("sise3_replicates", bench_sise3_replicates),
                        ("pfilter_request", lambda: bench_pfilter_request(args)),
                        ("abc_stream", lambda: bench_abc_stream(args, generations=2,
                                                                proposals=5000))):
            try:
                out[key] = fn()
            except Exception as e:          # reported, not fatal to the line
                out[key] = {"error": "%s: %s" % (type(e).__name__, e)}
            print("bench: %s done" % key, file=sys.stderr, flush=True)
        pc = out["pcie_inclusive"]
        if "error" not in pc:
            # The plain path (one blocking run per solver, U to the host)
            # against the batched launch with inputs resident.
            pc["plain_over_batched"] = (sum(pc[s]["seconds"] for s in solvers)
                                        / (out["ms_per_step"] / 1e3))
    if not args.no_cpu_baseline:
        T = host_cores()
        out["cpu_baseline"] = cpu_baseline(model, solvers, T)
        out["cpu_baseline_1core"] = cpu_baseline(model, solvers, 1, budget_s=10.0,
                                                 slice_nodes=100_000)
    return out


def pcie_inclusive(model):
    """The blocking entry point (SimInf_hip_run: plan, upload, run and the
    download of U into host memory), timed once per solver."""
    from siminf_amd import _native
    lib = _native.load()
    dev = lib.SimInf_hip_model_builtin(model.plugin.encode())
    out = {}
    for s in ("ssm", "aem"):
        b = _native.ArgsBundle(model, solver=0 if s == "ssm" else 1, seed=7)
        t0 = time.perf_counter()
        err = lib.SimInf_hip_run(ctypes.byref(b.args), ctypes.c_void_p(dev))
        dt = time.perf_counter() - t0
        if err:
            raise SystemExit("SimInf_hip_run failed: %d" % err)
        out[s] = {"seconds": dt, "node_days_per_s": model.Nn * _ndays(model) / dt,
                  "u_bytes_copied": int(b.U.nbytes)}
        del b
    return out


def bench_sise3(args, steps=10, spatial=False):
    """SISe3, 1600 herds x 1460 days with scheduled events (one replicate,
    grid-wide coupled kernel): mean of `steps` resident runs after the
    autotuner.  spatial: the same herds and events under SISe3_sp, each
    herd's phi also spreading to the two to eight herds within distance 1
    (every herd's state published each day, a second barrier per day)."""
    import siminf_amd as si
    from siminf_amd import workloads as W
    model = W.sise3_sp_herds() if spatial else W.sise3_herds()
    with si.Plan(model) as p:
        p.run(100)
        if not args.no_autotune:
            p.autotune(seed=200, reps=2)
        solo, pack = p.herd_map()
        nd = tr = 0
        ms = []
        t0 = time.perf_counter()
        for i in range(steps):
            a, b, c = p.run(1 + i)
            tr += a
            nd += b
            ms.append(c)
        wall = time.perf_counter() - t0
    out = {"workload": "%s 1600 herds x 1460 days, %d scheduled events"
                       % ("SISe3_sp" if spatial else "SISe3", len(model.events["event"])),
           "node_days_per_s": nd / wall, "transitions_per_s": tr / wall,
           "ms_per_run": wall / steps * 1e3, "kernel_ms_mean": statistics.mean(ms),
           "herd_map": {"solo_herds": solo, "herds_per_wavefront": pack},
           "algorithmic_bytes_per_run": algorithmic_bytes(model)}
    if not spatial and not args.no_cpu_baseline:
        # The checker at the reference's partition (node blocks over the
        # host's cores; the external transfers on one thread) and on one
        # core.
        out["cpu_baseline"] = cpu_baseline(model, ["ssm"], host_cores(), budget_s=6.0)
        out["cpu_baseline_1core"] = cpu_baseline(model, ["ssm"], 1, budget_s=4.0)
    return out


def bench_sise3_replicates(reps=(1, 2, 4, 8, 16, 32), runs=2):
    """SISe3, 1600 herds x 1460 days with scheduled events, R replicates
    in one plan (mssm): the grid-wide coupled kernel runs as many
    replicates per cooperative launch as the co-resident workgroups allow
    (default herd map: 64 herds to a wavefront).  Node-days/s per R, mean
    of `runs` resident runs."""
    import siminf_amd as si
    from siminf_amd import workloads as W
    rows = []
    for R in reps:
        model = W.sise3_herds(replicates=R)
        with si.Plan(model) as p:
            p.run(100)
            nd = 0
            t0 = time.perf_counter()
            for i in range(runs):
                nd += p.run(1 + i)[1]
            wall = time.perf_counter() - t0
        rows.append({"replicates": R, "ms_per_run": wall / runs * 1e3,
                     "node_days_per_s": nd / wall})
    return {"workload": "SISe3 1600 herds x 1460 days with events, R replicates "
                        "(default herd map)", "rows": rows}


def bench_pfilter_request(args):
    """One particle-filter request: SIR on 10,000 nodes, 2000 particles,
    30 weekly observations of the infected total, binomial observation
    process with p = 0.5 (observations thinned from one simulated
    trajectory of the same model)."""
    import siminf_amd as si
    from siminf_amd import workloads as W
    spec = W.pfilter_binomial_request(Nn=10_000, n_particles=2000, seed=23)
    t0 = time.perf_counter()
    res = si.pfilter(spec.model, 1, spec.times, spec.obs, spec.n_particles,
                     obs_prob=0.5, seed=spec.seed)
    dt = time.perf_counter() - t0
    return {"workload": "pfilter request: SIR 10000 nodes, 2000 particles, 30 weekly "
                        "intervals, binomial p=0.5",
            "seconds": dt, "particle_sims_per_s": spec.n_particles / dt,
            "particle_node_days_per_s": spec.n_particles * spec.model.Nn *
            (spec.times[-1] - spec.model.tspan[0]) / dt,
            "loglik": res.loglik, "min_ess": float(np.min(res.ess))}


def _ndays(model):
    ts = model.tspan
    f0 = np.floor(ts[0])
    k = 0
    while not (f0 + k + 1 > ts[-1]):
        k += 1
    return k + 1


def bench_pfilter_stream(args):
    """Particle-filter request stream through the serving loop.  A
    particle-simulation is one particle carried through all observation
    intervals of its request."""
    from siminf_amd import _native, serving
    from siminf_amd import workloads as W
    lib = _native.load()

    def request(spec):
        return serving.Request(spec.model, 1, spec.times, spec.obs, spec.n_particles,
                               obs_prob=spec.obs_prob, seed=spec.seed,
                               obs_dist=args.obs, obs_offset=1e-6)

    # Warm-up: the first launches of each kernel, outside the timed stream.
    for spec in W.pfilter_requests(max(args.warmup, 1), seed=99, nodes=(100, 1000),
                                   n_particles=args.particles):
        serving._run_filter(request(spec), lib)

    specs = W.pfilter_requests(args.requests, seed=5, n_particles=args.particles,
                               rate=args.rate)
    server = serving.PFilterServer(hbm_budget=args.hbm_budget_gib << 30,
                                   max_inflight=args.inflight, lib=lib)
    reqs = []
    with server:
        t0 = time.perf_counter()
        for spec in specs:
            delay = t0 + spec.arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            reqs.append(server.submit(request(spec)))
        for k, r in enumerate(reqs):
            r.done.wait()
            print("request %d: %d nodes x %d particles, %.3f s latency%s"
                  % (k, r.model.Nn, r.n_particles, r.latency,
                     "" if r.error is None else " error: %s" % r.error),
                  file=sys.stderr, flush=True)
        wall = max(r.t_done for r in reqs) - t0
    errors = [str(r.error) for r in reqs if r.error is not None]
    lat = [r.latency for r in reqs]
    psims = sum(r.n_particles for r in reqs)
    node_days = sum(r.n_particles * r.model.Nn * (r.times[-1] - r.model.tspan[0])
                    for r in reqs)

    # The same comparison sample through both paths: the served loop (all
    # submitted at once) and the plain path (one run() per interval).
    sample = W.pfilter_requests(args.plain_sample, seed=17, nodes=(100, 10000),
                                n_particles=args.particles, rate=args.rate)
    sreqs = []
    with serving.PFilterServer(hbm_budget=args.hbm_budget_gib << 30,
                               max_inflight=args.inflight, lib=lib) as srv:
        ts0 = time.perf_counter()
        for spec in sample:
            sreqs.append(srv.submit(request(spec)))
        for r in sreqs:
            r.done.wait()
        served_s = max(r.t_done for r in sreqs) - ts0
    tp0 = time.perf_counter()
    plain_ll = []
    for spec in sample:
        t1 = time.perf_counter()
        ll, _ = serving.host_pfilter(spec.model, 1, spec.times, spec.obs, spec.n_particles,
                                     obs_prob=spec.obs_prob, seed=spec.seed, lib=lib,
                                     obs_dist=args.obs, obs_offset=1e-6)
        plain_ll.append(ll)
        print("plain path: %d nodes, %.2f s" % (spec.model.Nn, time.perf_counter() - t1),
              file=sys.stderr, flush=True)
    plain_s = time.perf_counter() - tp0
    sample_psims = sum(s.n_particles for s in sample)

    out = {
        "metric": "particle-simulations/sec of the filter request stream (p50/p99 "
                  "request latency alongside)",
        "value": psims / wall,
        "unit": "particle-simulations/s",
        "n_gpus": 1,
        "steps": len(reqs),
        "warmup": max(args.warmup, 1),
        "ms_per_step": wall / len(reqs) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 state, f64 rates/weights",
        "data": "synthetic (seeded request generator; observations thinned from the "
                "mean-field SIR)",
        "config": {"workload": "pfilter request stream: SIR, nodes log-uniform in "
                               "[100, 100000], %d particles, 30 weekly observations of "
                               "the infected total, %s observation process, Poisson "
                               "arrivals at %.1f/s" % (args.particles, args.obs, args.rate),
                   "requests": len(reqs), "max_inflight": args.inflight,
                   "hbm_budget_gib": args.hbm_budget_gib},
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3)},
        "particle_node_days_per_s": node_days / wall,
        "wall_s": wall,
        "errors": errors,
        "peak_inflight": server.peak_inflight,
        "peak_hbm_bytes_admitted": server.peak_bytes,
        "plain_path_sample": {
            "requests": len(sample),
            "nodes": [int(s.model.Nn) for s in sample],
            "served_s": served_s, "plain_s": plain_s,
            "served_particle_sims_per_s": sample_psims / served_s,
            "plain_particle_sims_per_s": sample_psims / plain_s,
            "speedup": plain_s / served_s,
            "served_loglik": [r.result.loglik if r.result else None for r in sreqs],
            "plain_loglik": plain_ll},
    }
    print(json.dumps(out), flush=True)


def bench_abc_stream(args, generations=8, proposals=10_000, n_particles=1000, chunk=500):
    """The mixed ABC-SMC stream: SIR (1000 nodes x 150 days), SEIR (20,000
    nodes x 365 days) and SISe3 with scheduled events (500 herds x 365
    days), each generation's proposals of the three models run chunk by
    chunk in one batched launch, distances on daily prevalence from the
    device.  Reports proposals/s and p50/p99 per-simulation latency."""
    from siminf_amd import abc_stream as S
    problems = S.config_problems()
    res = S.run_stream(problems, generations=generations, proposals=proposals,
                       n_particles=n_particles, chunk=chunk, seed=7,
                       fuse=not args.no_fuse)
    lat = np.asarray(res.latencies)
    return {"workload": "ABC-SMC stream: SIR 1000 nodes x 150 days, SEIR 20000 nodes x 365 "
                        "days, SISe3 500 herds x 365 days with events; %d generations x %d "
                        "proposals per model, %d accepted, chunks of %d per batched launch"
                        % (generations, proposals, n_particles, chunk),
            "proposals": res.proposals, "wall_s": res.wall_s,
            "proposals_per_s": res.proposals / res.wall_s,
            "node_days_per_s": res.node_days / res.wall_s,
            "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                           "p99": float(np.percentile(lat, 99) * 1e3)},
            "launch_ms_mean": float(np.mean(res.launch_s) * 1e3),
            "split_s": res.split_s,
            "shape": res.shape,
            "fused_prevalence": dict(zip([pr.name for pr in problems], res.fused)),
            "tolerances": {pr.name: [g[pr.name][0] for g in res.generations]
                           for pr in problems},
            "posterior_means": {pr.name: dict(zip([q.parameter for q in pr.priors],
                                                  np.average(res.generations[-1][pr.name][1],
                                                             axis=0,
                                                             weights=res.generations[-1][
                                                                 pr.name][2]).tolist()))
                                for pr in problems}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="seir_batch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--shape", default=None,
                    help='seir_batch: launch shape instead of the autotuner, as JSON '
                         '{"ssm": shift, "aem": shift, "launch": "graph"|"streams"}')
    ap.add_argument("--no-extras", action="store_true",
                    help="seir_batch: skip the PCIe-inclusive, SISe3 and filter-request "
                         "measurements")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default lanes-per-wavefront rule of the node kernels")
    ap.add_argument("--no-fuse", action="store_true",
                    help="abc_stream: plans keep U and prevalence reads it back")
    ap.add_argument("--requests", type=int, default=24,
                    help="pfilter_stream: requests in the timed stream")
    ap.add_argument("--particles", type=int, default=2000)
    ap.add_argument("--generations", type=int, default=8,
                    help="abc_stream: ABC-SMC generations")
    ap.add_argument("--proposals", type=int, default=10_000,
                    help="abc_stream: proposals per model and generation")
    ap.add_argument("--rate", type=float, default=4.0, help="pfilter_stream: arrivals/s")
    ap.add_argument("--inflight", type=int, default=8)
    ap.add_argument("--hbm-budget-gib", type=int, default=96)
    ap.add_argument("--plain-sample", type=int, default=4)
    ap.add_argument("--obs", default="poisson", choices=("poisson", "binomial"))
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device (torch), for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * 2**20 // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_mib": args.copy_calibration}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(the multi-GPU sharding is a design, not a launch path)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "pfilter_stream":
        bench_pfilter_stream(args)
        return
    if args.workload == "abc_stream":
        out = {"metric": "proposals/sec of the mixed ABC-SMC stream (p50/p99 "
                         "per-simulation latency alongside)", "unit": "proposals/s",
               "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "int32 state, f64 rates/time",
               "data": "synthetic (observations simulated at the generating parameters)"}
        r = bench_abc_stream(args, generations=args.generations, proposals=args.proposals)
        out.update({"value": r["proposals_per_s"], "steps": r["proposals"], "warmup": 0,
                    "ms_per_step": 1e3 * r["wall_s"] / r["proposals"],
                    "config": {"workload": r["workload"]}, "abc_stream": r})
        print(json.dumps(out), flush=True)
        return
    if args.workload == "seir_batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.workload == "pfilter_request":
        # One configuration-4 request, repeated: warm-up requests, then
        # `steps` timed ones.
        for _ in range(args.warmup):
            bench_pfilter_request(args)
        runs = [bench_pfilter_request(args) for _ in range(args.steps)]
        secs = [r["seconds"] for r in runs]
        out = {"metric": "particle-simulations/sec of one particle-filter request",
               "value": 2000 * len(secs) / sum(secs), "unit": "particle-sims/s",
               "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
               "ms_per_step": 1e3 * sum(secs) / len(secs), "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None,
               "dtype": "int32 state, f64 rates/time", "data": "synthetic",
               "config": {"workload": runs[0]["workload"]},
               "p50_ms": float(np.percentile(secs, 50) * 1e3),
               "p99_ms": float(np.percentile(secs, 99) * 1e3), "request": runs[-1]}
        print(json.dumps(out), flush=True)
        return

    from siminf_amd import _native
    model, config = make_workload(args.workload)
    solver = config["solver"]
    lib = _native.load()
    dev = lib.SimInf_hip_model_builtin(model.plugin.encode())
    bundle = _native.ArgsBundle(model, solver=0 if solver == "ssm" else 1, seed=1)
    plan = ctypes.c_void_p()
    err = lib.SimInf_hip_plan_create(ctypes.byref(plan), ctypes.byref(bundle.args),
                                     ctypes.c_void_p(dev))
    if err:
        raise SystemExit("plan_create failed: %d" % err)

    split = []

    def step(seed):
        e = lib.SimInf_hip_plan_run(plan, seed)
        if e:
            raise SystemExit("run failed: %d %s" % (e, lib.SimInf_hip_error_message(e)))
        tr, nd, ms = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
        lib.SimInf_hip_plan_stats(plan, ctypes.byref(tr), ctypes.byref(nd), ctypes.byref(ms))
        sm, fm, fc = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        lib.SimInf_hip_plan_kernel_times(plan, ctypes.byref(sm), ctypes.byref(fm),
                                         ctypes.byref(fc))
        split.append((sm.value, fm.value, fc.value))
        return tr.value, nd.value, ms.value

    for i in range(args.warmup):
        step(100 + i)
    # The launch shape from the plan's autotuner, before the timed steps:
    # lanes per wavefront for the node kernels, the number of herds on a
    # wavefront of their own for a grid-wide coupled kernel.
    best = ctypes.c_int(-1)
    if not args.no_autotune:
        err = lib.SimInf_hip_plan_autotune(plan, 200, 2, ctypes.byref(best), None)
        if err:
            raise SystemExit("autotune failed: %d" % err)
    autotuned = {"lanes_per_wave_shift": best.value}
    if lib.SimInf_hip_plan_get_herd_map(plan, None, None) == 0:
        n_solo, pack = ctypes.c_int(), ctypes.c_int()
        lib.SimInf_hip_plan_get_herd_map(plan, ctypes.byref(n_solo), ctypes.byref(pack))
        if (best.value >= 0 and model.events is not None and
                (np.asarray(model.events["event"]) == 3).any()):
            autotuned = {"solo_herds": n_solo.value, "herds_per_wavefront": pack.value}
    del split[:]
    lat, kms, trans = [], [], 0
    node_days = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        s0 = time.perf_counter()
        tr, nd, ms = step(1 + i)
        lat.append(time.perf_counter() - s0)
        kms.append(ms)
        trans += tr
        node_days += nd
    wall = time.perf_counter() - t0
    lib.SimInf_hip_plan_destroy(plan)

    # Roofline of the dominant kernel.  The solver kernel's algorithmic
    # bytes are the whole run's less the trajectory columns the fill kernel
    # writes; the fill kernel's are those columns (Nc ints and Nd doubles
    # per node-column) plus its read of each frozen node's state.
    total_bytes = algorithmic_bytes(model)
    solver_ms = statistics.mean(x[0] for x in split)
    fill_ms = statistics.mean(x[1] for x in split)
    fill_cols = statistics.mean(x[2] for x in split)
    per_col = 4 * model.Nc + 8 * model.Nd
    fill_bytes = fill_cols * per_col
    if fill_ms > solver_ms:
        dominant, kernel_s, bytes_per_launch = "k_fill_frozen", fill_ms / 1e3, fill_bytes
    else:
        dominant, kernel_s = "solver", solver_ms / 1e3
        bytes_per_launch = total_bytes - fill_bytes
    achieved = bytes_per_launch / kernel_s / 1e9
    traffic = pmc_traffic(args.workload, dominant)
    out = {
        "metric": "node-days/sec (SSA transitions/sec and p50/p99 simulation latency alongside)",
        "value": node_days / wall,
        "unit": "node-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 state, f64 rates/time",
        "data": "synthetic (seeded generator of the configuration's shape)",
        "config": dict(config, nodes=model.Nn, replicates=model.replicates,
                       days=_ndays(model), events=int(len(model.events["event"]))),
        "transitions_per_s": trans / wall,
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3)},
        "kernel_ms_mean": statistics.mean(kms),
        "autotuned": autotuned,
        "kernel_ms_split": {"solver": solver_ms, "fill": fill_ms,
                            "fill_node_columns": fill_cols},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "kernel": dominant,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(model, (solver,), host_cores())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
    # Every plan is destroyed by now: end the device work and return the
    # pooled memory before the interpreter's exit handlers run.
    if "siminf_amd._native" in sys.modules:
        from siminf_amd import _native
        if _native._lib is not None:
            _native._lib.SimInf_hip_release()
