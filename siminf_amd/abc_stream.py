"""ABC-SMC over several models at once, batched across model types.

Each problem is a model (SIR, SEIR, SISe3, ...), priors on some of its
ldata or gdata parameters, and an observed daily prevalence series.  A
generation draws `proposals` parameter vectors per problem -- from the
priors in the first generation, else from the previous generation's
particles moved by a Gaussian kernel (SimInf_abc_proposals; the kernel's
covariance twice the particles' covariance, R/abc.R:897-901) -- and
simulates them as replicates of one plan per problem, each replicate with
its own parameters (Plan.set_replicate_data).  A chunk of every problem's
proposals runs in one batched launch (one Batch holding the problems'
plans), and each replicate's daily prevalence comes from the plan's
resident trajectory on the device (prevalence level 1 per column); the
distance is the root mean square difference to the observed series.
Plans whose solver runs node kernels keep only the sums that prevalence
needs instead of U (Plan.fuse_prevalence), so the trajectory is neither
written nor read back.

The tolerance of a generation is its n_particles-th smallest distance, so
that n_particles proposals are accepted from every generation (a
tolerance schedule set by the distances, as the reference's adaptive
rule sets it from the previous generation, R/abc.R:408-500); accepted
particles are weighted as the reference weights them
(SimInf_abc_weights, on the device).

Latency of a simulation: from the start of its chunk's launch to its
distance being on the host.
"""
import time

import numpy as np

from . import abc_smc as A
from .model import SimInfModel
from .plan import Batch, Plan


class Problem:
    """One ABC problem: a model of one replicate, priors on parameters
    (name -> (low, high), uniform), the prevalence formula and the
    observed series (one value per tspan point)."""

    def __init__(self, name, model, priors, formula, observed=None, solver="ssm"):
        self.name = name
        self.model = model
        self.formula = formula
        self.solver = solver
        self.priors = A.parse_priors(["%s ~ uniform(%r, %r)" % (k, lo, hi)
                                      for k, (lo, hi) in priors.items()])
        self.slot, self.rows = A.match_priors(model, self.priors)
        self.observed = observed

    def replicated(self, n):
        """The model with n replicates (u0, v0 repeated)."""
        m = self.model
        return SimInfModel(m.plugin, m.G, m.S, m.tspan, np.tile(m.u0, (1, n)),
                           ldata=m.ldata, gdata=m.gdata,
                           v0=np.tile(m.v0, (1, n)) if m.Nd else None, E=m.E, N=m.N,
                           events={k: m.events[k] for k in
                                   ("event", "time", "node", "dest", "n", "proportion",
                                    "select", "shift")},
                           replicates=n, compartments=m.compartments, v0_names=m.v0_names,
                           ldata_names=m.ldata_names, gdata_names=m.gdata_names)

    def replicate_data(self, xx):
        """(ldata, gdata) per replicate for the proposals xx (n x d)."""
        m, n = self.model, len(xx)
        if self.slot == "ldata":
            if (m.ldata == m.ldata[:, :1]).all():
                # The nodes share their ldata: one column per replicate.
                ld = np.tile(m.ldata[:, :1], (1, n))
                for k, r in enumerate(self.rows):
                    ld[r, :] = xx[:, k]
            else:
                ld = np.tile(m.ldata, (1, n))
                for k, r in enumerate(self.rows):
                    ld[r, :] = np.repeat(xx[:, k], m.Nn)
            return ld, None
        gd = np.tile(m.gdata.reshape(-1, 1), (1, n))
        for k, r in enumerate(self.rows):
            gd[r, :] = xx[:, k]
        return None, gd


def distances(prev, observed, n, tlen):
    """Root mean square distance of each replicate's prevalence series
    (prev: n*tlen values, replicate-major) to the observed one."""
    p = np.asarray(prev).reshape(n, tlen)
    return np.sqrt(np.mean((p - np.asarray(observed)[None, :]) ** 2, axis=1))


class StreamResult:
    def __init__(self):
        self.generations = []        # per generation: {problem: (eps, x, w)}
        self.latencies = []          # per simulation, seconds
        self.launch_s = []           # per batched launch
        self.proposals = 0
        self.node_days = 0
        self.wall_s = 0.0
        self.shape = None
        self.fused = []              # per problem: whether its plan fused
        # Seconds spent setting parameters, in the batched launches and in
        # the summaries (prevalence and distances).
        self.split_s = {"parameters": 0.0, "launch": 0.0, "summaries": 0.0}


def observe(problem, seed, lib=None):
    """The observed series: the problem's model run once at its own
    parameters."""
    with Plan(problem.model, solver=problem.solver, lib=lib) as p:
        p.run(seed)
        return p.prevalence(problem.formula, level=1)


def run_stream(problems, generations=8, proposals=10_000, n_particles=1000, chunk=1000,
               seed=1, lib=None, autotune=True, fuse=True):
    """ABC-SMC of every problem, generation by generation; a generation's
    proposals of all problems run chunk by chunk, each chunk of all
    problems in one batched launch.  fuse: node-kernel plans keep only
    their prevalence sums (Plan.fuse_prevalence).  Returns a
    StreamResult."""
    rng = np.random.default_rng(seed)
    for pr in problems:
        if pr.observed is None:
            pr.observed = observe(pr, int(rng.integers(1, 2**62)), lib)
    plans = [Plan(pr.replicated(chunk), solver=pr.solver, lib=lib) for pr in problems]
    res = StreamResult()
    try:
        # Plans that can keep the prevalence sums instead of U do (the
        # distances need nothing else).
        res.fused = [pl.fuse_prevalence(pr.formula) if fuse else False
                     for pr, pl in zip(problems, plans)]
        # Parameters per replicate before the batch is captured, so that
        # the graph reads them where they stay.
        for pr, pl in zip(problems, plans):
            xx0, _ = A.proposals(pr.priors, chunk, None, None, None, rng)
            ld, gd = pr.replicate_data(xx0)
            pl.set_replicate_data(ldata=ld, gdata=gd)
        batch = Batch(plans, lib=lib)
        if autotune:
            ms, shifts, graph = batch.autotune(seed=int(rng.integers(1, 2**62)), reps=1)
            res.shape = {"lanes_per_wave_shift": shifts, "graph": graph, "tuned_ms": ms}
        state = [dict(x=None, w=None, sigma=None) for _ in problems]
        t_start = time.perf_counter()
        for g in range(generations):
            xx_all, anc_all, d_all = [], [], []
            for k, pr in enumerate(problems):
                st = state[k]
                xx, anc = A.proposals(pr.priors, proposals, st["x"], st["w"], st["sigma"], rng)
                xx_all.append(xx)
                anc_all.append(anc)
                d_all.append(np.empty(proposals))
            for c0 in range(0, proposals, chunk):
                c1 = min(c0 + chunk, proposals)
                t0 = time.perf_counter()
                for k, (pr, pl) in enumerate(zip(problems, plans)):
                    xx = xx_all[k][c0:c1]
                    if len(xx) < chunk:
                        # The last chunk's spare replicates repeat its first
                        # proposal (their results are not used).
                        xx = np.vstack([xx, np.repeat(xx[:1], chunk - len(xx), axis=0)])
                    ld, gd = pr.replicate_data(xx)
                    pl.set_replicate_data(ldata=ld, gdata=gd)
                t1 = time.perf_counter()
                batch.run([int(x) for x in rng.integers(1, 2**62, len(plans))])
                t2 = time.perf_counter()
                t_launch = t2 - t1
                for k, (pr, pl) in enumerate(zip(problems, plans)):
                    prev = pl.prevalence(pr.formula, level=1)
                    d = distances(prev, pr.observed, chunk, len(pr.model.tspan))
                    d_all[k][c0:c1] = d[:c1 - c0]
                    res.node_days += (c1 - c0) * pr.model.Nn * len(pr.model.tspan)
                dt = time.perf_counter() - t0
                res.split_s["parameters"] += t1 - t0
                res.split_s["launch"] += t_launch
                res.split_s["summaries"] += dt - (t2 - t0)
                res.launch_s.append(t_launch)
                res.latencies.extend([dt] * ((c1 - c0) * len(problems)))
                res.proposals += (c1 - c0) * len(problems)
            gen = {}
            for k, pr in enumerate(problems):
                st = state[k]
                order = np.argsort(d_all[k], kind="stable")[:n_particles]
                eps = float(d_all[k][order[-1]])
                x_new, anc = xx_all[k][order], anc_all[k][order]
                w_new = A.weights(pr.priors, st["x"], x_new, anc, st["w"], st["sigma"],
                                  lib=lib)
                st["x"], st["w"] = x_new, w_new
                st["sigma"] = 2.0 * np.atleast_2d(np.cov(x_new, rowvar=False))
                gen[pr.name] = (eps, x_new, w_new)
            res.generations.append(gen)
        res.wall_s = time.perf_counter() - t_start
        batch.close()
    finally:
        for pl in plans:
            pl.close()
    return res


def config_problems(scale=1.0, seed=3):
    """The three problems of the mixed stream, with the structures of the
    single-model workloads: the parsed SIR of the README (1000 nodes x
    150 days, gdata beta, gamma), SEIR (20,000 nodes with sizes ~
    Poisson(500), 1 % seeded, 365 days, ldata beta, epsilon, gamma) and
    SISe3 with scheduled events (500 herds x 365 days, gdata upsilon_1,
    gamma_1); priors uniform over +-50 % of the generating values.  scale
    shrinks node counts (for tests)."""
    from . import workloads as W
    n = lambda x: max(10, int(round(x * scale)))
    sir = W.mparse_sir_readme(Nn=n(1000), days=150)
    seir = W.seir_million(Nn=n(20_000), days=365, seed=seed, seeded_frac=0.01)
    sise3 = W.sise3_herds(Nn=n(500), days=365, seed=seed)
    pm = lambda v: (0.5 * v, 1.5 * v)
    return [
        Problem("SIR", sir, {"beta": pm(0.16), "gamma": pm(0.077)}, "I ~ S + I + R"),
        Problem("SEIR", seir, {"beta": pm(0.16), "epsilon": pm(0.25), "gamma": pm(0.077)},
                "I ~ S + E + I + R"),
        Problem("SISe3", sise3, {"upsilon_1": pm(0.018), "gamma_1": pm(0.1)},
                "I_1 + I_2 + I_3 ~ S_1 + I_1 + S_2 + I_2 + S_3 + I_3"),
    ]
