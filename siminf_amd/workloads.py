"""Synthetic workloads with the shapes of the benchmark configurations.

  mparse_sir_readme   SIR via mparse, Nn x (99, 5, 0), beta 0.16,
                      gamma 0.077, tspan 1:150 (README.md:60-76), on the
                      hand-written plugin of that model
  mparse_sir_generated  the same model through the parser, on the plugin
                      generated from its transitions
  mparse_sis_env      a parsed SIS model with a continuous state updated by
                      a post-time-step function, ldata and enum constants
  sise3_herds         SISe3 on herds of NegBin(mean 130, size 1.5) animals
                      split 10/15/75 % by age, 1 % of herds seeded, with
                      scheduled enter / exit / internal / external events
                      (parameters of man/examples/SISe3.R)
  sise3_sp_herds      the sise3_herds herds and events with local spread
                      of phi among herds on a jittered grid (SISe3_sp)
  seir_million        SEIR on Nn ~ Poisson(500) nodes, 0.5 % seeded with
                      5 infected, beta 0.16, epsilon 0.25, gamma 0.077
  pfilter_requests    particle-filter requests: SIR on node counts
                      log-uniform in [100, 100000], beta ~ U(0.05, 0.4),
                      gamma ~ U(0.03, 0.2), 30 weekly observations of the
                      infected total thinned with p = 0.5, Poisson arrivals

Scheduled events must stay feasible: the reference aborts when an event
asks for more individuals than a node holds (SimInf_solver.c:102-105).
The SSA only moves animals between S_k and I_k, so the per-age totals of
a herd change through events alone; the generator tracks a lower and an
upper bound of every (herd, age) total through the event list and clamps
fixed counts to the lower bound.  Proportion draws (used for exits only)
remove a binomial number of animals; the lower bound then drops by
p*ub + 7 sd + 5, which the draw exceeds with probability below 1e-11.
"""
import numpy as np

from . import models


def mparse_sir_readme(Nn=1000, days=150, replicates=1):
    return models.SIR_mparse(u0={"S": np.full(Nn, 99), "I": np.full(Nn, 5),
                                 "R": np.zeros(Nn, int)},
                             tspan=np.arange(1, days + 1, dtype=float),
                             beta=0.16, gamma=0.077, replicates=replicates)


SIR_TRANSITIONS = ["S -> beta*S*I/N -> I", "I -> gamma*I -> R", "N <- S+I+R"]


def mparse_sir_generated(Nn=1000, days=150, replicates=1):
    """The README model built by the parser: its plugin is generated from
    the transitions and compiled for the device."""
    from .modelparser import mparse
    return mparse(SIR_TRANSITIONS, ["S", "I", "R"], gdata={"beta": 0.16, "gamma": 0.077},
                  u0={"S": np.full(Nn, 99), "I": np.full(Nn, 5), "R": np.zeros(Nn, int)},
                  tspan=np.arange(1, days + 1, dtype=float), replicates=replicates)


PTS_DECAY = ["v_new[PHI] = v[PHI] * (1.0 - gdata[DECAY]) +",
             "    gdata[SHED] * u[I] / (u[S] + u[I] + 1.0);",
             "if (!R_FINITE(v_new[PHI]))",
             "    return SIMINF_ERR_V_IS_NOT_FINITE;",
             "if (v_new[PHI] < 0.0)",
             "    return SIMINF_ERR_V_IS_NEGATIVE;",
             "return 1;"]


def mparse_sis_env(Nn=200, days=120, seed=3, replicates=1):
    """A parsed SIS model with an environmental infectious pressure phi
    (a continuous state v0) that the post time step updates, per-node
    recovery rates (ldata) and named constants (use_enum)."""
    from .modelparser import mparse
    rng = np.random.default_rng(seed)
    S = rng.integers(20, 400, Nn)
    I = (rng.random(Nn) < 0.1) * rng.integers(1, 10, Nn)
    return mparse(["S -> upsilon*phi*S -> I", "I -> gamma*I -> S"], ["S", "I"],
                  ldata={"gamma": rng.uniform(0.05, 0.2, Nn)},
                  gdata={"upsilon": 0.05, "DECAY": 0.1, "SHED": 0.3},
                  v0={"phi": np.zeros(Nn)}, u0={"S": S, "I": I},
                  tspan=np.arange(1, days + 1, dtype=float), pts_fun=PTS_DECAY,
                  use_enum=True, replicates=replicates)


def seir_million(Nn=1_000_000, days=365, seed=7, seeded_frac=0.005):
    rng = np.random.default_rng(seed)
    N = rng.poisson(500, Nn)
    I = np.zeros(Nn, dtype=np.int64)
    k = max(1, int(round(seeded_frac * Nn)))
    idx = rng.choice(Nn, k, replace=False)
    I[idx] = np.minimum(5, N[idx])
    S = N - I
    return models.SEIR(u0={"S": S, "E": np.zeros(Nn, int), "I": I,
                           "R": np.zeros(Nn, int)},
                       tspan=np.arange(1, days + 1, dtype=float),
                       beta=0.16, epsilon=0.25, gamma=0.077)


SISE3_GDATA = dict(upsilon_1=1.8e-2, upsilon_2=1.8e-2, upsilon_3=1.8e-2,
                   gamma_1=0.1, gamma_2=0.1, gamma_3=0.1, alpha=1.0,
                   beta_t1=0.1, beta_t2=0.1, beta_t3=0.125, beta_t4=0.125,
                   epsilon=0.0)


def _power_law_degrees(rng, n, exponent=2.3, kmax=1000):
    k = np.arange(1, kmax + 1, dtype=float)
    p = k ** (-exponent)
    p /= p.sum()
    return rng.choice(k, size=n, p=p)


def sise3_events(u_age, days, rng, enter_rate=0.1, exit_rate=0.08,
                 internal_every=7, external_rate=0.2, prop_frac=0.1):
    """Feasible SISe3 events for herds whose per-age totals are u_age
    (Nn x 3).  Returns the event columns (one-based node/dest/select/shift,
    as the R layer stores them)."""
    Nn = u_age.shape[0]
    lb = u_age.astype(np.int64).copy()
    ub = u_age.astype(np.int64).copy()
    w = _power_law_degrees(rng, Nn)
    w = w / w.sum()
    out = {c: [] for c in ("event", "time", "node", "dest", "n", "proportion",
                           "select", "shift")}

    def push(ev, t, node, dest, n, prop, select, shift):
        out["event"].append(ev)
        out["time"].append(t)
        out["node"].append(node + 1)
        out["dest"].append(dest + 1)
        out["n"].append(n)
        out["proportion"].append(prop)
        out["select"].append(select)
        out["shift"].append(shift)

    def leave_bound(a, p):
        sd = np.sqrt(max(a * p * (1 - p), 0.0))
        return int(min(a, np.ceil(p * a + 7 * sd + 5)))

    for t in range(1, days + 1):
        # Within a day the solver runs E1 before E2, each in sorted order:
        # exit (0), enter (1), internal (2), external (3); the bounds are
        # updated in the same order.
        day = []
        n_exit = rng.poisson(exit_rate * Nn)
        for node in rng.integers(0, Nn, n_exit):
            age = int(rng.integers(0, 3))
            day.append((0, node, age))
        n_enter = rng.poisson(enter_rate * Nn)
        for node in rng.integers(0, Nn, n_enter):
            day.append((1, node, 0))
        if internal_every and t % internal_every == 0:
            for node in range(Nn):
                day.append((2, node, 0))
                day.append((2, node, 1))
        n_ext = rng.poisson(external_rate * Nn)
        src = rng.choice(Nn, n_ext, p=w)
        dst = rng.choice(Nn, n_ext, p=w)
        for s_, d_ in zip(src, dst):
            if s_ != d_:
                day.append((3, int(s_), int(d_)))
        # Sort like SimInf_events: (event, select) within the day.
        def sel_of(e):
            return {0: 4 + e[2], 1: 1, 2: 4 + e[2]}.get(e[0], None)
        keyed = []
        for e in day:
            if e[0] == 3:
                age = int(rng.integers(0, 3))
                keyed.append((3, 4 + age, e, age))
            else:
                keyed.append((e[0], sel_of(e), e, e[2]))
        keyed.sort(key=lambda x: (x[0], x[1]))
        for ev, select, e, age in keyed:
            node = e[1]
            if ev == 0:
                if rng.random() < prop_frac:
                    p = float(rng.uniform(0.01, 0.05))
                    push(0, t, node, node, 0, p, select, 0)
                    lb[node, age] -= min(lb[node, age], leave_bound(ub[node, age], p))
                else:
                    n = min(int(rng.integers(1, 11)), int(lb[node, age]))
                    if n <= 0:
                        continue
                    push(0, t, node, node, n, 0.0, select, 0)
                    lb[node, age] -= n
                    ub[node, age] -= n
            elif ev == 1:
                n = int(rng.integers(1, 11))
                push(1, t, node, node, n, 0.0, 1, 0)
                lb[node, 0] += n
                ub[node, 0] += n
            elif ev == 2:
                n = min(int(rng.integers(1, 11)), int(lb[node, age]))
                if n <= 0:
                    continue
                # select 4 with shift 1 moves age 1 -> 2; select 5 with
                # shift 2 moves age 2 -> 3 (R/SISe3.R:158-160).
                push(2, t, node, node, n, 0.0, select, age + 1)
                lb[node, age] -= n
                ub[node, age] -= n
                lb[node, age + 1] += n
                ub[node, age + 1] += n
            else:
                dest = e[2]
                n = min(int(rng.integers(1, 11)), int(lb[node, age]))
                if n <= 0:
                    continue
                push(3, t, node, dest, n, 0.0, select, 0)
                lb[node, age] -= n
                ub[node, age] -= n
                lb[dest, age] += n
                ub[dest, age] += n
    return {k: np.asarray(v) for k, v in out.items()}


def sise3_herds(Nn=1600, days=1460, seed=11, events=True, replicates=1,
                **event_rates):
    rng = np.random.default_rng(seed)
    size = 1.5
    mean = 130.0
    herd = rng.negative_binomial(size, size / (size + mean), Nn)
    a1 = np.floor(0.10 * herd).astype(int)
    a2 = np.floor(0.15 * herd).astype(int)
    a3 = herd - a1 - a2
    I1 = np.zeros(Nn, dtype=int)
    seeded = rng.choice(Nn, max(1, Nn // 100), replace=False)
    I1[seeded] = 10
    u0 = {"S_1": a1, "I_1": I1, "S_2": a2, "I_2": np.zeros(Nn, int),
          "S_3": a3, "I_3": np.zeros(Nn, int)}
    ev = None
    if events:
        u_age = np.stack([a1 + I1, a2, a3], axis=1)
        ev = sise3_events(u_age, days, rng, **event_rates)
    return models.SISe3(u0=u0, tspan=np.arange(1, days + 1, dtype=float),
                        events=ev, phi=0.0, end_t1=91, end_t2=182, end_t3=273,
                        end_t4=365, replicates=replicates, **SISE3_GDATA)


def sise3_sp_herds(Nn=1600, days=1460, seed=11, events=True, replicates=1,
                   coupling=0.0005, **event_rates):
    """The herds and events of sise3_herds on a grid of spacing 0.8 with
    each point moved by up to 0.3 along each axis, each herd coupled
    (SISe3_sp, local spread of phi) to the two to eight herds within
    distance 1."""
    from . import spatial
    base = sise3_herds(Nn, days, seed, events=events, replicates=1, **event_rates)
    rng = np.random.default_rng(seed + 1000)
    k = int(np.ceil(np.sqrt(Nn)))
    gx, gy = np.meshgrid(np.arange(k), np.arange(k))
    x = 0.8 * gx.ravel()[:Nn] + rng.uniform(-0.3, 0.3, Nn)
    y = 0.8 * gy.ravel()[:Nn] + rng.uniform(-0.3, 0.3, Nn)
    distance = spatial.distance_matrix(x, y, 1.0)
    g = dict(zip(base.gdata_names, base.gdata))
    return models.SISe3_sp(
        u0=base.u0[:, :Nn].T, tspan=base.tspan, events=base.events if events else None,
        phi=0.0, upsilon_1=g["upsilon_1"], upsilon_2=g["upsilon_2"],
        upsilon_3=g["upsilon_3"], gamma_1=g["gamma_1"], gamma_2=g["gamma_2"],
        gamma_3=g["gamma_3"], alpha=g["alpha"], beta_t1=g["beta_t1"],
        beta_t2=g["beta_t2"], beta_t3=g["beta_t3"], beta_t4=g["beta_t4"],
        end_t1=91, end_t2=182, end_t3=273, end_t4=365, distance=distance,
        coupling=coupling, replicates=replicates)


class FilterSpec:
    """A particle-filter request of the request-stream workload: the model,
    the observation times and counts, and the arrival offset (seconds)."""

    def __init__(self, model, times, obs, n_particles, obs_prob, seed, arrival, beta,
                 gamma):
        self.model = model
        self.times = times
        self.obs = obs
        self.n_particles = n_particles
        self.obs_prob = obs_prob
        self.seed = seed
        self.arrival = arrival
        self.beta = beta
        self.gamma = gamma


def _sir_mean_infected(beta, gamma, S0, I0, R0, times, t0, substeps=16):
    """I(t) of the mean-field SIR of one node (Euler, substeps per day)."""
    S, I, R = float(S0), float(I0), float(R0)
    N = S + I + R
    out = []
    t = float(t0)
    h = 1.0 / substeps
    for te in times:
        while t < te - 1e-12:
            inf = beta * S * I / N * h
            rec = gamma * I * h
            S, I, R = S - inf, I + inf - rec, R + rec
            t += h
        out.append(I)
    return np.asarray(out)


def pfilter_requests(n_requests=24, seed=5, nodes=(100, 100000), n_particles=2000,
                     n_obs=30, every=7, obs_prob=0.5, rate=4.0):
    """Requests of the particle-filter stream: SIR (S=99, I=5, R=0 per
    node) with node counts log-uniform in ``nodes``; observations are the
    binomially thinned mean-field infected total at days 1 + every*k."""
    rng = np.random.default_rng(seed)
    times = 1.0 + every * np.arange(1, n_obs + 1, dtype=float)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    arrivals -= arrivals[0]
    out = []
    for k in range(n_requests):
        Nn = int(round(np.exp(rng.uniform(np.log(nodes[0]), np.log(nodes[1])))))
        beta = float(rng.uniform(0.05, 0.4))
        gamma = float(rng.uniform(0.03, 0.2))
        I = _sir_mean_infected(beta, gamma, 99, 5, 0, times, 1.0)
        obs = rng.binomial(np.round(I * Nn).astype(np.int64), obs_prob)
        model = models.SIR_mparse(u0={"S": np.full(Nn, 99), "I": np.full(Nn, 5),
                                      "R": np.zeros(Nn, int)},
                                  tspan=np.array([1.0, times[-1]]), beta=beta, gamma=gamma)
        out.append(FilterSpec(model, times, obs.reshape(-1, 1), n_particles, obs_prob,
                              int(rng.integers(0, 2**62)), float(arrivals[k]), beta, gamma))
    return out


def pfilter_binomial_request(Nn=10_000, n_particles=2000, seed=23, n_obs=30, every=7,
                             obs_prob=0.5, min_infected=1000, lib=None):
    """A particle-filter request of the request-stream's shape with a
    binomial observation process: SIR (S=99, I=5, R=0 per node), beta ~
    U(0.05, 0.4), gamma ~ U(0.03, 0.2), the infected total observed at
    days 1 + every*k as Binomial(I, obs_prob) of one trajectory simulated
    on the GPU with the same parameters.  The reference's filter stops
    with "Invalid observation process." as soon as one particle's weight
    is -Inf (R/pfilter.R:437-442), which a binomial observation gives any
    particle holding fewer infected than observed; so (beta, gamma) are
    drawn until the simulated infected total stays at or above
    min_infected at every observation time, where particles near the
    truth hold several times the observed count."""
    from .model import run as run_model
    from .particle_filter import pfilter_tspan
    rng = np.random.default_rng(seed)
    times = 1.0 + every * np.arange(1, n_obs + 1, dtype=float)
    u0 = {"S": np.full(Nn, 99), "I": np.full(Nn, 5), "R": np.zeros(Nn, int)}
    intervals = pfilter_tspan(1.0, times)
    for _ in range(200):
        beta = float(rng.uniform(0.05, 0.4))
        gamma = float(rng.uniform(0.03, 0.2))
        # The truth runs the filter's intervals one after another, each
        # from the previous one's final state (as a particle does), so
        # that it follows the filter's model exactly.
        state = models.SIR_mparse(u0=u0, tspan=[1.0], beta=beta, gamma=gamma).u0
        I = []
        for tb, te in intervals:
            ts = [te] if np.isnan(tb) else [tb, te]
            step = models.SIR_mparse(u0={"S": state[0], "I": state[1], "R": state[2]},
                                     tspan=ts, beta=beta, gamma=gamma)
            run_model(step, seed=int(rng.integers(0, 2**62)), lib=lib)
            state = step.U[:, -1].reshape(Nn, 3).T
            I.append(int(state[1].sum()))
        I = np.asarray(I)
        if I.min() >= min_infected:
            break
    else:
        raise RuntimeError("no parameters of the prior keep the infected total above %d"
                           % min_infected)
    obs = rng.binomial(I, obs_prob)
    model = models.SIR_mparse(u0=u0, tspan=np.array([1.0, times[-1]]), beta=beta,
                              gamma=gamma)
    return FilterSpec(model, times, obs.reshape(-1, 1), n_particles, obs_prob,
                      int(rng.integers(0, 2**62)), 0.0, beta, gamma)
