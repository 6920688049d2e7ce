"""A serving loop for particle-filter requests: the inner loop of pfilter
and PMCMC (R/pfilter.R:315-499, called once per proposal at
R/pmcmc.R:351 and R/pmcmc.R:725) as a stream of independent requests.

Continuous admission: a request is admitted as soon as a slot (at most
``max_inflight`` requests on the device) and its HBM footprint are free;
it then runs the device-resident filter (`SimInf_hip_pfilter`) on its own
non-blocking HIP stream with stream-ordered allocations, so requests that
arrive while others run are launched beside them and their kernels
overlap on the GPU.  A request's result does not depend on what else is in
flight: every random stream of the filter is keyed by the request's own
seed, and the filter's sums run in a fixed order.

`host_pfilter` is the plain path: one `run()` per observation interval
(plan, upload, launch, download), the weights, log-sum-exp and systematic
resampling in numpy, the particles' states uploaded again for the next
interval -- what the reference's R loop does around run().
"""
import queue
import threading
import time

import numpy as np

from . import _native, particle_filter
from .model import SimInfModel, run as run_model


def footprint(model, n_particles):
    """HBM bytes a filter request allocates: per (particle, node) lane the
    initial state, the two trajectory columns and the resampled state of u
    (4 bytes per compartment) and the same of v (8 bytes per continuous
    state, at least one), plus the per-particle weights and ancestors and
    the model's ldata and small tables (bounded by 1 MiB)."""
    lanes = int(model.Nn) * int(n_particles)
    nd = max(int(model.Nd), 1)
    per_lane = 4 * model.Nc * 4 + 8 * nd * 4
    return (lanes * per_lane + 16 * int(n_particles) + 8 * model.Nld * model.Nn
            + (1 << 20))


class Request:
    """One filter request; ``result`` (PFilterResult) or ``error`` is set
    when ``done`` is."""

    def __init__(self, model, obs_compartment, times, obs, n_particles,
                 obs_prob=0.5, seed=0, tag=None, obs_dist="binomial", obs_offset=0.0):
        self.model = model
        self.obs_dist = obs_dist
        self.obs_offset = obs_offset
        self.obs_compartment = obs_compartment
        self.times = times
        self.obs = obs
        self.n_particles = int(n_particles)
        self.obs_prob = obs_prob
        self.seed = seed
        self.tag = tag
        self.bytes = footprint(model, n_particles)
        self.t_arrival = None
        self.t_admit = None
        self.t_start = None
        self.t_done = None
        self.result = None
        self.error = None
        self.done = threading.Event()

    @property
    def latency(self):
        return self.t_done - self.t_arrival

    def wait(self, timeout=None):
        if not self.done.wait(timeout):
            raise TimeoutError("request not done")
        if self.error is not None:
            raise self.error
        return self.result


def _run_filter(req, lib):
    return particle_filter.pfilter(req.model, req.obs_compartment, req.times, req.obs,
                                   req.n_particles, obs_prob=req.obs_prob, seed=req.seed,
                                   lib=lib, obs_dist=req.obs_dist,
                                   obs_offset=req.obs_offset)


class PFilterServer:
    """Admission by slots and HBM footprint, first come first served: a
    request that does not fit waits, and the requests behind it wait for it
    (no overtaking, so a large request is not starved)."""

    def __init__(self, hbm_budget=64 << 30, max_inflight=8, lib=None, runner=None):
        self.runner = runner or _run_filter
        self.lib = lib if lib is not None or runner is not None else _native.load()
        self.budget = int(hbm_budget)
        self.free = int(hbm_budget)
        self.max_inflight = int(max_inflight)
        self.inflight = 0
        self.peak_inflight = 0
        self.peak_bytes = 0
        self._queue = queue.Queue()
        self._cv = threading.Condition()
        self._dispatcher = None
        self._workers = []

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    def start(self):
        self._dispatcher = threading.Thread(target=self._dispatch, daemon=True)
        self._dispatcher.start()
        return self

    def submit(self, req):
        if req.bytes > self.budget:
            raise ValueError("request needs %d bytes of HBM, over the budget of %d"
                             % (req.bytes, self.budget))
        req.t_arrival = time.perf_counter()
        self._queue.put(req)
        return req

    def _dispatch(self):
        while True:
            req = self._queue.get()
            if req is None:
                return
            with self._cv:
                while self.inflight >= self.max_inflight or self.free < req.bytes:
                    self._cv.wait()
                self.free -= req.bytes
                self.inflight += 1
                self.peak_inflight = max(self.peak_inflight, self.inflight)
                self.peak_bytes = max(self.peak_bytes, self.budget - self.free)
                req.t_admit = time.perf_counter()
            t = threading.Thread(target=self._serve, args=(req,), daemon=True)
            self._workers.append(t)
            t.start()

    def _serve(self, req):
        req.t_start = time.perf_counter()
        try:
            req.result = self.runner(req, self.lib)
        except Exception as e:        # handed to the caller with the request
            req.error = e
        req.t_done = time.perf_counter()
        with self._cv:
            self.free += req.bytes
            self.inflight -= 1
            self._cv.notify_all()
        req.done.set()

    def stop(self):
        self._queue.put(None)
        if self._dispatcher is not None:
            self._dispatcher.join()
        for t in self._workers:
            t.join()
        self._workers = []


def host_pfilter(model, obs_compartment, times, obs, n_particles, obs_prob=0.5,
                 seed=0, lib=None, obs_dist="binomial", obs_offset=0.0):
    """The filter driven from the host with one run() per interval (the
    plain path).  Returns (loglik, ess).  Models with scheduled events are
    refused: this baseline does not split them over the intervals."""
    from scipy import stats
    if len(model.events["event"]):
        raise ValueError("host_pfilter runs models without scheduled events.")
    lib = lib or _native.load()
    tspan2, obs, obs_nodes = particle_filter.prepare(model, obs_compartment, times, obs)
    P = int(n_particles)
    Nn, Nc, Nd = model.Nn, model.Nc, model.Nd
    rng = np.random.default_rng(seed)
    u = np.tile(model.u0, (1, P))
    v = np.tile(model.v0, (1, P)) if Nd else None
    loglik = 0.0
    ess = np.zeros(len(tspan2))
    for i, (tb, te) in enumerate(tspan2):
        ts = np.array([te]) if np.isnan(tb) else np.array([tb, te])
        m = SimInfModel(model.plugin, model.G, model.S, ts, u, ldata=model.ldata,
                        gdata=model.gdata, v0=v, E=model.E, replicates=P,
                        compartments=model.compartments)
        run_model(m, seed=int(rng.integers(0, 2**63 - 1)), lib=lib)
        tlen = len(ts)
        last = m.U.reshape(Nn * Nc, P, tlen)[:, :, -1]          # row node*Nc+c
        x = last.reshape(Nn, Nc, P)[:, obs_compartment, :]      # Nn x P
        if obs_dist == "poisson":
            def logw(o, X):
                return stats.poisson.logpmf(o, obs_prob * X + obs_offset)
        else:
            def logw(o, X):
                return stats.binom.logpmf(o, X, obs_prob)
        if obs_nodes == 1:
            w = logw(obs[i, 0], x.sum(axis=0)) if obs[i, 0] >= 0 else np.zeros(P)
        else:
            o = obs[i][:, None]
            w = np.where(o >= 0, logw(np.maximum(o, 0), x), 0.0).sum(axis=0)
        if not np.all(np.isfinite(w)):
            raise ValueError("Invalid observation process.")
        mw = w.max()
        w = np.exp(w - mw)
        loglik += mw + np.log(w.sum()) - np.log(P)
        w /= w.sum()
        ess[i] = 1.0 / np.sum(w * w)
        c = np.cumsum(w)
        pos = (rng.random() + np.arange(P)) / P
        anc = np.minimum(np.searchsorted(c, pos * c[-1]), P - 1)
        # Column rep*Nn + node of the next u0 is the ancestor's node column.
        u = np.asfortranarray(last[:, anc].reshape(Nc, Nn * P, order="F"))
        if Nd:
            lastv = m.V.reshape(Nn * Nd, P, tlen)[:, :, -1]
            v = np.asfortranarray(lastv[:, anc].reshape(Nd, Nn * P, order="F"))
    return loglik, ess
