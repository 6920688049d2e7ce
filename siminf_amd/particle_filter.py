"""Bootstrap particle filter on the device (pfilter, R/pfilter.R:503-560
and pfilter_internal, R/pfilter.R:315-499).

The state of every particle stays in HBM for the whole filter: per
observation interval one launch runs all particles' SSA, and the weights,
the log-likelihood increment, the effective sample size, systematic
resampling (SimInf_particle_filter.c:32-66) and the ancestor gather run on
the device behind it.  The observation process is the binomial or the
poisson form of pfilter's formula (R/pfilter.R:279-295) on the count of one
compartment, per node or of the sum over the nodes.
"""
import ctypes

import numpy as np

from . import _native
from .model import SimInfError, _raise, device_plugin


OBS_DISTRIBUTIONS = {"binomial": 0, "poisson": 1}


class PFilterResult:
    """loglik, ess (one per interval) and the zero-based ancestors
    (intervals x particles) of a filter run."""

    def __init__(self, loglik, ess, ancestors, n_particles):
        self.loglik = loglik
        self.ess = ess
        self.ancestors = ancestors
        self.n_particles = n_particles


def pfilter_tspan(tspan0, times):
    """Intervals of the filter (pfilter_tspan, R/pfilter.R:177-193): row i
    is (begin, end), begin NaN when the interval is one time unit."""
    times = np.asarray(times, dtype=np.float64)
    if len(times) < 1 or np.any(np.diff(times) <= 0):
        raise ValueError("'time' column in data must be an increasing vector.")
    if times[0] < tspan0:
        raise ValueError("data$time[1] must be >= tspan[1].")
    rows = []
    for i, t in enumerate(times):
        if i == 0:
            rows.append((tspan0, t) if tspan0 < t else (np.nan, t))
        elif int(t) - int(times[i - 1]) > 1:
            rows.append((times[i - 1], t))
        else:
            rows.append((np.nan, t))
    return np.asarray(rows, dtype=np.float64)


def prepare(model, obs_compartment, times, obs):
    """Checked filter inputs: (tspan2, obs matrix, obs_nodes)."""
    if model.replicates != 1:
        raise ValueError("the filter runs its particles as replicates of a "
                         "model with one replicate")
    if not 0 <= int(obs_compartment) < model.Nc:
        raise ValueError("obs_compartment out of range")
    tspan2 = pfilter_tspan(model.tspan[0], times)
    obs = np.asarray(obs)
    if obs.ndim == 1:
        obs = obs.reshape(-1, 1)
    if obs.shape[0] != len(tspan2) or obs.shape[1] not in (1, model.Nn):
        raise ValueError("obs must be (number of times) x (1 or Nn)")
    obs = np.ascontiguousarray(np.where(np.isnan(obs.astype(float)), -1, obs)
                               .astype(np.int32))
    return np.ascontiguousarray(tspan2), obs, obs.shape[1]


def pfilter(model, obs_compartment, times, obs, n_particles, obs_prob=0.5,
            seed=None, lib=None, obs_dist="binomial", obs_offset=0.0):
    """Run the filter on the GPU.  obs_compartment is zero-based; obs is
    (number of times) x Nn, or one column for the sum over the nodes;
    missing observations are negative or NaN.  The observed count x of a
    compartment holding X is Binomial(size = X, prob = obs_prob), or with
    obs_dist "poisson" Poisson(lambda = obs_prob * X + obs_offset).  Raises
    SimInfError with the reference's message on failure."""
    if obs_dist not in OBS_DISTRIBUTIONS:
        raise ValueError("Unknown distribution: '%s'." % obs_dist)
    lib = lib or _native.load()
    tspan2, obs, obs_nodes = prepare(model, obs_compartment, times, obs)
    if seed is None:
        seed = int(np.random.randint(0, 2**32 - 1))
    dev = device_plugin(model, lib)
    b = _native.ArgsBundle(model, solver=0, seed=int(seed))
    n_int = len(tspan2)
    loglik = ctypes.c_double()
    ess = np.zeros(n_int)
    anc = np.zeros((n_int, int(n_particles)), dtype=np.int32)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    err = lib.SimInf_hip_pfilter(
        ctypes.byref(b.args), ctypes.c_void_p(dev), int(n_particles), n_int,
        tspan2.ctypes.data_as(dp), int(obs_compartment), OBS_DISTRIBUTIONS[obs_dist],
        float(obs_prob), float(obs_offset), obs_nodes, obs.ctypes.data_as(ip),
        int(seed) & 0xFFFFFFFFFFFFFFFF, ctypes.byref(loglik), ess.ctypes.data_as(dp),
        anc.ctypes.data_as(ip))
    if err:
        _raise(lib, err)
    return PFilterResult(loglik.value, ess, anc, int(n_particles))


__all__ = ["pfilter", "pfilter_tspan", "PFilterResult", "SimInfError",
           "OBS_DISTRIBUTIONS"]
