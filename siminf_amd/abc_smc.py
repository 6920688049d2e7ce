"""Approximate Bayesian computation with sequential Monte Carlo (ABC-SMC,
Toni et al. 2009) over the GPU solver.

Mirrors (what, not how):
  abc / abc_internal         R/abc.R:742-960
  abc_ldata                  R/abc.R:634-731: one node per proposal, the
                             first node replicated n times (10 x
                             n_particles at first, doubling up to 1e5
                             while fewer than half are accepted)
  abc_gdata                  R/abc.R:562-632: one run per proposal
  replicate_first_node       R/abc.R:307-332
  abc_init_tolerance         R/abc.R:373-406
  init_model / post_gen      R/abc.R:598-600, :647-650, :833-834: a
                             gdata proposal's model passes through
                             init_model before its run (ldata: not
                             implemented, the reference's message);
                             post_gen sees the result after each
                             generation
  abc_distance / abc_accept  R/abc.R:517-560
  parse_priors               R/distributions.R:170-232 (gamma, lognormal,
                             normal, uniform priors)
  proposals                  SimInf_abc_proposals, SimInf_abc.c:79-273
  weights                    SimInf_abc_weights, SimInf_abc.c:283-406, on
                             the device (SimInf_hip_abc_weights)

The simulations run on the GPU, one launch per batch of proposals; the
weights, O(n_particles^2) kernel evaluations per generation, run on the
GPU.  Proposals and acceptance are host bookkeeping over arrays, as in the
reference.  Random numbers come from numpy's generator seeded by
``seed`` (the reference draws from R's), so results agree in distribution
only.  With tolerance None and n_init, the tolerance is adaptive
(abc_first_epsilon R/abc.R:408-431, abc_adaptive_tolerance :446-495,
KLIEP R/density_ratio.R:19-135: the first generation keeps the n_particles
closest of n_init proposals; later ones take the quantile the estimated
density ratio of successive generations gives, and stop once it exceeds
0.99 from generation 3).
"""
import ctypes
import re

import numpy as np

from . import _native
from .model import SimInfModel, _raise, run as run_model

KINDS = {"gamma": "g", "lognormal": "l", "normal": "n", "uniform": "u"}


class ABCError(ValueError):
    """An argument the reference rejects (same message)."""


class Prior:
    def __init__(self, parameter, distribution, p1, p2):
        self.parameter = parameter
        self.distribution = distribution
        self.p1 = p1
        self.p2 = p2

    @property
    def kind(self):
        return KINDS[self.distribution]


_PRIOR = re.compile(r"^\s*([A-Za-z.][A-Za-z0-9._]*)?\s*~\s*([A-Za-z_][A-Za-z0-9_]*)\s*"
                    r"\((.*)\)\s*$")


def _check_hyperparameters(dist, h):
    if dist not in KINDS:
        raise ABCError("Unknown distribution: '%s'." % dist)
    if len(h) != 2 or any(np.isnan(h)):
        raise ABCError("Invalid formula specification for %s distribution." % dist)
    if dist == "gamma" and not all(v > 0 for v in h):
        raise ABCError("Invalid distribution: gamma hyperparameters must be > 0.")
    if dist == "lognormal" and h[1] < 0:
        raise ABCError("Invalid distribution: lognormal standard deviation must be >= 0.")
    if dist == "normal" and h[1] < 0:
        raise ABCError("Invalid distribution: normal variance must be >= 0.")
    if dist == "uniform" and h[0] >= h[1]:
        raise ABCError("Invalid distribution: uniform bounds in wrong order.")


def parse_priors(priors):
    """Priors from 'parameter ~ distribution(p1, p2)' strings
    (parse_priors, R/distributions.R:170-232)."""
    if isinstance(priors, str):
        priors = [priors]
    if not isinstance(priors, (list, tuple)) or not all(isinstance(p, str) for p in priors):
        raise ABCError("'priors' must be a formula or a list with formula items.")
    out = []
    for text in priors:
        m = _PRIOR.match(text)
        if not m:
            raise ABCError("Invalid formula specification for distribution.")
        parameter, dist, args = m.group(1) or "", m.group(2), m.group(3)
        parts = [a.strip() for a in args.split(",")]
        if len(parts) > 2:
            raise ABCError("Invalid formula specification for distribution.")
        try:
            h = [float(a) for a in parts] + [np.nan] * (2 - len(parts))
        except ValueError:
            if dist not in KINDS:
                raise ABCError("Unknown distribution: '%s'." % dist)
            raise ABCError("Invalid formula specification for 'priors'.")
        _check_hyperparameters(dist, h)
        out.append(Prior(parameter, dist, h[0], h[1]))
    names = [p.parameter for p in out]
    if len(set(names)) != len(names) or any(not n for n in names):
        raise ABCError("'priors' must have non-duplicated parameter names.")
    return out


def match_priors(model, priors):
    """('ldata' or 'gdata', row indices) of the prior parameters
    (match_priors, R/distributions.R:236-250)."""
    names = [p.parameter for p in priors]
    if all(n in model.ldata_names for n in names):
        return "ldata", [model.ldata_names.index(n) for n in names]
    if all(n in model.gdata_names for n in names):
        return "gdata", [model.gdata_names.index(n) for n in names]
    raise ABCError("All parameters in 'priors' must be either in 'gdata' or 'ldata'.")


def _prior_logdens(p, x, mean):
    """Vectorised log-density of a prior (the support test of proposals)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        if p.kind == "g":
            from scipy import stats
            return stats.gamma.logpdf(x, p.p1, scale=1.0 / p.p2)
        if p.kind == "l":
            from scipy import stats
            return stats.lognorm.logpdf(x, p.p2, scale=np.exp(mean))
        if p.kind == "n":
            from scipy import stats
            return stats.norm.logpdf(x, mean, p.p2)
        return np.where((x >= p.p1) & (x <= p.p2), -np.log(p.p2 - p.p1), -np.inf)


def proposals(priors, n, x, w, sigma, rng):
    """(n x d proposals, ancestors) (SimInf_abc_proposals): from the
    priors for the first generation (x None), else a particle of x drawn
    by weight and moved by N(0, sigma), redrawn until every prior density
    is positive and finite.  Ancestors are zero-based (-1 for the first
    generation)."""
    d = len(priors)
    if x is None:
        xx = np.empty((n, d))
        for k, p in enumerate(priors):
            if p.kind == "g":
                xx[:, k] = rng.gamma(p.p1, 1.0 / p.p2, n)
            elif p.kind == "l":
                xx[:, k] = rng.lognormal(p.p1, p.p2, n)
            elif p.kind == "n":
                xx[:, k] = rng.normal(p.p1, p.p2, n)
            else:
                xx[:, k] = rng.uniform(p.p1, p.p2, n)
        return xx, np.full(n, -1, dtype=np.int64)
    w = np.asarray(w, dtype=np.float64)
    if not np.all(np.isfinite(w)) or np.any(w < 0):
        raise ABCError("Invalid weight detected (non-finite or < 0.0).")
    cdf = np.cumsum(w)
    L = np.linalg.cholesky(np.asarray(sigma, dtype=np.float64))
    xx = np.empty((n, d))
    anc = np.empty(n, dtype=np.int64)
    todo = np.arange(n)
    while len(todo):
        j = np.searchsorted(cdf, rng.random(len(todo)) * cdf[-1], side="right")
        j = np.minimum(j, len(cdf) - 1)
        prop = x[j] + rng.standard_normal((len(todo), d)) @ L.T
        ok = np.ones(len(todo), dtype=bool)
        for k, p in enumerate(priors):
            lp = _prior_logdens(p, prop[:, k], x[j, k])
            ok &= np.isfinite(lp)
        xx[todo[ok]] = prop[ok]
        anc[todo[ok]] = j[ok]
        todo = todo[~ok]
    return xx, anc


def call_weights(fn, distribution, p1, p2, x, xx, w, sigma):
    """Marshal the arguments of SimInf_hip_abc_weights (or the checker's
    restatement) and call fn; returns (error code, weights)."""
    xx = np.asfortranarray(np.asarray(xx, dtype=np.float64))
    n, d = xx.shape
    dp = ctypes.POINTER(ctypes.c_double)
    ww = np.zeros(n)
    if w is None:
        args = (distribution.encode(), None, None, n, d, None, None, None, None)
    else:
        x = np.asfortranarray(np.asarray(x, dtype=np.float64))
        p1 = np.ascontiguousarray(p1, dtype=np.float64)
        p2 = np.ascontiguousarray(p2, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        sigma = np.asfortranarray(np.asarray(sigma, dtype=np.float64).reshape(d, d))
        args = (distribution.encode(), p1.ctypes.data_as(dp), p2.ctypes.data_as(dp), n, d,
                x.ctypes.data_as(dp), xx.ctypes.data_as(dp), w.ctypes.data_as(dp),
                sigma.ctypes.data_as(dp))
    err = fn(*args, ww.ctypes.data_as(dp))
    return err, ww


def weights(priors, x, xx, ancestor, w, sigma, lib=None):
    """abc_weights (R/abc.R:733-740) on the device: the previous particles
    taken in the order of the current particles' ancestors."""
    lib = lib or _native.load()
    dist = "".join(p.kind for p in priors)
    xa = None if x is None else np.asarray(x)[np.asarray(ancestor)]
    err, ww = call_weights(lib.SimInf_hip_abc_weights, dist,
                           [p.p1 for p in priors], [p.p2 for p in priors], xa, xx, w, sigma)
    if err:
        _raise(lib, err)
    return ww


def init_tolerance(tolerance):
    """(summaries x generations) schedule (abc_init_tolerance)."""
    t = np.asarray(tolerance, dtype=np.float64)
    if t.dtype.kind not in "fiu":
        raise ABCError("'tolerance' must have non-negative values.")
    if t.ndim < 2:
        t = t.reshape(1, -1)
    if t.shape[1] == 0:
        raise ABCError("'tolerance' must have columns.")
    if np.any(np.isnan(t)) or np.any(t < 0):
        raise ABCError("'tolerance' must have non-negative values.")
    if t.shape[1] > 1 and np.any(np.diff(t, axis=1) >= 0):
        raise ABCError("'tolerance' must be a decreasing vector.")
    return t


def check_distance(distance, n):
    """abc_distance (R/abc.R:517-551)."""
    d = np.asarray(distance)
    if d.dtype.kind not in "fiu":
        raise ABCError("The result from the ABC distance function must be numeric.")
    d = d.astype(np.float64)
    if d.ndim == 1:
        d = d.reshape(-1, 1)
    if d.shape[0] != n:
        raise ABCError("Invalid dimension of the result from the ABC distance function.")
    if np.any(np.isnan(d)) or np.any(d < 0):
        raise ABCError("The result from the ABC distance function must be non-negative.")
    return d


def accept(distance, tolerance):
    if tolerance is None:
        # Every proposal of the adaptive first generation is kept.
        if distance.shape[1] != 1:
            raise ABCError("Adaptive tolerance must have one summary statistic.")
        return np.ones(distance.shape[0], dtype=bool)
    """abc_accept (R/abc.R:553-560)."""
    if distance.shape[1] != len(tolerance):
        raise ABCError("Mismatch between the number of summary statistics and tolerance.")
    return np.all(distance <= tolerance, axis=1)


def replicate_first_node(model, n):
    """The model's first node n times (R/abc.R:307-332), events of the
    first node repeated per copy with node offsets."""
    u0 = model.u0[:, [0] * n]
    v0 = model.v0[:, [0] * n] if model.Nd else None
    ld = model.ldata[:, [0] * n] if model.Nld else None
    ev = None
    k = len(model.events["event"])
    if k:
        ev = {c: np.tile(model.events[c], n) for c in
              ("event", "time", "node", "dest", "n", "proportion", "select", "shift")}
        ev["node"] = ev["node"] + np.repeat(np.arange(n), k)
    return SimInfModel(model.plugin, model.G, model.S, model.tspan, u0, ldata=ld,
                       gdata=model.gdata, v0=v0, E=model.E, N=model.N, events=ev,
                       compartments=model.compartments, v0_names=model.v0_names,
                       ldata_names=model.ldata_names, gdata_names=model.gdata_names)


class ABCResult:
    """Accepted particles x (generations, n_particles, parameters), their
    weights, distances and tolerances per generation, the ESS and the
    number of proposals per generation."""

    def __init__(self, parameters):
        self.parameters = parameters
        self.x, self.weight, self.distance, self.tolerance = [], [], [], []
        self.ess, self.nprop = [], []


def _generation_ldata(model, pars, priors, n_particles, fn, generation, eps, x, w, sigma,
                      rng, data, solver, lib):
    n = 10 * n_particles
    batch = replicate_first_node(model, n)
    xs, ancs, ds = [], [], []
    got = nprop = 0
    while got < n_particles:
        if n < 100000 and nprop > 2 * n:
            n = min(100000, 2 * n)
            batch = replicate_first_node(model, n)
        xx, anc = proposals(priors, n, x, w, sigma, rng)
        for k, r in enumerate(pars):
            batch.ldata[r, :] = xx[:, k]
        run_model(batch, solver=solver, seed=int(rng.integers(0, 2**63 - 1)), lib=lib)
        d = check_distance(fn(batch, generation, data), n)
        ok = accept(d, eps)
        csum = np.cumsum(ok) + got
        hit = np.nonzero(csum == n_particles)[0]
        j = np.nonzero(ok)[0]
        if len(hit):
            j = j[j <= hit[0]]
            nprop += int(hit[0]) + 1
        else:
            nprop += n
        xs.append(xx[j])
        ancs.append(anc[j])
        ds.append(d[j])
        got += len(j)
    return np.vstack(xs), np.concatenate(ancs), np.vstack(ds), nprop


def _generation_gdata(model, pars, priors, n_particles, fn, generation, eps, x, w, sigma,
                      rng, data, solver, lib, init_model=None):
    xs, ancs, ds = [], [], []
    nprop = 0
    while len(xs) < n_particles:
        xx, anc = proposals(priors, 1, x, w, sigma, rng)
        # Each proposal's model owns copies of the arrays an init_model
        # callback may change in place (the reference's models are R
        # values, copied on modification).
        events = {k: model.events[k].copy() for k in
                  ("event", "time", "node", "dest", "n", "proportion", "select", "shift")}
        m = SimInfModel(model.plugin, model.G, model.S, model.tspan, model.u0.copy(),
                        ldata=model.ldata.copy(), gdata=model.gdata.copy(),
                        v0=model.v0.copy() if model.Nd else None, E=model.E.copy(),
                        N=None if model.N is None else model.N.copy(),
                        events=events, compartments=model.compartments,
                        v0_names=model.v0_names, ldata_names=model.ldata_names,
                        gdata_names=model.gdata_names)
        for k, r in enumerate(pars):
            m.gdata[r] = xx[0, k]
        if init_model is not None:
            m = init_model(m)
        run_model(m, solver=solver, seed=int(rng.integers(0, 2**63 - 1)), lib=lib)
        d = check_distance(fn(m, generation, data), 1)
        nprop += 1
        if accept(d, eps)[0]:
            xs.append(xx[0])
            ancs.append(anc[0])
            ds.append(d[0])
    return np.vstack(xs), np.asarray(ancs), np.vstack(ds), nprop


def kliep_kernel(x, y, sigma):
    """Gaussian kernel matrix exp(-|x_i - y_j|^2 / (2 sigma^2))."""
    d2 = ((x[:, None, :] - y[None, :, :]) ** 2).sum(axis=2)
    return np.exp(-d2 / (2.0 * sigma * sigma))


def _ginv1(a):
    return 0.0 if a == 0.0 else 1.0 / a


def kliep_alpha(alpha, mean_xde):
    """Projection of the kernel weights (KLIEP_alpha)."""
    alpha = alpha + mean_xde * (1.0 - float(mean_xde @ alpha)) * _ginv1(float(mean_xde @ mean_xde))
    alpha = np.maximum(alpha, 0.0)
    return alpha * _ginv1(float(mean_xde @ alpha))


def kliep_learning(mean_xde, xnu):
    """Gradient ascent on the mean log density ratio (KLIEP_learning)."""
    alpha = kliep_alpha(np.ones(xnu.shape[1]), mean_xde)
    xa = xnu @ alpha
    score = np.mean(np.log(xa))
    for eps in 10.0 ** np.arange(3, -4, -1):
        for _ in range(100):
            new = kliep_alpha(alpha + eps * (xnu.T @ (1.0 / xa)), mean_xde)
            xa_new = xnu @ new
            with np.errstate(divide="ignore"):
                score_new = np.mean(np.log(xa_new))
            if not score_new > score:
                break
            alpha, xa, score = new, xa_new, score_new
    return alpha


def kliep(xnu, xde, rng, fold=5, kernels=100):
    """Kullback-Leibler importance estimation with cross validation of the
    kernel width (KLIEP); returns (centers, sigma, weights)."""
    if not (isinstance(xnu, np.ndarray) and xnu.ndim == 2 and xnu.dtype.kind in "fiu"):
        raise ABCError("'xnu' must be a numeric matrix.")
    if not (isinstance(xde, np.ndarray) and xde.ndim == 2 and xde.dtype.kind in "fiu"):
        raise ABCError("'xde' must be a numeric matrix.")
    if xnu.shape[1] != xde.shape[1]:
        raise ABCError("'xnu' and 'xde' must have the same dimension.")
    xnu = xnu.astype(np.float64)
    xde = xde.astype(np.float64)
    kernels = min(kernels, xnu.shape[0])
    centers = xnu[rng.choice(xnu.shape[0], size=kernels, replace=False)]
    sigma, score = 10.0, -np.inf
    for e in np.arange(np.log10(sigma) - 1, -6, -1):
        for _ in range(9):
            sigma_new = sigma - 10.0 ** e
            Xnu = kliep_kernel(xnu, centers, sigma_new)
            mean_xde = kliep_kernel(xde, centers, sigma_new).mean(axis=0)
            cv = rng.permutation(xnu.shape[0]) % fold
            score_new = 0.0
            for i in range(fold):
                a = kliep_learning(mean_xde, Xnu[cv != i])
                with np.errstate(divide="ignore"):
                    score_new += np.mean(np.log(Xnu[cv == i] @ a)) / fold
            if not score_new > score:
                break
            sigma, score = sigma_new, score_new
    Xnu = kliep_kernel(xnu, centers, sigma)
    mean_xde = kliep_kernel(xde, centers, sigma).mean(axis=0)
    return centers, sigma, kliep_learning(mean_xde, Xnu)


def kliep_density_ratio(x, centers, sigma, weights):
    return kliep_kernel(np.atleast_2d(x), centers, sigma) @ weights


def first_epsilon(distance, n_particles):
    """The n_particles-th smallest distance (abc_first_epsilon)."""
    i = np.argsort(distance.sum(axis=1), kind="stable")
    return distance[i[n_particles - 1]]


def adaptive_tolerance(xnu, xde, distance, generation, rng):
    """Next tolerance from the estimated density ratio of the current
    (xnu) and previous (xde) particles, or None when the stopping rule
    applies (abc_adaptive_tolerance)."""
    from scipy import optimize
    centers, sigma, wts = kliep(xnu, xde, rng)
    f = lambda z: -float(kliep_density_ratio(np.atleast_1d(z), centers, sigma, wts)[0])
    if xnu.shape[1] > 1:
        res = optimize.minimize(f, xnu[0], method="Nelder-Mead")
        c_t = -res.fun
    else:
        res = optimize.minimize_scalar(f, bounds=(xnu.min(), xnu.max()), method="bounded")
        c_t = -res.fun
    if not (np.isfinite(c_t) and c_t > 0.0):
        raise ABCError("Unable to compute the adaptive tolerance: the density ratio's "
                       "supremum is not finite and positive.")
    q_t = 1.0 / c_t
    if q_t > 0.99 and generation >= 3:
        return None
    # The rank is clamped to [1, n]: an optimizer that stops below the
    # supremum of the ratio (q_t > 1) takes the largest distance.
    i = np.argsort(distance.sum(axis=1), kind="stable")
    n = distance.shape[0]
    rank = min(max(int(np.ceil(q_t * n)), 1), n)
    return distance[i[rank - 1]]


def abc(model, priors, n_particles, distance, tolerance=None, data=None, seed=None,
        solver="ssm", lib=None, post_gen=None, init_model=None, n_init=None):
    """ABC-SMC (abc, R/abc.R:742-960).  With a tolerance schedule (one
    column per generation) the generations follow it; with tolerance None
    and n_init the tolerance is adaptive (see the module notes).
    distance(result, generation, data) returns one row of non-negative
    distances per proposal: per node for ldata parameters (the run model
    holds one node per proposal), one row for gdata parameters.
    init_model(model) -> model is applied to each gdata proposal's model
    before its run; post_gen(result) is called after each generation."""
    if init_model is not None and not callable(init_model):
        raise ABCError("'init_model' must be 'NULL' or a 'function'.")
    if post_gen is not None and not callable(post_gen):
        raise ABCError("'post_gen' must be 'NULL' or a 'function'.")
    if tolerance is None and n_init is None:
        raise ABCError("Both 'n_init' and 'tolerance' can not be NULL.")
    lib = lib or _native.load()
    n_particles = int(n_particles)
    if n_particles < 2:
        raise ABCError("'n_particles' must be an integer > 1.")
    if n_init is not None:
        n_init = int(n_init)
        if n_init <= 1:
            raise ABCError("'n_init' must be an integer > 1.")
        if n_init <= n_particles:
            raise ABCError("'n_init' must be an integer > 'n_particles'.")
        if tolerance is not None:
            raise ABCError("'tolerance' must be NULL for adaptive distance.")
    priors = parse_priors(priors)
    target, pars = match_priors(model, priors)
    tol = None if tolerance is None else init_tolerance(tolerance)
    rng = np.random.default_rng(seed)
    if target == "ldata" and init_model is not None:
        raise ABCError("'init_model' callback is not implemented for 'ldata'.")
    step = _generation_ldata if target == "ldata" else _generation_gdata
    extra = {} if target == "ldata" else {"init_model": init_model}
    out = ABCResult([p.parameter for p in priors])
    x = w = None
    eps = None if tol is None else tol[:, 0]
    g = 0
    while True:
        adaptive_first = tol is None and g == 0
        sigma = None if x is None else 2.0 * np.atleast_2d(np.cov(x, rowvar=False))
        xx, anc, d, nprop = step(model, pars, priors, n_init if adaptive_first else n_particles,
                                 distance, g + 1, eps, x, w, sigma, rng, data, solver, lib,
                                 **extra)
        if adaptive_first:
            eps = first_epsilon(d, n_particles)
            keep = np.argsort(d.sum(axis=1), kind="stable")[:n_particles]
            x_old = xx
            xx, anc, d = xx[keep], np.asarray(anc)[keep], d[keep]
        else:
            x_old = x
        w = weights(priors, x, xx, anc, w, sigma, lib)
        x = xx
        out.x.append(xx)
        out.weight.append(w)
        out.distance.append(d)
        out.tolerance.append(np.asarray(eps, dtype=np.float64))
        out.ess.append(1.0 / np.sum(w * w))
        out.nprop.append(nprop)
        if post_gen is not None:
            post_gen(out)
        g += 1
        if tol is not None:
            if g >= tol.shape[1]:
                break
            eps = tol[:, g]
        else:
            eps = adaptive_tolerance(x, x_old, d, g + 1, rng)
            if eps is None:
                break
    return out


__all__ = ["abc", "ABCResult", "ABCError", "parse_priors", "proposals", "weights", "kliep",
           "adaptive_tolerance", "first_epsilon"]
