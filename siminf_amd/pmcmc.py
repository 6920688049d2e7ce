"""Particle Markov chain Monte Carlo over the device particle filter
(R/pmcmc.R: pmcmc :109-190, continue_pmcmc :399-486, pmcmc_proposal
:346-359): a random-walk Metropolis-Hastings chain over the parameters
that the priors name, whose likelihood is the bootstrap filter's estimate
at the proposal.

Each iteration is one filter request.  Several chains run side by side,
one thread each, through a `serving.PFilterServer`, so the requests of
different chains are admitted as slots and HBM free up and their kernels
overlap on the GPU; a chain's result does not depend on what else runs,
since each filter request's random streams are keyed by its own seed.

Restated from the reference:
  * the first iteration filters at theta (from the priors when not
    given) and records logPost, logLik, logPrior, accept = 0, theta;
  * a proposal is theta + z Q with z standard normal and Q'Q the
    proposal covariance: the initial covmat (by default
    diag((theta / 10)^2 / npars)) while the iteration is within the first
    `adaptive`, nothing has been accepted yet, or with probability
    adaptmix, and otherwise 2.38^2 / npars times the chain's empirical
    covariance (running mean and scatter updated every iteration);
  * a proposal outside the priors' support is rejected without a filter
    run; otherwise it is accepted with probability
    exp(logLik' + logPrior' - logLik - logPrior).
Q here is the symmetric square root of the covariance (eigenvalues below
zero taken as zero) where the reference takes a pivoted Cholesky factor:
both give z Q the distribution N(0, covariance).  Priors are the
R-level ones (R/distributions.R:256-300: gamma with shape p1 and rate
1 / p2, lognormal meanlog / sdlog, normal mean / sd, uniform min / max).
R's random numbers are not available, so a chain is reproducible from its
seed within this library, not against the reference's.
"""
import copy
import math
import threading

import numpy as np

from . import particle_filter, serving
from .abc_smc import ABCError, match_priors, parse_priors

COLUMNS = ("logPost", "logLik", "logPrior", "accept")


class PMCMCError(ValueError):
    """An argument the reference rejects (same message)."""


def dpriors(theta, priors):
    """Sum of the priors' log-densities at theta (R/distributions.R:274-300)."""
    from scipy import stats
    total = 0.0
    for x, p in zip(np.atleast_1d(theta), priors):
        if p.distribution == "gamma":
            total += stats.gamma.logpdf(x, p.p1, scale=p.p2)
        elif p.distribution == "lognormal":
            total += stats.lognorm.logpdf(x, p.p2, scale=math.exp(p.p1))
        elif p.distribution == "normal":
            total += stats.norm.logpdf(x, p.p1, p.p2)
        else:
            total += stats.uniform.logpdf(x, p.p1, p.p2 - p.p1)
    return float(total)


def rpriors(priors, rng):
    """One draw from each prior (R/distributions.R:256-270)."""
    out = []
    for p in priors:
        if p.distribution == "gamma":
            out.append(rng.gamma(p.p1, p.p2))
        elif p.distribution == "lognormal":
            out.append(rng.lognormal(p.p1, p.p2))
        elif p.distribution == "normal":
            out.append(rng.normal(p.p1, p.p2))
        else:
            out.append(rng.uniform(p.p1, p.p2))
    return np.array(out, dtype=float)


def cov_root(covmat):
    """Q with Q'Q = covmat (symmetric square root; negative eigenvalues of
    a covariance that is not quite positive semidefinite count as 0)."""
    c = np.asarray(covmat, dtype=float)
    w, v = np.linalg.eigh((c + c.T) / 2.0)
    return (v * np.sqrt(np.clip(w, 0.0, None))) @ v.T


class PMCMCResult:
    """One chain: ``chain`` (iterations x (4 + parameters)) with columns
    ``columns`` (logPost, logLik, logPrior, accept, then the parameters),
    and the settings it ran with."""

    def __init__(self, parameters, n_particles, covmat, adaptmix, adaptive, target, pars):
        self.parameters = list(parameters)
        self.columns = list(COLUMNS) + self.parameters
        self.n_particles = int(n_particles)
        self.covmat = np.asarray(covmat, dtype=float)
        self.adaptmix = float(adaptmix)
        self.adaptive = int(adaptive)
        self.target = target
        self.pars = list(pars)
        self.chain = np.zeros((0, len(self.columns)))
        self.latencies = []       # seconds from submission to result, per filter request

    def __len__(self):
        return self.chain.shape[0]

    def acceptance_ratio(self):
        """Mean of the accept column (acceptance_ratio, R/pmcmc.R:72-74)."""
        return float(np.nanmean(self.chain[:, 3])) if len(self) else float("nan")

    def theta(self, i=-1):
        return self.chain[i, 4:].copy()


def _check_integer(name, x):
    """check_integer_arg (R/check_arguments.R:57-78)."""
    if x is None:
        raise PMCMCError("'%s' is missing." % name)
    a = np.atleast_1d(np.asarray(x, dtype=object))
    ok = all(isinstance(v, (int, float, np.integer, np.floating)) and
             not isinstance(v, (bool, np.bool_)) and
             math.isfinite(float(v)) and float(v).is_integer() for v in a)
    if not ok:
        raise PMCMCError("'%s' must be integer." % name)
    return np.array([int(v) for v in a])


def _check_args(n_particles, n_iterations, adaptmix, adaptive):
    """The reference's checks and messages (R/check_arguments.R:365-371,
    R/pmcmc.R:194-236)."""
    npart = _check_integer("n_particles", n_particles)
    if npart.size != 1 or npart[0] <= 1:
        raise PMCMCError("'n_particles' must be an integer > 1.")
    a = np.atleast_1d(np.asarray(adaptmix, dtype=float))
    if a.size != 1 or not (0.0 < a[0] < 1.0):
        raise PMCMCError("'adaptmix' must be a value > 0 and < 1.")
    ad = _check_integer("adaptive", adaptive)
    if ad.size != 1 or ad[0] < 0:
        raise PMCMCError("'adaptive' must be an integer >= 0.")
    it = _check_integer("n_iterations", n_iterations)
    if it.size != 1 or it[0] <= 0:
        raise PMCMCError("'n_iterations' must be an integer > 0.")
    return int(npart[0]), int(it[0]), float(a[0]), int(ad[0])


def with_theta(model, target, pars, theta):
    """A copy of the model with the parameters set (set_proposal,
    R/pmcmc.R:312-323); the copy owns its ldata / gdata."""
    m = copy.copy(model)
    if target == "gdata":
        g = np.array(model.gdata, dtype=float, copy=True)
        g[list(pars)] = theta
        m.gdata = g
    else:
        ld = np.array(model.ldata, dtype=float, copy=True, order="F")
        for k, x in zip(pars, theta):
            ld[k, :] = x
        m.ldata = ld
    m.U = None
    m.V = None
    return m


def _proposal(res, i, n_accepted, emp_cov, rng):
    """pmcmc_proposal (R/pmcmc.R:346-359); i is the 1-based iteration."""
    npars = len(res.parameters)
    if (res.adaptive == 0 or i <= res.adaptive or n_accepted == 0 or
            rng.uniform() < res.adaptmix):
        cov = res.covmat
    else:
        cov = 2.38 ** 2 / npars * emp_cov
    return res.chain[i - 2, 4:] + rng.standard_normal(npars) @ cov_root(cov)


def run_chain(res, priors, theta, n_iterations, loglik_fn, rng):
    """Fill res.chain with n_iterations iterations from theta
    (pmcmc :150-178 and continue_pmcmc :409-483).  loglik_fn(theta, seed)
    returns the filter's log-likelihood at theta."""
    npars = len(res.parameters)
    chain = np.full((n_iterations, 4 + npars), np.nan)
    res.chain = chain

    def filter_seed():
        return int(rng.integers(0, 2**63 - 1))

    logLik = float(loglik_fn(theta, filter_seed()))
    logPrior = dpriors(theta, priors)
    chain[0] = np.concatenate([[logLik + logPrior, logLik, logPrior, 0.0], theta])
    n_accepted = 0
    M = np.array(theta, dtype=float)
    S = np.zeros((npars, npars))
    emp_cov = np.zeros((npars, npars))
    for i in range(2, n_iterations + 1):
        accept = 0.0
        prop = _proposal(res, i, n_accepted, emp_cov, rng)
        lp_prop = dpriors(prop, priors)
        if math.isfinite(lp_prop):
            ll_prop = float(loglik_fn(prop, filter_seed()))
            with np.errstate(over="ignore", invalid="ignore"):
                alpha = float(np.exp(ll_prop + lp_prop - logLik - logPrior))
            # A ratio that is not finite (including one that overflows) is
            # not accepted, as in the reference.
            if math.isfinite(alpha) and rng.uniform() < alpha:
                logLik, logPrior, theta = ll_prop, lp_prop, prop
                accept = 1.0
                n_accepted += 1
        chain[i - 1] = np.concatenate([[logLik + logPrior, logLik, logPrior, accept], theta])
        d = theta - M
        M = M + d / i
        S = S + np.outer(d, theta - M)
        emp_cov = S / (i - 1)
    return res


def pmcmc(model, obs_compartment, times, obs, priors, n_particles, n_iterations,
          theta=None, covmat=None, adaptmix=0.05, adaptive=100, obs_prob=0.5,
          obs_dist="binomial", obs_offset=0.0, seed=None, chains=1, server=None,
          loglik_fn=None, lib=None):
    """Run `chains` PMCMC chains of n_iterations each (pmcmc, R/pmcmc.R:109-190).

    The observation process is the filter's (`particle_filter.pfilter`:
    compartment obs_compartment, binomial with obs_prob or Poisson with
    obs_prob * X + obs_offset).  Chain c draws from numpy's generator
    seeded with (seed, c).  With chains > 1 the chains run concurrently
    through `server` (a started `serving.PFilterServer`, or one made here).
    loglik_fn(model, seed) replaces the device filter (the tests pass the
    checker's).  Returns a PMCMCResult, or a list of them with chains > 1.
    """
    n_particles, n_iterations, adaptmix, adaptive = _check_args(
        n_particles, n_iterations, adaptmix, adaptive)
    pri = parse_priors(priors)
    target, pars = match_priors(model, pri)
    names = [p.parameter for p in pri]
    if seed is None:
        seed = int(np.random.randint(0, 2**31 - 1))
    rngs = [np.random.default_rng([int(seed), c]) for c in range(int(chains))]

    thetas = []
    for rng in rngs:
        if theta is None:
            th = rpriors(pri, rng)
        elif (isinstance(theta, dict) and all(n in theta for n in names) and
              all(isinstance(theta[n], (int, float, np.integer, np.floating)) and
                  not isinstance(theta[n], (bool, np.bool_)) for n in names)):
            th = np.array([float(theta[n]) for n in names])
        elif (not isinstance(theta, (dict, str)) and np.ndim(theta) == 1 and
              len(theta) == len(names)):
            th = np.asarray(theta, dtype=float)      # in the priors' order
        else:
            raise PMCMCError("'theta' must be a vector with initial values for the parameters.")
        thetas.append(th)
    results = []
    for th in thetas:
        cov = (np.diag((th / 10.0) ** 2 / len(th)) if covmat is None
               else np.asarray(covmat, dtype=float).reshape(len(th), len(th)))
        results.append(PMCMCResult(names, n_particles, cov, adaptmix, adaptive, target, pars))

    own_server = None
    if loglik_fn is None and int(chains) > 1 and server is None:
        own_server = server = serving.PFilterServer(max_inflight=int(chains), lib=lib).start()

    def device_loglik(res):
        def f(th, s):
            m = with_theta(model, target, pars, th)
            if server is not None:
                req = serving.Request(m, obs_compartment, times, obs, n_particles,
                                      obs_prob=obs_prob, seed=s, obs_dist=obs_dist,
                                      obs_offset=obs_offset)
                ll = server.submit(req).wait().loglik
                res.latencies.append(req.latency)
                return ll
            return particle_filter.pfilter(m, obs_compartment, times, obs, n_particles,
                                           obs_prob=obs_prob, seed=s, lib=lib,
                                           obs_dist=obs_dist, obs_offset=obs_offset).loglik
        return f

    def fn_for(res):
        if loglik_fn is not None:
            return lambda th, s: loglik_fn(with_theta(model, target, pars, th), s)
        return device_loglik(res)

    errors = []

    def go(k):
        try:
            run_chain(results[k], pri, thetas[k], n_iterations, fn_for(results[k]), rngs[k])
        except Exception as e:      # raised to the caller below
            errors.append(e)

    try:
        if len(results) == 1:
            go(0)
        else:
            ts = [threading.Thread(target=go, args=(k,)) for k in range(len(results))]
            for t in ts:
                t.start()
            for t in ts:
                t.join()
    finally:
        if own_server is not None:
            own_server.stop()
    if errors:
        raise errors[0]
    return results[0] if int(chains) == 1 else results


__all__ = ["pmcmc", "PMCMCResult", "PMCMCError", "dpriors", "rpriors", "cov_root",
           "with_theta", "run_chain", "ABCError"]
