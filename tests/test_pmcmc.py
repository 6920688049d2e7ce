"""CPU tests of the PMCMC driver (siminf_amd/pmcmc.py, R/pmcmc.R): the
reference's argument checks and messages (tests/pmcmc.R:26-166), the
prior densities against their closed forms, the proposal square root,
and the Metropolis-Hastings chain against an exact posterior: with a
Gaussian log-likelihood standing in for the filter, the chain's draws
must have the conjugate posterior's mean and variance."""
import importlib
import math

import numpy as np
import pytest

import siminf_amd as si
from siminf_amd.abc_smc import parse_priors

P = importlib.import_module("siminf_amd.pmcmc")

# tests/pmcmc.R:10-24: one SIR node and its observed infected.
TIMES = np.array([1, 4, 7, 10, 13, 16, 19, 22, 25, 28, 31, 34, 37, 40, 43, 46, 49, 52,
                  55, 58, 61, 64, 67, 70, 73, 76, 79, 82, 85, 88, 91, 94, 97, 100])
IOBS = np.array([1, 2, 2, 3, 3, 3, 3, 2, 3, 3, 6, 10, 9, 16, 17, 15, 15, 13, 18, 18, 16,
                 13, 12, 12, 11, 10, 9, 10, 7, 7, 6, 3, 3, 2])


def reference_model():
    return si.SIR(u0={"S": [99], "I": [1], "R": [0]}, tspan=np.arange(1, 101, 3.0),
                  beta=0.16, gamma=0.077)


PRIORS = ["beta ~ uniform(0, 1)", "gamma ~ uniform(0, 1)"]


def _call(**over):
    kw = dict(model=reference_model(), obs_compartment=1, times=TIMES,
              obs=IOBS.reshape(-1, 1), priors=PRIORS, n_particles=200, n_iterations=200,
              obs_prob=1.0, obs_dist="poisson", obs_offset=1e-6,
              loglik_fn=lambda m, s: 0.0)
    kw.update(over)
    return P.pmcmc(**kw)


@pytest.mark.parametrize("over, msg", [
    (dict(n_particles=200.1), "'n_particles' must be integer."),
    (dict(n_particles=float("nan")), "'n_particles' must be integer."),
    (dict(n_particles=-200), "'n_particles' must be an integer > 1."),
    (dict(n_particles=[200, 200]), "'n_particles' must be an integer > 1."),
    (dict(adaptmix=-1), "'adaptmix' must be a value > 0 and < 1."),
    (dict(adaptmix=[0.5, 0.5]), "'adaptmix' must be a value > 0 and < 1."),
    (dict(adaptive=-1), "'adaptive' must be an integer >= 0."),
    (dict(n_iterations=0, theta={"beta": 0.16, "gamma": 0.077}),
     "'n_iterations' must be an integer > 0."),
    (dict(n_iterations=[200, 200], theta={"beta": 0.16, "gamma": 0.077}),
     "'n_iterations' must be an integer > 0."),
    (dict(theta={"beta": "A", "gamma": 0.077}),
     "'theta' must be a vector with initial values for the parameters."),
    (dict(theta={"gamma": 0.077}),
     "'theta' must be a vector with initial values for the parameters."),
])
def test_reference_argument_errors(over, msg):
    """tests/pmcmc.R:26-166."""
    with pytest.raises(P.PMCMCError) as e:
        _call(**over)
    assert str(e.value) == msg


def test_prior_densities_closed_forms():
    """dpriors of each family against its written-out log-density, with
    the R-level parameters: gamma(shape p1, rate 1 / p2)."""
    x = 0.7
    cases = [
        ("a ~ gamma(2, 3)", (2 - 1) * math.log(x) - x / 3 - math.lgamma(2) - 2 * math.log(3)),
        ("a ~ lognormal(0.5, 2)",
         -math.log(x * 2 * math.sqrt(2 * math.pi)) - (math.log(x) - 0.5) ** 2 / (2 * 4)),
        ("a ~ normal(1, 0.5)", -math.log(0.5 * math.sqrt(2 * math.pi)) - (x - 1) ** 2 / 0.5),
        ("a ~ uniform(0.5, 2.5)", -math.log(2.0)),
    ]
    for text, want in cases:
        assert abs(P.dpriors([x], parse_priors([text])) - want) < 1e-12, text
    assert P.dpriors([3.0], parse_priors(["a ~ uniform(0.5, 2.5)"])) == -math.inf
    both = parse_priors(["a ~ normal(1, 0.5)", "b ~ uniform(0.5, 2.5)"])
    assert abs(P.dpriors([x, x], both) - (cases[2][1] + cases[3][1])) < 1e-12


def test_prior_draws_moments():
    rng = np.random.default_rng(3)
    pri = parse_priors(["a ~ gamma(2, 3)", "b ~ uniform(1, 2)", "c ~ normal(4, 0.5)",
                        "d ~ lognormal(0, 0.25)"])
    x = np.array([P.rpriors(pri, rng) for _ in range(40000)])
    np.testing.assert_allclose(x.mean(axis=0), [6.0, 1.5, 4.0, math.exp(0.25 ** 2 / 2)],
                               rtol=0.02)


def test_covariance_root():
    rng = np.random.default_rng(1)
    A = rng.normal(size=(3, 3))
    for c in (A @ A.T, np.diag([0.2, 0.0, 1.0]), np.outer([1.0, 2.0, 3.0], [1.0, 2.0, 3.0])):
        Q = P.cov_root(c)
        np.testing.assert_allclose(Q.T @ Q, c, atol=1e-12)


def _gaussian_loglik(y1, y2):
    """A stand-in for the filter: observations y1 of a normal with mean
    gdata[0] (beta) and y2 of a normal with mean gdata[1] (gamma), both of
    sd 1, so that under flat priors the two parameters have independent
    normal posteriors."""
    def f(m, seed):
        return float(-0.5 * np.sum((y1 - m.gdata[0]) ** 2) - 0.5 * np.sum((y2 - m.gdata[1]) ** 2))
    return f


def _mparse_model():
    return si.SIR_mparse(u0={"S": [99], "I": [1], "R": [0]}, tspan=np.arange(1, 11.0),
                         beta=0.2, gamma=0.1)


def test_chain_targets_the_posterior():
    """Flat priors and independent normal observations of beta (50) and
    gamma (200): the chain's draws have each parameter's posterior mean
    (the sample mean) and sd (1 / sqrt(n)); the proposals adapt after
    100 iterations."""
    rng = np.random.default_rng(5)
    y1, y2 = rng.normal(0.3, 1.0, 50), rng.normal(0.1, 1.0, 200)
    res = P.pmcmc(_mparse_model(), 1, TIMES, IOBS.reshape(-1, 1),
                  ["beta ~ uniform(-5, 5)", "gamma ~ uniform(-5, 5)"], 100, 40000,
                  theta={"beta": 1.0, "gamma": 0.5}, seed=11, loglik_fn=_gaussian_loglik(y1, y2))
    assert res.chain.shape == (40000, 6) and res.columns[4:] == ["beta", "gamma"]
    x = res.chain[2000:, 4:]
    sd = np.array([1 / math.sqrt(50), 1 / math.sqrt(200)])
    np.testing.assert_allclose(x.mean(axis=0), [y1.mean(), y2.mean()], atol=0.1 * sd.max())
    np.testing.assert_allclose(x.std(axis=0) / sd, 1.0, atol=0.06)
    assert abs(np.corrcoef(x.T)[0, 1]) < 0.05
    assert 0.1 < res.acceptance_ratio() < 0.8
    # logPost = logLik + logPrior on every row; accepted rows change theta.
    np.testing.assert_allclose(res.chain[:, 0], res.chain[:, 1] + res.chain[:, 2])
    moved = np.any(np.diff(res.chain[:, 4:], axis=0) != 0, axis=1)
    np.testing.assert_array_equal(moved, res.chain[1:, 3] == 1)


def test_one_parameter_posterior():
    """Priors on gamma alone (beta stays the model's): gamma's posterior."""
    rng = np.random.default_rng(6)
    y2 = rng.normal(-0.2, 1.0, 80)
    res = P.pmcmc(_mparse_model(), 1, TIMES, IOBS.reshape(-1, 1), ["gamma ~ uniform(-5, 5)"],
                  100, 30000, theta={"gamma": 0.3}, seed=12,
                  loglik_fn=_gaussian_loglik(np.zeros(0), y2))
    g = res.chain[2000:, 4]
    sd = 1 / math.sqrt(80)
    assert abs(g.mean() - y2.mean()) < 0.1 * sd
    assert abs(g.std() / sd - 1.0) < 0.06


def test_chains_alone_and_together():
    """A chain's draws depend on its seed only: chain 1 of a two-chain
    run equals chain 1 run alone."""
    y = np.random.default_rng(2).normal(0.5, 1.0, 30)
    kw = dict(model=_mparse_model(), obs_compartment=1, times=TIMES,
              obs=IOBS.reshape(-1, 1), priors=["beta ~ uniform(0, 1)", "gamma ~ uniform(0, 1)"],
              n_particles=100, n_iterations=300, seed=4, loglik_fn=_gaussian_loglik(y, y))
    both = P.pmcmc(chains=2, **kw)
    one = P.pmcmc(chains=1, **kw)
    np.testing.assert_array_equal(both[0].chain, one.chain)
    assert not np.array_equal(both[0].chain, both[1].chain)


def test_proposals_outside_the_prior_skip_the_filter():
    """A proposal with log-prior -Inf is rejected without a filter run."""
    calls = []

    def f(m, seed):
        calls.append(m.gdata[0])
        return 0.0
    res = P.pmcmc(_mparse_model(), 1, TIMES, IOBS.reshape(-1, 1),
                  ["beta ~ uniform(0, 0.01)", "gamma ~ uniform(0, 1)"], 100, 200,
                  theta={"beta": 0.005, "gamma": 0.5}, covmat=np.diag([1.0, 1e-6]), seed=3,
                  loglik_fn=f)
    assert all(0 <= b <= 0.01 for b in calls)
    assert len(calls) < 200
    assert np.all((res.chain[:, 4] >= 0) & (res.chain[:, 4] <= 0.01))
