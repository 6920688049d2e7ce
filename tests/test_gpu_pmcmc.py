"""GPU tests of PMCMC over the device particle filter (R/pmcmc.R) on the
reference's own example (tests/pmcmc.R:10-24: one SIR node, Poisson
observations of I + 1e-6): the chain equals, row for row, the chain the
same seeds give with the checker's filter as its likelihood; chains run
together through the serving loop equal the same chains run alone."""
import importlib

import numpy as np
import pytest

from tests.test_pmcmc import IOBS, PRIORS, TIMES, reference_model

P = importlib.import_module("siminf_amd.pmcmc")

pytestmark = pytest.mark.gpu

KW = dict(obs_compartment=1, times=TIMES, obs=IOBS.reshape(-1, 1), priors=PRIORS,
          n_particles=200, obs_prob=1.0, obs_dist="poisson", obs_offset=1e-6)


def test_chain_equals_the_checker_chain(oracle):
    def checker(m, seed):
        err, ll, _, _ = oracle.pfilter(m, 1, TIMES, IOBS.reshape(-1, 1), 200, obs_prob=1.0,
                                       seed=seed, obs_dist="poisson", obs_offset=1e-6)
        assert err == 0
        return ll
    print("case: pmcmc SIR 1 node, 200 particles, 60 iterations", flush=True)
    dev = P.pmcmc(reference_model(), n_iterations=60, seed=123,
                  theta={"beta": 0.16, "gamma": 0.077}, **KW)
    ref = P.pmcmc(reference_model(), n_iterations=60, seed=123,
                  theta={"beta": 0.16, "gamma": 0.077}, loglik_fn=checker, **KW)
    np.testing.assert_array_equal(dev.chain, ref.chain)
    assert np.all(np.isfinite(dev.chain[:, 1]))
    assert 0.0 < dev.acceptance_ratio() < 1.0


def test_chains_through_the_server_equal_chains_alone():
    """Four chains side by side (their filter requests overlap on the
    device) give the chains each seed gives alone."""
    print("case: pmcmc 4 chains x 30 iterations through the serving loop", flush=True)
    together = P.pmcmc(reference_model(), n_iterations=30, seed=7, chains=4, **KW)
    assert len(together) == 4
    alone = P.pmcmc(reference_model(), n_iterations=30, seed=7, chains=1, **KW)
    np.testing.assert_array_equal(together[0].chain, alone.chain)
    assert not np.array_equal(together[0].chain, together[1].chain)
    # One filter request per iteration whose proposal is inside the priors.
    assert all(1 <= len(r.latencies) <= 30 for r in together)
