"""GPU tests of the benchmark configurations at their full size.

* SEIR on 1,000,000 nodes (sizes ~ Poisson(500), 0.5 % seeded with 5
  infected) over 365 days, under ssm and under aem: U and the transition
  count equal the checker's bit for bit (the checker runs its node loops
  on the host's cores; every node keeps its own stream, so the thread
  count does not change its results), and every node keeps its
  population.
* A particle-filter request: SIR on 10,000 nodes, 2000 particles, 30
  weekly observations of the infected total with a binomial observation
  process (p = 0.5): finite log-likelihood, effective sample sizes in
  [1, 2000], ancestors sorted (systematic resampling), the same result
  twice, alone and beside another request; and bit for bit the checker's
  filter for 100 particles of the same request.
"""
import os

import numpy as np
import pytest

import siminf_amd as si
from siminf_amd import particle_filter, serving
from siminf_amd import workloads as W
from tests.test_gpu_parity import _gpu

pytestmark = pytest.mark.gpu


def _threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


@pytest.mark.parametrize("solver", ["ssm", "aem"])
def test_seir_million_nodes_full_year(oracle, solver):
    m = W.seir_million()
    print("case: SEIR 1e6 nodes x 365 days %s seed 77" % solver, flush=True)
    err, b, transitions = _gpu(m, solver=solver, seed=77)
    assert err == 0
    U = b.U
    del b
    ref = oracle.run(m, solver=solver, seed=77, threads=_threads())
    assert ref.err == 0
    assert transitions == ref.transitions > 5_000_000
    assert np.array_equal(U, ref.U)
    del ref
    # Population per node in every column.
    tot = U.reshape(m.Nn, 4, -1).sum(axis=1, dtype=np.int64)
    assert (tot == m.u0.sum(axis=0)[:, None]).all()


def test_pfilter_request_binomial_2000_particles():
    spec = W.pfilter_binomial_request(Nn=10_000, n_particles=2000, seed=23)
    kw = dict(obs_prob=0.5, seed=spec.seed, obs_dist="binomial")
    print("case: pfilter 10000 nodes x 2000 particles x 30 intervals", flush=True)
    r = si.pfilter(spec.model, 1, spec.times, spec.obs, 2000, **kw)
    assert np.isfinite(r.loglik)
    assert r.ess.shape == (30,) and (r.ess >= 1.0 - 1e-9).all() and (r.ess <= 2000 + 1e-6).all()
    assert r.ancestors.shape == (30, 2000)
    assert (np.diff(r.ancestors, axis=1) >= 0).all()
    assert ((r.ancestors >= 0) & (r.ancestors < 2000)).all()
    again = si.pfilter(spec.model, 1, spec.times, spec.obs, 2000, **kw)
    assert again.loglik == r.loglik
    np.testing.assert_array_equal(again.ancestors, r.ancestors)
    # Alone equals in the mix (beside a smaller request of another seed).
    other = W.pfilter_binomial_request(Nn=3000, n_particles=2000, seed=24)
    with serving.PFilterServer(hbm_budget=64 << 30, max_inflight=2) as srv:
        q1 = srv.submit(serving.Request(spec.model, 1, spec.times, spec.obs, 2000,
                                        obs_prob=0.5, seed=spec.seed))
        q2 = srv.submit(serving.Request(other.model, 1, other.times, other.obs, 2000,
                                        obs_prob=0.5, seed=other.seed))
        mixed, _ = q1.wait(300), q2.wait(300)
    assert mixed.loglik == r.loglik
    np.testing.assert_array_equal(mixed.ess, r.ess)
    np.testing.assert_array_equal(mixed.ancestors, r.ancestors)


def test_pfilter_request_binomial_against_checker(oracle):
    """The same request's structure with 100 particles, against the
    checker's filter bit for bit."""
    spec = W.pfilter_binomial_request(Nn=10_000, n_particles=100, seed=23)
    kw = dict(obs_prob=0.5, seed=spec.seed, obs_dist="binomial")
    print("case: pfilter 10000 nodes x 100 particles vs checker", flush=True)
    err, ll, ess, anc = oracle.pfilter(spec.model, 1, spec.times, spec.obs, 100, **kw)
    assert err == 0
    r = particle_filter.pfilter(spec.model, 1, spec.times, spec.obs, 100, **kw)
    assert r.loglik == ll
    np.testing.assert_array_equal(r.ess, ess)
    np.testing.assert_array_equal(r.ancestors, anc)
