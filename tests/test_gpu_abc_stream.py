"""GPU tests of the mixed ABC-SMC stream (SIR, SEIR and SISe3 problems,
each generation's proposals batched across the model types).

* Each proposal's simulation is the same alone and in the mixed batch:
  a chunk of every problem run together in one batched launch gives
  every plan's trajectory bit for bit as the plan run by itself, and as
  the checker given the same per-replicate parameters.
* A short stream runs end to end: n_particles accepted per generation,
  normalised weights, tolerances that do not grow, and particles inside
  the priors.
"""
import numpy as np
import pytest

from siminf_amd import abc_smc as A
from siminf_amd import abc_stream as S
from siminf_amd.plan import Batch, Plan

pytestmark = pytest.mark.gpu


def test_chunk_same_alone_and_in_the_mix(oracle):
    problems = S.config_problems(scale=0.05)
    rng = np.random.default_rng(4)
    chunk = 40
    data = []
    for pr in problems:
        xx, _ = A.proposals(pr.priors, chunk, None, None, None, rng)
        data.append(pr.replicate_data(xx))
    seeds = [11, 12, 13]
    alone = []
    for pr, (ld, gd), sd in zip(problems, data, seeds):
        print("case: %s alone, %d proposals" % (pr.name, chunk), flush=True)
        with Plan(pr.replicated(chunk)) as p:
            p.set_replicate_data(ldata=ld, gdata=gd)
            p.run(sd)
            alone.append(p.download().U.copy())
    plans = [Plan(pr.replicated(chunk)) for pr in problems]
    try:
        for p, (ld, gd) in zip(plans, data):
            p.set_replicate_data(ldata=ld, gdata=gd)
        for graph in (True, False):
            print("case: mixed batch, graph=%s" % graph, flush=True)
            with Batch(plans) as b:
                b.set_graph(graph)
                b.run(seeds)
                for p, want in zip(plans, alone):
                    np.testing.assert_array_equal(p.download().U, want)
    finally:
        for p in plans:
            p.close()
    for pr, (ld, gd), sd, want in zip(problems, data, seeds, alone):
        ref = oracle.run(pr.replicated(chunk), seed=sd, rep_ldata=ld, rep_gdata=gd)
        assert ref.err == 0
        np.testing.assert_array_equal(want, ref.U)


def test_short_stream():
    problems = S.config_problems(scale=0.05)
    res = S.run_stream(problems, generations=3, proposals=400, n_particles=40, chunk=200,
                       seed=6, autotune=False)
    assert res.proposals == 3 * 400 * 3
    assert len(res.latencies) == res.proposals and min(res.latencies) > 0
    for pr in problems:
        eps = [g[pr.name][0] for g in res.generations]
        assert eps[-1] <= eps[0] + 1e-12, (pr.name, eps)
        for eps_g, x, w in (g[pr.name] for g in res.generations):
            assert x.shape == (40, len(pr.priors))
            assert abs(w.sum() - 1.0) < 1e-9 and (w >= 0).all()
            for k, p in enumerate(pr.priors):
                assert (x[:, k] >= p.p1).all() and (x[:, k] <= p.p2).all()
