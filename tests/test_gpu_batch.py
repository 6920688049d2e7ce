"""GPU tests of batched launches: several plans of different models,
solvers and sizes captured in one graph and launched once.  Every plan's
trajectory and transition count must equal its own run with the same
seed bit for bit (and the checker's where it is run), on every replay of
the graph with new seeds."""
import numpy as np
import pytest

import siminf_amd as si
from siminf_amd import workloads as W

pytestmark = pytest.mark.gpu


def _copy_u(plan):
    m = plan.download()
    return None if m.U is None else m.U.copy(), None if m.V is None else m.V.copy()


def _separate(models, seeds):
    out = []
    for (model, solver), seed in zip(models, seeds):
        with si.Plan(model, solver=solver) as p:
            tr, _, _ = p.run(seed)
            out.append(_copy_u(p) + (tr,))
    return out


def _batched(models, seed_sets):
    plans = [si.Plan(m, solver=s) for m, s in models]
    try:
        with si.Batch(plans) as b:
            results = []
            for seeds in seed_sets:
                print("batch launch seeds=%s" % (list(seeds),), flush=True)
                ms = b.run(seeds)
                assert ms > 0
                res = []
                for p in plans:
                    tr, _, _ = _stats(p)
                    res.append(_copy_u(p) + (tr,))
                results.append(res)
            return results
    finally:
        for p in plans:
            p.close()


def _stats(p):
    import ctypes
    tr, nd, ms = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_double()
    p.lib.SimInf_hip_plan_stats(p._handle, ctypes.byref(tr), ctypes.byref(nd),
                                ctypes.byref(ms))
    return tr.value, nd.value, ms.value


def _same(a, b):
    for x, y in zip(a[:2], b[:2]):
        if x is None:
            assert y is None
        else:
            np.testing.assert_array_equal(x, y)
    assert a[2] == b[2]


def test_seir_ssm_and_aem_in_one_launch(oracle):
    """SEIR with sizes ~ Poisson(500) and 0.5 % of nodes seeded, under ssm
    and aem in one launch (the frozen-node fill and active passes of both
    run side by side)."""
    m = W.seir_million(Nn=100_000, days=120, seed=3, seeded_frac=0.005)
    models = [(m, "ssm"), (m, "aem")]
    seed_sets = [(5, 6), (7, 8)]
    got = _batched(models, seed_sets)
    for seeds, res in zip(seed_sets, got):
        for a, b in zip(_separate(models, seeds), res):
            _same(a, b)
    # And the checker, for the first launch.
    small = W.seir_million(Nn=20_000, days=60, seed=4, seeded_frac=0.01)
    res = _batched([(small, "ssm"), (small, "aem")], [(11, 12)])[0]
    for (solver, seed), r in zip((("ssm", 11), ("aem", 12)), res):
        ref = oracle.run(small, solver=solver, seed=seed)
        assert ref.err == 0
        np.testing.assert_array_equal(r[0], ref.U)
        assert r[2] == ref.transitions


def test_mixed_models_and_sizes_in_one_launch(oracle):
    """Models of different types, node counts and replicate counts in one
    launch: the parsed SIR, SEIR, SISe3 with scheduled E1 events, SIR
    replicates (mssm) and SISe3 replicates with external transfers (one
    workgroup per replicate)."""
    models = [
        (W.mparse_sir_readme(Nn=1000, days=150), "ssm"),
        (W.seir_million(Nn=20_000, days=90, seed=2, seeded_frac=0.02), "aem"),
        (W.sise3_herds(Nn=300, days=200, seed=3, external_rate=0.0), "ssm"),
        (W.mparse_sir_readme(Nn=30, days=60, replicates=50), "ssm"),
        (W.sise3_herds(Nn=120, days=100, seed=8, replicates=6), "ssm"),
    ]
    seed_sets = [(1, 2, 3, 4, 5), (9, 8, 7, 6, 5)]
    got = _batched(models, seed_sets)
    for seeds, res in zip(seed_sets, got):
        for a, b in zip(_separate(models, seeds), res):
            _same(a, b)
    for (model, solver), seed, r in zip(models, seed_sets[1], got[1]):
        ref = oracle.run(model, solver=solver, seed=seed)
        assert ref.err == 0
        np.testing.assert_array_equal(r[0], ref.U)
        if r[1] is not None and r[1].size:
            np.testing.assert_array_equal(r[1], ref.V)


def test_grid_coupled_plans_are_refused():
    """A one-replicate plan with external transfers runs a cooperative
    grid and cannot join a batch; the same plan twice is refused too."""
    m = W.sise3_herds(Nn=300, days=20, seed=5)
    with si.Plan(m) as p:
        with pytest.raises(si.SimInfError, match="Invalid model."):
            si.Batch([p])
    m2 = W.mparse_sir_readme(Nn=10, days=5)
    with si.Plan(m2) as q:
        with pytest.raises(si.SimInfError, match="Invalid model."):
            si.Batch([q, q])


def test_batch_autotune_keeps_results(oracle):
    """The batch autotuner changes launch shapes only: after it, every
    plan's trajectory equals its own run with the same seed, under the
    kept launch (graph or direct enqueue)."""
    models = [(W.seir_million(Nn=50_000, days=80, seed=5, seeded_frac=0.01), "ssm"),
              (W.seir_million(Nn=50_000, days=80, seed=5, seeded_frac=0.01), "aem"),
              (W.mparse_sir_readme(Nn=500, days=60), "ssm")]
    plans = [si.Plan(m, solver=s) for m, s in models]
    try:
        with si.Batch(plans) as b:
            ms, shifts, graph = b.autotune(seed=40, reps=1)
            print("autotuned", ms, shifts, graph, flush=True)
            assert ms > 0 and len(shifts) == 3
            assert all(-1 <= x <= 6 for x in shifts)
            assert all(p.fill_block() in (4, 8, 16, 32, 64) for p in plans)
            for use_graph in (graph, not graph):
                b.set_graph(use_graph)
                b.run([21, 22, 23])
                got = [_copy_u(p) + (_stats(p)[0],) for p in plans]
                for a, c in zip(_separate(models, (21, 22, 23)), got):
                    _same(a, c)
    finally:
        for p in plans:
            p.close()
