"""Fused level-1 prevalence (Plan.fuse_prevalence): a node-kernel plan
that keeps exact integer sums per replicate and U column instead of U
gives, bit for bit, the prevalence the same plan unfused computes from
its resident U (itself equal to the checker on the oracle's U):

* ssm with replicates, fewer than 64 nodes (each lane adds its frozen
  counts alone) and more (a wavefront's lanes span two replicates);
* nodes frozen from the start (classify pass), nodes that freeze during
  the run, tspan with gaps (several columns in one day), E1 events;
* the all-events method;
* batched launches (graph and streams), a plan fused after its batch was
  captured, repeated runs (the sums start from zero each run);
* the errors of a fused plan, and plans that cannot fuse.
"""
import numpy as np
import pytest

import siminf_amd as si
from oracle import prevalence as P
from siminf_amd import summaries as S
from siminf_amd import workloads as W
from siminf_amd.plan import Batch, Plan
from tests.golden_data import seir_enter_model

pytestmark = pytest.mark.gpu


def _case(msg):
    print("case: " + msg, flush=True)


def _seir(Nn, days=60, seed=3, frac=0.05, replicates=1, tspan=None):
    m = W.seir_million(Nn=Nn, days=days, seed=seed, seeded_frac=frac)
    u = m.u0
    return si.SEIR(u0={"S": u[0], "E": u[1], "I": u[2], "R": u[3]},
                   tspan=m.tspan if tspan is None else tspan, beta=0.16, epsilon=0.25,
                   gamma=0.077, replicates=replicates)


def _unfused(model, formula, seed, solver="ssm"):
    with Plan(model, solver=solver) as p:
        p.run(seed)
        return p.prevalence(formula, level=1), p.download().U.copy()


def _fused(model, formula, seed, solver="ssm", runs=1):
    with Plan(model, solver=solver) as p:
        assert p.fuse_prevalence(formula)
        out = []
        for _ in range(runs):
            p.run(seed)
            out.append(p.prevalence(formula, level=1))
        return out


CASES = [
    ("SIR 40 nodes x 3 replicates",
     lambda: si.SIR(u0={"S": np.arange(20, 60), "I": np.arange(40) % 4 * (np.arange(40) < 9),
                        "R": 0},
                    tspan=np.arange(1, 31.0), beta=0.3, gamma=0.1, replicates=3),
     "I ~ S + I + R"),
    ("SEIR 3000 nodes x 4 replicates, 5 % seeded",
     lambda: _seir(3000, replicates=4), "I ~ S + E + I + R"),
    ("SEIR 1000 nodes x 3 replicates, tspan with gaps",
     lambda: _seir(1000, replicates=3,
                   tspan=np.array([0.5, 1.0, 1.5, 2.0, 7.0, 7.25, 19.0, 40.0, 41.0, 90.0])),
     "E + I ~ S + E + I + R"),
    ("SEIR 777 nodes, every node seeded", lambda: _seir(777, frac=1.0), "I ~ S + E + I + R"),
    ("SEIR with enter events", seir_enter_model, "I ~ S + E + I + R"),
]


@pytest.mark.parametrize("name,make,formula", CASES, ids=[c[0] for c in CASES])
def test_fused_equals_unfused_ssm(name, make, formula):
    _case("fused ssm: " + name)
    m = make()
    want, U = _unfused(m, formula, seed=5)
    lhs, rhs, _ = S.parse_formula(formula, m.compartments)
    np.testing.assert_array_equal(want, P.prevalence(U, m.compartments, lhs, rhs, 1, None,
                                                     None))
    got = _fused(m, formula, seed=5, runs=2)
    for g in got:
        np.testing.assert_array_equal(g, want)


@pytest.mark.parametrize("Nn", [50, 2000])
def test_fused_equals_unfused_aem(Nn):
    _case("fused aem: SEIR %d nodes" % Nn)
    m = _seir(Nn, days=50, frac=0.1)
    want, _ = _unfused(m, "I ~ S + E + I + R", seed=8, solver="aem")
    got = _fused(m, "I ~ S + E + I + R", seed=8, solver="aem")
    np.testing.assert_array_equal(got[0], want)


def test_fused_in_batches():
    ms = [_seir(1500, replicates=2, seed=4), _seir(300, replicates=5, seed=6, frac=0.2)]
    formulas = ["I ~ S + E + I + R", "E ~ S + E + I + R"]
    seeds = [21, 22]
    want = [_unfused(m, f, s)[0] for m, f, s in zip(ms, formulas, seeds)]
    plans = [Plan(m) for m in ms]
    try:
        assert plans[0].fuse_prevalence(formulas[0])
        for graph in (True, False):
            _case("fused plan in a batch, graph=%s" % graph)
            with Batch(plans) as b:
                b.set_graph(graph)
                b.run(seeds)
                for p, f, w in zip(plans, formulas, want):
                    np.testing.assert_array_equal(p.prevalence(f, level=1), w)
                # Fused after the batch was captured: captured again.
                if plans[1].fuse_prevalence(formulas[1]):
                    b.run(seeds)
                    for p, f, w in zip(plans, formulas, want):
                        np.testing.assert_array_equal(p.prevalence(f, level=1), w)
    finally:
        for p in plans:
            p.close()


def test_fused_errors_and_refusals():
    _case("fused plan errors")
    m = _seir(200, days=20)
    with Plan(m) as p:
        assert p.fuse_prevalence("I ~ S + E + I + R")
        with pytest.raises(si.SimInfError):
            p.prevalence("I ~ S + E + I + R")      # not run yet
        p.run(3)
        p.prevalence("I ~ S + E + I + R")
        for bad in (lambda: p.prevalence("I ~ S + E + I + R", level=2),
                    lambda: p.prevalence("E ~ S + E + I + R"),
                    lambda: p.prevalence("I ~ S + E + I + R", index=[1, 2]),
                    lambda: p.prevalence("I ~ S + E + I + R | S > 10"),
                    lambda: p.download(),
                    lambda: p.trajectory(),
                    lambda: p.fuse_prevalence("I ~ S + E + I + R | S > 10")):
            with pytest.raises(si.SimInfError):
                bad()
        # Other masks: the plan must run again.
        assert p.fuse_prevalence("E ~ S + E + I + R")
        with pytest.raises(si.SimInfError):
            p.prevalence("E ~ S + E + I + R")
        p.run(3)
        want, _ = _unfused(m, "E ~ S + E + I + R", seed=3)
        np.testing.assert_array_equal(p.prevalence("E ~ S + E + I + R"), want)
    # A coupled plan (E2 events) keeps its trajectory.
    _case("coupled plan does not fuse")
    mc = W.sise3_herds(Nn=100, days=20, seed=2)
    with Plan(mc) as p:
        assert not p.fuse_prevalence("I_1 ~ .")
        p.run(1)
        p.download()
