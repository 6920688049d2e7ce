"""On-device summaries of a resident trajectory: prevalence at levels 1-3
with node indices and conditions, and trajectory extraction, equal bit
for bit to the checker (oracle/prevalence.py) on the downloaded U, and
to the reference's own prevalence fixtures."""
import numpy as np
import pytest

import siminf_amd as si
from oracle import prevalence as P
from siminf_amd import summaries as S
from tests.golden_data import load, seir_enter_model
from tests.test_summaries import CONDITIONS

pytestmark = pytest.mark.gpu


def _case(msg):
    print("case: " + msg, flush=True)


def test_seir_fixture_on_device():
    _case("seir enter events prevalence")
    d = load("prevalence.json")["seir"]
    with si.Plan(seir_enter_model()) as plan:
        with pytest.raises(si.SimInfError, match="Please run the model first"):
            plan.prevalence("I ~ .")
        plan.run(1)
        assert np.all(np.abs(plan.prevalence("I ~ S + E + I + R") - d["level1"]) < d["tol"])
        assert np.array_equal(plan.prevalence("I ~ .", level=2), d["level2"])
        p3 = plan.prevalence("I ~ .", level=3)
        l3 = d["level3"]
        got = p3[np.array(l3["node"]) - 1, np.array(l3["time"], dtype=int)]
        assert np.all(np.abs(got - l3["prevalence"]) < d["tol"])
        p3s = plan.prevalence("I ~ .", level=3, index=[2, 3])
        sub = d["level3_index_2_3"]
        got = p3s[np.array(sub["node"]) - 2, np.array(sub["time"], dtype=int)]
        assert np.all(np.abs(got - sub["prevalence"]) < d["tol"])
        t = plan.trajectory(["E", "R"], index=[2, 4])
        assert t[:, 0].tolist() == [220, 240, 420, 440] and t[:, 9].tolist() == [229, 249, 429, 449]


def _compare(plan, model, formula, level, index, cond):
    U = model.U
    lhs, rhs, _ = S.parse_formula(formula, model.compartments)
    zidx = None if index is None else (np.unique(index) - 1)
    want = P.prevalence(U, model.compartments, lhs, rhs, level, zidx, cond)
    got = plan.prevalence(formula, level=level, index=index)
    np.testing.assert_array_equal(got, want)


def test_sise3_with_events_matches_checker():
    from siminf_amd import workloads as W
    m = W.sise3_herds(Nn=300, days=120, seed=4)
    _case("sise3 300 herds x 120 days")
    with si.Plan(m) as plan:
        plan.run(7)
        plan.download()
        cases = "I_1 + I_2 + I_3 ~ ."
        cond = ("S_1 > 10 & I_3 == 0",
                lambda e: (e["S_1"] > 10) & (e["I_3"] == 0))
        for level in (1, 2, 3):
            for index in (None, [5, 17, 300, 2]):
                _compare(plan, m, cases, level, index, None)
                _compare(plan, m, cases + " | " + cond[0], level, index, cond[1])
        got = plan.trajectory(["I_1", "S_3"], index=[3, 1])
        np.testing.assert_array_equal(got, P.trajectory(m.U, m.compartments,
                                                        ["I_1", "S_3"], [0, 2]))
        np.testing.assert_array_equal(plan.trajectory(), m.U)


def test_conditions_and_replicates():
    m = si.SIR(u0={"S": np.arange(20, 60), "I": np.arange(40) % 4, "R": 0},
               tspan=np.arange(1, 31.0), beta=0.3, gamma=0.1, replicates=3)
    _case("SIR 40 nodes x 3 replicates, conditions")
    with si.Plan(m) as plan:
        plan.run(11)
        plan.download()
        assert plan.prevalence("I ~ .").shape == (90,)
        for text, ref in CONDITIONS:
            for level in (1, 2, 3):
                _compare(plan, m, "I ~ S + I | " + text, level, None, ref)
        with pytest.raises(S.SummaryError):
            plan.prevalence("I ~ .", level=4)


def test_large_population_level1():
    from siminf_amd import workloads as W
    m = W.seir_million(Nn=200_000, days=60)
    _case("SEIR 2e5 nodes x 60 days")
    with si.Plan(m) as plan:
        plan.run(3)
        p = plan.prevalence("I ~ .")
        p2 = plan.prevalence("E + I ~ .", level=2)
        plan.download()
        np.testing.assert_array_equal(p, P.prevalence(m.U, m.compartments, ["I"],
                                                      m.compartments, 1))
        np.testing.assert_array_equal(p2, P.prevalence(m.U, m.compartments, ["E", "I"],
                                                       m.compartments, 2))


def test_trajectory_of_continuous_state():
    """trajectory(model, "phi", index, format = "matrix") gathers rows of
    V (R/trajectory.R:73-85, :227-228); U and V names together raise the
    reference's message."""
    _case("trajectory of V (SISe phi), 2 replicates")
    n = 12
    m = si.SISe(u0={"S": np.arange(50, 50 + n), "I": np.arange(n) % 4},
                tspan=np.arange(1, 41.0), phi=np.linspace(0.0, 1.0, n), upsilon=0.02,
                gamma=0.1, alpha=1.0, beta_t1=0.12, beta_t2=0.1, beta_t3=0.12, beta_t4=0.1,
                end_t1=91, end_t2=182, end_t3=273, end_t4=365, epsilon=0.001,
                replicates=2)
    assert m.Nd == 1 and m.v0_names == ["phi"]
    with si.Plan(m) as plan:
        plan.run(seed=5)
        plan.download()
        V = np.asarray(m.V)
        assert V.shape == (n, 2 * 40) and np.any(V != V[:, :1])
        np.testing.assert_array_equal(plan.trajectory("phi"), V)
        np.testing.assert_array_equal(plan.trajectory(["phi"], index=[7, 2, 11]),
                                      V[[1, 6, 10]])
        with pytest.raises(S.SummaryError, match="Cannot combine data from different slots."):
            plan.trajectory(["I", "phi"])
        with pytest.raises(S.SummaryError, match="Non-existing compartment"):
            plan.trajectory(["phi", "Q"])
