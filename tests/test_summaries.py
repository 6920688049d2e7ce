"""Summaries of a trajectory, on the CPU.

The checker (oracle/prevalence.py) is pinned by the reference's own
prevalence fixtures: the hand-made U with conditions of
tests/prevalence.R:31-95 and levels 1-3 of the SEIR enter-events
trajectory in tests/SEIR.R:327-410 (an RNG-free U, itself a fixture).
The formula checks reproduce the reference's messages
(tests/prevalence.R:33-81, tests/SimInf_model.R:708-730), and the
condition compiler is checked by running its postfix program against the
same condition written directly in numpy."""
import numpy as np
import pytest

from oracle import prevalence as P
from siminf_amd import summaries as S
from tests.golden_data import load, seir_enter_expected

HAND_CONDITIONS = {
    "I ~ .": None,
    "I ~ . | R == 0": lambda e: e["R"] == 0,
    "I ~ . | R > 0": lambda e: e["R"] > 0,
    "I ~ . | (I + R) < 4": lambda e: (e["I"] + e["R"]) < 4,
    "I ~ . | S == 0 | R == 0": lambda e: (e["S"] == 0) | (e["R"] == 0),
    "I ~ . | R == 5": lambda e: e["R"] == 5,
}


def _hand_U():
    d = load("prevalence.json")["hand"]
    return np.array(d["U"], dtype=np.int64).reshape(d["nrow"], d["ncol"]), d


def _expected(vals):
    return np.array([np.nan if v is None else v for v in vals])


def test_checker_against_prevalence_fixtures():
    U, d = _hand_U()
    comp = d["compartments"]
    for case in d["cases"]:
        lhs, rhs, _ = S.parse_formula(case["formula"], comp)
        index = None if case["index"] is None else [i - 1 for i in case["index"]]
        p = P.prevalence(U, comp, lhs, rhs, 1, index, HAND_CONDITIONS[case["formula"]])
        exp = _expected(case["expected"])
        assert np.array_equal(np.isnan(p), np.isnan(exp)), case
        ok = ~np.isnan(exp)
        assert np.all(np.abs(p[ok] - exp[ok]) < 1e-8), case


def test_checker_against_seir_fixtures():
    d = load("prevalence.json")["seir"]
    U = seir_enter_expected()
    comp = ["S", "E", "I", "R"]
    for rhs in (["S", "E", "I", "R"], comp):
        p1 = P.prevalence(U, comp, ["I"], rhs, 1)
        assert np.all(np.abs(p1 - d["level1"]) < d["tol"])
        assert np.array_equal(P.prevalence(U, comp, ["I"], rhs, 2), d["level2"])
    p3 = P.prevalence(U, comp, ["I"], comp, 3)           # nodes x time
    l3 = d["level3"]
    got = p3[np.array(l3["node"]) - 1, np.array(l3["time"], dtype=int)]
    assert np.all(np.abs(got - l3["prevalence"]) < d["tol"])
    sub = d["level3_index_2_3"]
    p3s = P.prevalence(U, comp, ["I"], comp, 3, index=[1, 2])
    rows = {2: 0, 3: 1}
    got = p3s[[rows[n] for n in sub["node"]], np.array(sub["time"], dtype=int)]
    assert np.all(np.abs(got - sub["prevalence"]) < d["tol"])


def test_trajectory_checker_against_seir_fixture():
    """tests/SEIR.R:318-325: E and R of nodes 2 and 4, format matrix."""
    U = seir_enter_expected()
    got = P.trajectory(U, ["S", "E", "I", "R"], ["E", "R"], index=[1, 3])
    exp = np.array([220, 240, 420, 440, 221, 241, 421, 441, 222, 242, 422, 442, 223, 243,
                    423, 443, 224, 244, 424, 444, 225, 245, 425, 445, 226, 246, 426, 446,
                    227, 247, 427, 447, 228, 248, 428, 448, 229, 249, 429, 449]
                   ).reshape(10, 4).T
    assert np.array_equal(got, exp)


def test_formula_errors_match_reference():
    comp = ["S", "I", "R"]
    cases = [("I ~ . | R == 0", 4, "'level' must be an integer with a value 1, 2 or 3."),
             ("~I", 1, "Invalid 'formula' specification."),
             ("A + S ~ S", 1, "Non-existing compartment(s) in model: 'A'."),
             ("S ~ A + S", 1, "Non-existing compartment(s) in model: 'A'."),
             ("I ~ . | TRUE == 0", 1, S._CONDITION_ERROR),
             ("I ~ . | R + 1", 1, S._CONDITION_ERROR)]
    for formula, level, msg in cases:
        with pytest.raises(S.SummaryError) as e:
            S.prevalence_spec(formula, comp, level)
        assert str(e.value) == msg, formula
    for bad in (["A", "S"], 3.4, 0, 10):
        with pytest.raises(S.SummaryError, match="node index must be an integer"):
            S.node_index(bad, 6)
    assert S.node_index([3, 1, 3], 6).tolist() == [0, 2]


def test_masks_and_dot():
    comp = ["S", "E", "I", "R"]
    assert S.prevalence_spec("I ~ .", comp, 1)[:2] == (4, 15)
    assert S.prevalence_spec("I + E ~ S + E + I + R", comp, 2)[:2] == (6, 15)
    assert S.prevalence_spec(". ~ I", comp, 3)[:2] == (15, 4)
    assert S.trajectory_mask(["R", "E"], comp) == 10


def _run_program(ops, args, env):
    """Evaluate a postfix program over arrays (the device's semantics:
    doubles, comparisons and logicals give 1.0 / 0.0)."""
    st = []
    names = list(env)
    for op, a in zip(ops, args):
        if op == S.OP_U:
            st.append(env[names[int(a)]].astype(np.float64))
        elif op == S.OP_CONST:
            st.append(np.float64(a))
        elif op == S.OP_NEG:
            st.append(-st.pop())
        elif op == S.OP_NOT:
            st.append((st.pop() == 0).astype(np.float64))
        else:
            b, x = st.pop(), st.pop()
            f = {S.OP_ADD: np.add, S.OP_SUB: np.subtract, S.OP_MUL: np.multiply,
                 S.OP_DIV: np.divide, S.OP_EQ: np.equal, S.OP_NE: np.not_equal,
                 S.OP_LT: np.less, S.OP_LE: np.less_equal, S.OP_GT: np.greater,
                 S.OP_GE: np.greater_equal,
                 S.OP_AND: lambda p, q: (p != 0) & (q != 0),
                 S.OP_OR: lambda p, q: (p != 0) | (q != 0)}[op]
            with np.errstate(divide="ignore", invalid="ignore"):
                st.append(np.asarray(f(x, b), dtype=np.float64))
    assert len(st) == 1
    return st[0] != 0


CONDITIONS = [
    ("S == 0 | R == 0", lambda e: (e["S"] == 0) | (e["R"] == 0)),
    ("!R == 0", lambda e: ~(e["R"] == 0)),
    ("-I < 2 & R >= 1 | S != 3", lambda e: ((-e["I"] < 2) & (e["R"] >= 1)) | (e["S"] != 3)),
    ("I + R * 2 > S / 3", lambda e: e["I"] + e["R"] * 2 > e["S"] / 3),
    ("(S > 1) & !(I == 0)", lambda e: (e["S"] > 1) & ~(e["I"] == 0)),
    ("S - I - R <= 1", lambda e: (e["S"] - e["I"]) - e["R"] <= 1),
    ("S / I / 2 >= 1 && TRUE", lambda e: (e["S"] / e["I"]) / 2 >= 1),
    ("!(S < 2 || I < 2) & R == 1e0", lambda e: ~((e["S"] < 2) | (e["I"] < 2)) & (e["R"] == 1)),
    ("I * -1 < -2", lambda e: e["I"] * -1 < -2),
]


def test_condition_compiler_precedence():
    rng = np.random.default_rng(1)
    env = {c: rng.integers(0, 5, (40, 7)) for c in ("S", "I", "R")}
    for text, ref in CONDITIONS:
        ops, args = S.compile_condition(text, ["S", "I", "R"])
        with np.errstate(divide="ignore", invalid="ignore"):
            expected = np.asarray(ref(env), dtype=bool)
        assert np.array_equal(_run_program(ops, args, env), expected), text
        assert S.stack_depth(ops) <= S.COND_STACK


def test_condition_errors():
    for text in ("S ==", "S == 1 == 2", "(S > 1", "S > 1)", "S % 2 == 0", "X > 1"):
        with pytest.raises(S.SummaryError):
            S.compile_condition(text, ["S", "I", "R"])
    deep = "(" * 20 + "S" + " + 1)" * 20 + " > 0"
    with pytest.raises(S.SummaryError, match="too long"):
        S.compile_condition(" | ".join(["S > %d" % k for k in range(30)]), ["S"])
    S.compile_condition(deep, ["S"])


def test_checker_node_prevalence_by_hand():
    """Level 2 on the hand-made U: only node 2 ever has a case and all
    three nodes hold individuals, so 1/3 per column; with R == 0 node 3
    drops out (1/2 while node 2 has no R, then 0/1)."""
    U, d = _hand_U()
    comp = d["compartments"]
    assert np.allclose(P.prevalence(U, comp, ["I"], comp, 2), 1 / 3)
    p = P.prevalence(U, comp, ["I"], comp, 2, condition=lambda e: e["R"] == 0)
    assert np.allclose(p, [0.5, 0.5, 0.5, 0.0, 0.0])
