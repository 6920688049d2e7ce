"""Sparse output templates (punchcard<-, R/punchcard.R:128-256) on the
CPU: the RNG-free template patterns and error messages of
tests/punchcard.R, and the checker run with templates equals the dense
run at the template's positions."""
import numpy as np
import pytest

import siminf_amd as si
from siminf_amd.punchcard import PunchcardError, create_template, punchcard


def _sir6():
    return si.SIR(u0={"S": np.arange(100, 106), "I": np.arange(1, 7), "R": 0},
                  tspan=np.arange(1, 11.0), beta=0.16, gamma=0.077)


def _sise(S, I, phi, tspan=np.arange(1, 11.0)):
    return si.SISe(u0={"S": S, "I": I}, tspan=tspan, phi=phi, upsilon=0, gamma=0,
                   alpha=0, epsilon=0, beta_t1=0, beta_t2=0, beta_t3=0, beta_t4=0,
                   end_t1=91, end_t2=182, end_t3=273, end_t4=365)


def _pattern(m):
    c = m.tocoo()
    return sorted(zip((c.row + 1).tolist(), (c.col + 1).tolist()))


def test_errors_match_reference():
    """tests/punchcard.R:41-48 and :232-235."""
    m = _sir6()
    for value, msg in ((5, "'value' argument is not a 'data.frame'."),
                       ({"node": [10], "time": [3]}, "Unable to match all nodes."),
                       ({"node": [3], "time": [11]}, "Unable to match all time-points to tspan."),
                       ({"a": [3], "b": [11]}, "'value' must have the columns 'time' and 'node'.")):
        with pytest.raises(PunchcardError) as e:
            punchcard(m, value)
        assert str(e.value) == msg


def test_replicates_pattern():
    """tests/punchcard.R:354-385: 6 columns of u0 as 2 replicates of 3
    nodes; U and V patterns repeat per replicate with a tlen offset."""
    base = _sise(np.arange(101, 107), np.arange(1, 7), [0.1, 0.2, 0.3, 0.4, 0.5, 0.6])
    m = si.SimInfModel("SISe", base.G, base.S, base.tspan, base.u0,
                       ldata=base.ldata[:, :3], gdata=base.gdata, v0=base.v0, E=base.E,
                       replicates=2, compartments=base.compartments,
                       v0_names=base.v0_names)
    assert m.Nn == 3
    punchcard(m, {"node": np.tile([1, 3], 10), "time": np.repeat(np.arange(1, 11), 2),
                  "S": [True, False, False, True], "I": [False, True, True, False],
                  "phi": [True, False, False, True]})
    i = [1, 6, 2, 5] * 10
    j = list(np.repeat(np.arange(1, 21), 2))
    assert m.U_sparse.shape == (6, 20) and _pattern(m.U_sparse) == sorted(zip(i, j))
    i = [1, 3] * 10
    j = list(range(1, 21))
    assert m.V_sparse.shape == (3, 20) and _pattern(m.V_sparse) == sorted(zip(i, j))


def test_dense_when_everything_is_selected_and_empty_templates():
    m = _sir6()
    punchcard(m, {"node": np.repeat(np.arange(1, 7), 10), "time": np.tile(np.arange(1, 11), 6)})
    assert m.U_sparse is None                       # every node, compartment, time
    punchcard(m, {})                                # tests/punchcard.R:209-219
    assert m.U_sparse.shape == (18, 10) and m.U_sparse.nnz == 0
    assert m.V_sparse.shape == (0, 10) and m.V_sparse.nnz == 0
    punchcard(m, None)
    assert m.U_sparse is None and m.V_sparse is None


def test_na_entries_and_slots():
    """tests/punchcard.R:172-193 (NA is FALSE) and :302-325 (U and V
    columns select their own slot; only node/time selects everything)."""
    m = _sir6()
    punchcard(m, {"node": [1, 1, 2, 2, 3, 3, 4, 4], "time": [5, 6, 6, 7, 8, 9, 9, 10],
                  "S": [True, None, True, None, True, None, True, None],
                  "I": [True, None, None, True, True, None, None, True],
                  "R": [None, True, None, True, None, True, None, True]})
    exp = [(1, 5), (2, 5), (3, 6), (4, 6), (5, 7), (6, 7), (7, 8), (8, 8), (9, 9),
           (10, 9), (11, 10), (12, 10)]
    assert _pattern(m.U_sparse) == sorted(exp)
    s = _sise([100, 100], [0, 0], [1.0, 2.0])
    punchcard(s, {"node": [1, 1, 2, 2], "time": [2, 4, 6, 8], "S": [True, True, False, False],
                  "I": [True, True, False, False], "phi": [False, False, True, True]})
    assert _pattern(s.U_sparse) == [(1, 2), (1, 4), (2, 2), (2, 4)]
    assert _pattern(s.V_sparse) == [(2, 6), (2, 8)]
    punchcard(s, {"node": [1, 1], "time": [4, 6]})
    assert _pattern(s.U_sparse) == [(1, 4), (1, 6), (2, 4), (2, 6)]
    assert _pattern(s.V_sparse) == [(1, 4), (1, 6)]


def test_checker_sparse_run_equals_dense_run(oracle):
    m = _sir6()
    dense = oracle.run(m, seed=9)
    assert dense.err == 0
    U = dense.U.copy()
    rng = np.random.default_rng(2)
    value = {"node": rng.integers(1, 7, 25), "time": rng.integers(1, 11, 25),
             "S": rng.random(25) < 0.5, "I": rng.random(25) < 0.7}
    punchcard(m, value)
    sparse = oracle.run(m, seed=9)
    assert sparse.err == 0
    c = m.U_sparse.tocoo()
    got = sp_values(m.U_sparse, sparse.prU)
    assert np.array_equal(got, U[c.row, c.col].astype(float))


def sp_values(pattern, pr):
    """Values of a CSC pattern's entries in COO (row, col) order."""
    m = pattern.tocsc().copy()
    m.data = np.asarray(pr, dtype=float)
    return m.tocoo().data
