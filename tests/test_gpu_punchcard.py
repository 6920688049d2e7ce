"""Sparse output through punchcard templates on the GPU: the values at the
template's positions equal the dense run's (same seed) and the checker's,
and the RNG-free SISe cases of tests/punchcard.R:267-300 give their
known values."""
import numpy as np
import pytest

import siminf_amd as si
from siminf_amd.punchcard import punchcard
from tests.test_punchcard import _sise, sp_values

pytestmark = pytest.mark.gpu


def test_sparse_equals_dense_and_checker(oracle):
    rng = np.random.default_rng(5)
    m = si.SIR(u0={"S": rng.integers(50, 150, 300), "I": rng.integers(0, 6, 300), "R": 0},
               tspan=np.arange(1, 61.0), beta=0.3, gamma=0.1, replicates=2)
    print("case: SIR 300 nodes x 2 replicates, punchcard", flush=True)
    si.run(m, seed=4)
    U = m.U.copy()
    value = {"node": rng.integers(1, 301, 500), "time": rng.integers(1, 61, 500),
             "S": rng.random(500) < 0.4, "I": rng.random(500) < 0.8}
    punchcard(m, value)
    si.run(m, seed=4)
    c = m.U_sparse.tocoo()
    got = c.data
    np.testing.assert_array_equal(got, U[c.row, c.col].astype(float))
    ref = oracle.run(m, seed=4)
    assert ref.err == 0
    np.testing.assert_array_equal(got, sp_values(m.U_sparse, ref.prU))


def test_known_values_sise():
    """tests/punchcard.R:267-300: zero rates, so S, I and phi keep their
    initial values at every recorded point."""
    m = _sise([100, 100], [0, 0], [1.0, 2.0])
    print("case: SISe punchcard known values", flush=True)
    punchcard(m, {"node": [1, 1], "time": [4, 6]})
    si.run(m, seed=1)
    assert m.U_sparse.tocoo().data.tolist() == [100.0, 0.0, 100.0, 0.0]
    assert m.V_sparse.tocoo().data.tolist() == [1.0, 1.0]
    punchcard(m, {"node": [1, 1, 2, 2], "time": [2, 4, 6, 8], "S": [True, True, False, False],
                  "I": [True, True, False, False], "phi": [False, False, True, True]})
    si.run(m, seed=1)
    assert m.U_sparse.tocoo().data.tolist() == [100.0, 0.0, 100.0, 0.0]
    assert m.V_sparse.tocoo().data.tolist() == [2.0, 2.0]
