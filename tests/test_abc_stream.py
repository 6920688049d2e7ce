"""CPU tests of the mixed ABC-SMC stream's host logic: each proposal's
parameters land in its replicate's ldata or gdata, distances are per
replicate, and the replicated model repeats the first replicate's state
and events."""
import numpy as np

from siminf_amd import abc_stream as S


def test_replicate_data_layout():
    sir, seir, sise3 = S.config_problems(scale=0.01)
    xx = np.array([[0.1, 0.05], [0.2, 0.06], [0.3, 0.07]])
    ld, gd = sir.replicate_data(xx)
    assert ld is None and gd.shape == (2, 3)
    np.testing.assert_array_equal(gd.T, xx)
    x3 = np.array([[0.1, 0.2, 0.05], [0.15, 0.3, 0.06]])
    ld, gd = seir.replicate_data(x3)
    # The nodes share their ldata, so one column per replicate.
    assert gd is None and ld.shape == (3, 2)
    np.testing.assert_array_equal(ld.T, x3)
    seir.model.ldata[0, 1] = 0.5
    ld, _ = seir.replicate_data(x3)
    Nn = seir.model.Nn
    assert ld.shape == (3, Nn * 2)
    np.testing.assert_array_equal(ld[:, :Nn], np.tile(x3[0][:, None], (1, Nn)))
    np.testing.assert_array_equal(ld[:, Nn:], np.tile(x3[1][:, None], (1, Nn)))
    ld, gd = sise3.replicate_data(np.array([[0.01, 0.09], [0.02, 0.11]]))
    assert ld is None and gd.shape == (12, 2)
    assert gd[0, 1] == 0.02 and gd[3, 1] == 0.11 and gd[1, 0] == sise3.model.gdata[1]


def test_replicated_model_and_distances():
    _, _, sise3 = S.config_problems(scale=0.02)
    m = sise3.replicated(3)
    assert m.replicates == 3 and m.Nn == sise3.model.Nn
    np.testing.assert_array_equal(m.u0[:, m.Nn:2 * m.Nn], sise3.model.u0)
    assert len(m.events["event"]) == len(sise3.model.events["event"])
    prev = np.array([0.1, 0.2, 0.3, 0.1, 0.2, 0.5])
    d = S.distances(prev, [0.1, 0.2, 0.3], 2, 3)
    np.testing.assert_allclose(d, [0.0, np.sqrt(0.04 / 3)])
