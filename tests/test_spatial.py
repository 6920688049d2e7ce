"""CPU tests of the spatial models SISe_sp and SISe3_sp: the distance
matrix and neighbour ldata against the reference's own expectations
(tests/ldata_sp.R, tests/SISe_sp.R), and the checker's local spread
(src/misc/SimInf_local_spread.c) pinned independently of the plugin code
it shares with the device: the linear map it reduces to in matrix form,
a two-herd closed form, conservation of sum(N * phi), equality with the
non-spatial models at coupling 0, and the error codes -7 and -10."""
import numpy as np
import pytest
import scipy.sparse as sp

import siminf_amd as si
from siminf_amd.spatial import SpatialError, ldata_sp
from tests import spatial_cases as C

TOL = 1e-8

# tests/ldata_sp.R:31-45: distance_matrix(1:10, 1:10, 3).
D_I = [1, 2, 0, 2, 3, 0, 1, 3, 4, 1, 2, 4, 5, 2, 3, 5, 6, 3, 4, 6, 7, 4, 5, 7, 8,
       5, 6, 8, 9, 6, 7, 9, 7, 8]
D_P = [0, 2, 5, 9, 13, 17, 21, 25, 29, 32, 34]
R2, R8 = 1.4142135623731, 2.82842712474619
D_X = [R2, R8, R2, R2, R8, R8, R2, R2, R8, R8, R2, R2, R8, R8, R2, R2, R8, R8,
       R2, R2, R8, R8, R2, R2, R8, R8, R2, R2, R8, R8, R2, R2, R8, R2]

# tests/ldata_sp.R:146-159: ldata with metric "degree" (0).
LDATA_DEGREE = np.array([
    91, 182, 273, 365, 1, 3, 2, 4, -1, 0, 0, 0, 0, 0,
    91, 182, 273, 365, 0, 2, 2, 4, 3, 4, -1, 0, 0, 0,
    91, 182, 273, 365, 0, 2, 1, 3, 3, 4, 4, 4, -1, 0,
    91, 182, 273, 365, 1, 3, 2, 4, 4, 4, 5, 4, -1, 0,
    91, 182, 273, 365, 2, 4, 3, 4, 5, 4, 6, 4, -1, 0,
    91, 182, 273, 365, 3, 4, 4, 4, 6, 4, 7, 4, -1, 0,
    91, 182, 273, 365, 4, 4, 5, 4, 7, 4, 8, 3, -1, 0,
    91, 182, 273, 365, 5, 4, 6, 4, 8, 3, 9, 2, -1, 0,
    91, 182, 273, 365, 6, 4, 7, 4, 9, 2, -1, 0, 0, 0,
    91, 182, 273, 365, 7, 4, 8, 3, -1, 0, 0, 0, 0, 0], float).reshape(10, 14).T

# tests/ldata_sp.R:163-195: ldata with metric "distance" (1).
LDATA_DIST = np.array([
    91, 182, 273, 365, 1, R2, 2, R8, -1, 0, 0, 0, 0, 0,
    91, 182, 273, 365, 0, R2, 2, R2, 3, R8, -1, 0, 0, 0,
    91, 182, 273, 365, 0, R8, 1, R2, 3, R2, 4, R8, -1, 0,
    91, 182, 273, 365, 1, R8, 2, R2, 4, R2, 5, R8, -1, 0,
    91, 182, 273, 365, 2, R8, 3, R2, 5, R2, 6, R8, -1, 0,
    91, 182, 273, 365, 3, R8, 4, R2, 6, R2, 7, R8, -1, 0,
    91, 182, 273, 365, 4, R8, 5, R2, 7, R2, 8, R8, -1, 0,
    91, 182, 273, 365, 5, R8, 6, R2, 8, R2, 9, R8, -1, 0,
    91, 182, 273, 365, 6, R8, 7, R2, 9, R2, -1, 0, 0, 0,
    91, 182, 273, 365, 7, R8, 8, R2, -1, 0, 0, 0, 0, 0], float).reshape(10, 14).T

L = np.array([[91.0] * 10, [182.0] * 10, [273.0] * 10, [365.0] * 10])


def _reference_d():
    return sp.csc_matrix((np.array(D_X), np.array(D_I), np.array(D_P)), shape=(10, 10))


def test_distance_matrix_golden():
    d = si.distance_matrix(np.arange(1, 11), np.arange(1, 11), 3)
    assert list(d.indices) == D_I and list(d.indptr) == D_P
    assert np.all(np.abs(d.data - np.array(D_X)) < TOL)


def test_distance_matrix_errors():
    """tests/ldata_sp.R:47-108."""
    cases = [
        (dict(x=np.ones(10), y=np.ones(10), cutoff=3, min_dist="min_dist"),
         "Invalid 'min_dist' argument. Please provide 'min_dist' > 0."),
        (dict(x=np.ones(10), y=np.ones(10), cutoff=3, min_dist=-1),
         "Invalid 'min_dist' argument. Please provide 'min_dist' > 0."),
        (dict(x=[], y=[1], cutoff=2), "'x' must be a numeric vector with length > 0."),
        (dict(x=[1, 2, 3], y=[1, 2], cutoff=2), "'y' must be a numeric vector with length 3."),
        (dict(x=[1, 2, 3], y=[1, 2, 3], cutoff=-2), "'cutoff' must be > 0."),
        (dict(x=[1, 2, 3], y=[1, 2, 3], cutoff=np.inf), "'cutoff' must be > 0."),
        (dict(x=[1, 2, 3], y=[4, np.nan, 6], cutoff=1), "Invalid distance for i=0 and j=1."),
        (dict(x=[1, 2, 3], y=[4, np.nan, 6], cutoff=3, na_fail=1),
         "'na_fail' must be TRUE or FALSE."),
    ]
    for kw, msg in cases:
        with pytest.raises(SpatialError) as e:
            si.distance_matrix(**kw)
        assert str(e.value) == msg


def test_distance_matrix_na_skipped():
    """tests/ldata_sp.R:87-99: with na_fail FALSE a missing coordinate
    drops the pairs it is in."""
    d = si.distance_matrix([1, 2, 3], [4, np.nan, 6], 3, na_fail=False)
    assert list(d.indices) == [2, 0] and list(d.indptr) == [0, 1, 1, 2]
    assert np.all(np.abs(d.data - 2.828427125) < TOL)


def test_distance_matrix_min_dist():
    """Coincident points take min_dist."""
    d = si.distance_matrix([0, 0, 3], [0, 0, 0], 1.0, min_dist=0.25)
    assert list(d.indptr) == [0, 1, 2, 2] and list(d.indices) == [1, 0]
    assert list(d.data) == [0.25, 0.25]


def test_ldata_sp_golden():
    d = _reference_d()
    assert np.all(np.abs(ldata_sp(L, d, 0) - LDATA_DEGREE) < TOL)
    got = si.add_spatial_coupling_to_ldata(np.arange(1, 11), np.arange(1, 11), 3, ldata=L)
    assert np.all(np.abs(got - LDATA_DIST) < TOL)
    got = si.add_spatial_coupling_to_ldata(np.arange(1, 11), np.arange(1, 11), 3)
    assert np.all(np.abs(got - LDATA_DIST[4:]) < TOL)
    # metric 2: 1 / distance^2
    got = ldata_sp(L, d, 2)
    assert np.all(np.abs(got[5, :2] - [1 / R2 ** 2, 1 / R2 ** 2]) < 1e-6)


def test_ldata_sp_errors():
    """tests/ldata_sp.R:110-144."""
    d = _reference_d()
    cases = [((None, d, 0), "Invalid 'data' argument."),
             ((d, d, 0), "Invalid 'data' argument."),
             ((np.arange(1, 11), d, 0), "Invalid 'data' argument."),
             ((L, None, 0), "Invalid 'distance' argument."),
             ((L, L, 0), "Invalid 'distance' argument."),
             ((L, d, None), "Invalid 'metric' argument."),
             ((L, d, 0.0), "Invalid 'metric' argument."),
             ((L, d, [0, 0]), "Invalid 'metric' argument."),
             ((L[:, 1:], d, 0), "The number of nodes in 'data' and 'distance' are not equal.")]
    for args, msg in cases:
        with pytest.raises(SpatialError) as e:
            ldata_sp(*args)
        assert str(e.value) == msg


def test_distance_argument_checks():
    """tests/SISe_sp.R:298-334 and the dgCMatrix check (:950-968)."""
    u0 = {"S": np.arange(9), "I": np.zeros(9, int)}
    for dist, msg in ((None, "'distance' is missing."),
                      (np.eye(9), "The 'distance' argument must be of type 'dgCMatrix'."),
                      (sp.csc_matrix(-np.eye(9)), "All values in the 'distance' matrix must be >= 0.")):
        with pytest.raises(ValueError) as e:
            si.SISe_sp(u0, np.arange(10.0), coupling=0.0005, distance=dist, **C.SISE_SP)
        assert str(e.value) == msg


def test_sise_sp_susceptible_only_golden(oracle):
    """tests/SISe_sp.R:990-1037: nine herds of 0..8 susceptible animals on
    a 3 x 3 grid (distance cutoff 2), phi 0: S stays 0..8 and I 0 over
    tspan 0..9."""
    g = np.array([(x, y) for y in range(1, 4) for x in range(1, 4)], float)
    dist = si.distance_matrix(g[:, 0], g[:, 1], 2)
    m = si.SISe_sp({"S": np.arange(9), "I": np.zeros(9, int)}, np.arange(10.0),
                   phi=np.zeros(9), coupling=0.0005, distance=dist, **C.SISE_SP)
    for solver in ("ssm", "aem"):
        r = oracle.run(m, solver=solver, seed=123)
        assert r.err == 0
        U = r.U.reshape(9, 2, 10)
        assert (U[:, 0, :] == np.arange(9)[:, None]).all()
        assert (U[:, 1, :] == 0).all()


def test_local_spread_is_the_linear_map(oracle):
    """With no transitions, decay or infected animals, phi moves by the
    local spread alone, a linear map computed here in matrix form; one
    herd is empty (it neither gives nor takes)."""
    n, days, D = 40, 30, 0.02
    dist = C.scattered(n, 5)
    rng = np.random.default_rng(2)
    N = rng.integers(1, 80, n)
    N[3] = 0
    phi0 = rng.uniform(0, 2, n)
    m = C.frozen_sise_sp(N, phi0, dist, D, days)
    want = C.spread_by_matrix(N, phi0, dist, D, days)
    for solver, threads in (("ssm", 1), ("aem", 1), ("ssm", 4)):
        r = oracle.run(m, solver=solver, seed=1, threads=threads)
        assert r.err == 0
        np.testing.assert_allclose(r.V, want, rtol=1e-12, atol=1e-15)
    # sum(N * phi) is conserved: every pair's flows cancel (symmetric d).
    tot = (want * N[:, None]).sum(axis=0)
    np.testing.assert_allclose(tot, (phi0 * N).sum(), rtol=1e-12)


def test_local_spread_two_herds_closed_form(oracle):
    """Two herds of 10 at distance 2, coupling 0.1: each day the gap in
    phi shrinks by 1 - 2 * 0.1 / 2 = 0.9 and the mean stays; a function
    that read the neighbour's new phi instead of its old one would break
    both."""
    dist = si.distance_matrix([0.0, 2.0], [0.0, 0.0], 3.0)
    m = C.frozen_sise_sp([10, 10], [1.0, 0.0], dist, 0.1, 25)
    r = oracle.run(m, solver="ssm", seed=1)
    assert r.err == 0
    t = np.arange(1, 26)
    np.testing.assert_allclose(r.V[0] - r.V[1], 0.9 ** t, rtol=1e-12)
    np.testing.assert_allclose(r.V[0] + r.V[1], 1.0, rtol=1e-14)


def test_local_spread_uses_the_days_counts(oracle):
    """The neighbours' N_j after the day's events: a herd that receives
    animals by an external transfer pulls less phi per animal.  Two herds
    at distance 1; on day 1 herd 2 sends 5 of its 10 animals to herd 1."""
    dist = si.distance_matrix([0.0, 1.0], [0.0, 0.0], 2.0)
    m = C.frozen_sise_sp([10, 10], [0.0, 1.0], dist, 0.1, 2)
    from siminf_amd.model import make_events
    ev = {"event": [3], "time": [1], "node": [2], "dest": [1], "n": [5],
          "proportion": [0.0], "select": [1], "shift": [0]}
    m.events = make_events(m.E, None, ev)
    r = oracle.run(m, solver="ssm", seed=1)
    assert r.err == 0
    # after the transfer N = (15, 5), phi = (0, 1)
    phi1 = 0.0 + (1.0 * 5 - 0.0 * 15) * 0.1 / (15 * 1.0)
    phi2 = 1.0 + (0.0 * 15 - 1.0 * 5) * 0.1 / (5 * 1.0)
    assert abs(r.V[0, 0] - phi1) < 1e-15 and abs(r.V[1, 0] - phi2) < 1e-15


def test_zero_coupling_equals_sise(oracle):
    """At coupling 0 SISe_sp is SISe with epsilon 0 and SISe3_sp is SISe3
    with epsilon 0, step for step (same streams)."""
    m_sp = C.sise_sp_model(n=80, days=120, seed=4, coupling=0.0)
    kw = dict(C.SISE_SP)
    u0 = m_sp.u0.T
    m = si.SISe(u0, m_sp.tspan, phi=m_sp.v0[0], epsilon=0.0, **kw)
    for solver in ("ssm", "aem"):
        a = oracle.run(m_sp, solver=solver, seed=9)
        b = oracle.run(m, solver=solver, seed=9)
        assert a.err == 0 and b.err == 0
        assert (a.U == b.U).all() and (a.V == b.V).all(), solver
        assert a.transitions > 0
    m3 = C.sise3_sp_model(n=200, days=200, seed=5, coupling=0.0)
    g = dict(zip(m3.gdata_names, m3.gdata))
    g.pop("coupling")
    base = si.SISe3(m3.u0.T, m3.tspan, events=m3.events, phi=0.0, epsilon=0.0,
                    end_t1=91, end_t2=182, end_t3=273, end_t4=365, **g)
    a = oracle.run(m3, solver="ssm", seed=3)
    b = oracle.run(base, solver="ssm", seed=3)
    assert a.err == 0 and b.err == 0
    assert (a.U == b.U).all() and (a.V == b.V).all()


def test_coupling_changes_the_epidemic(oracle):
    """With coupling the trajectories differ from coupling 0 (the spread
    term is not silently dropped)."""
    a = oracle.run(C.sise_sp_model(n=80, days=120, seed=4, coupling=0.01), seed=9)
    b = oracle.run(C.sise_sp_model(n=80, days=120, seed=4, coupling=0.0), seed=9)
    assert a.err == 0 and b.err == 0
    assert not np.array_equal(a.V, b.V)


def test_sise3_sp_population_conserved_without_events(oracle):
    m = C.sise3_sp_model(n=150, days=150, seed=6, events=False, coupling=0.01)
    r = oracle.run(m, solver="ssm", seed=2)
    assert r.err == 0
    U = r.U.reshape(150, 6, 150)
    assert (U.sum(axis=1) == m.u0.sum(axis=0)[:, None]).all()
    assert (r.V >= 0).all()


@pytest.mark.parametrize("coupling, code", [(-1.0, -10), (np.inf, -7)])
def test_spread_errors(oracle, coupling, code):
    """A negative phi (-10) or a non-finite phi (-7) from the spread term:
    herd 1 (phi 0) next to herd 2 (phi 1) at distance 1."""
    dist = si.distance_matrix([0.0, 1.0], [0.0, 0.0], 2.0)
    m = C.frozen_sise_sp([10, 10], [0.0, 1.0], dist, coupling, 5)
    for solver in ("ssm", "aem"):
        r = oracle.run(m, solver=solver, seed=1)
        assert r.err == code, solver


def test_distance_matrix_many_points_same_as_pairwise():
    """Above 2048 points the k-d tree path gives the pairwise loop's
    matrix exactly (indices, order and distances), points on the cutoff
    included; and coincident points take min_dist."""
    import siminf_amd.spatial as S
    rng = np.random.default_rng(4)
    n = 3000
    x = np.round(rng.uniform(0, 40, n), 1)          # many exact ties at the cutoff
    y = np.round(rng.uniform(0, 40, n), 1)
    x[5], y[5] = x[6], y[6]
    tree = si.distance_matrix(x, y, 1.0, min_dist=0.01)
    loop = S.distance_matrix(x[:2048], y[:2048], 1.0, min_dist=0.01)   # pairwise path
    assert tree.shape == (n, n)
    sub = tree[:2048, :2048].tocsc()
    sub.sort_indices()
    want = sp.csc_matrix(loop)
    want.sort_indices()
    assert (sub != want).nnz == 0
    # the first 2048 points' columns hold only neighbours among themselves
    # or beyond: compare the full tree against a brute force on all pairs
    d = np.hypot(x[:, None] - x[None, :], y[:, None] - y[None, :])
    np.fill_diagonal(d, np.inf)
    d[d <= 0] = 0.01
    ii, jj = np.nonzero(d.T <= 1.0)
    assert tree.nnz == len(ii)
    np.testing.assert_array_equal(tree.toarray().T[ii, jj] > 0, True)
    assert list(tree.indices[tree.indptr[0]:tree.indptr[1]]) == sorted(
        tree.indices[tree.indptr[0]:tree.indptr[1]])
