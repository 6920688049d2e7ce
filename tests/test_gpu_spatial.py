"""GPU parity of the spatial models SISe_sp and SISe3_sp, whose
post-time-step function reads other nodes' state (SISe_sp.c:90-138):
they run in the coupled kernel, which publishes every node's u and v
each day and passes the function pointers into those arrays.  Every case
equals the checker bit for bit (error code, U, V, transition count):
grid launches (one replicate; several replicates per launch), one
workgroup per replicate, ssm and aem, with and without scheduled events,
autotuned herd maps, and the local spread's linear map and error codes
(tests/spatial_cases.py)."""
import numpy as np
import pytest

import siminf_amd as si
from tests import spatial_cases as C
from tests.test_gpu_parity import _check

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", ["ssm", "aem"])
def test_sise_sp_one_replicate(oracle, solver):
    """One replicate of 300 herds: the grid-wide coupled kernel."""
    m = C.sise_sp_model(n=300, days=200, seed=3)
    ref, _ = _check(m, oracle, solver=solver, seed=11)
    assert ref.err == 0 and ref.transitions > 0


def test_sise_sp_replicates_in_workgroups(oracle):
    """Four replicates of 300 herds: one workgroup per replicate."""
    m = C.sise_sp_model(n=300, days=150, seed=4, replicates=4)
    ref, _ = _check(m, oracle, seed=12)
    assert ref.err == 0 and ref.transitions > 0


@pytest.mark.parametrize("solver", ["ssm", "aem"])
def test_sise_sp_with_events(oracle, solver):
    """Enter events and external transfers change the counts the spread
    reads on the days they happen."""
    m = C.sise_sp_model(n=250, days=120, seed=5, events=True)
    ref, _ = _check(m, oracle, solver=solver, seed=13)
    assert ref.err == 0


@pytest.mark.parametrize("solver", ["ssm", "aem"])
def test_sise3_sp_herds(oracle, solver):
    """1600 herds with the SISe3 herd workload's scheduled events (enter,
    exit, internal and external transfers) over a year."""
    m = C.sise3_sp_model(n=1600, days=365, seed=11, coupling=0.002)
    ref, _ = _check(m, oracle, solver=solver, seed=14)
    assert ref.err == 0 and ref.transitions > 0


def test_sise3_sp_replicates_per_launch(oracle):
    """700 herds, 3 replicates sharing cooperative launches."""
    m = C.sise3_sp_model(n=700, days=60, seed=12, coupling=0.002, replicates=3)
    ref, _ = _check(m, oracle, seed=15)
    assert ref.err == 0


def test_sise3_sp_autotuned_herd_map(oracle):
    """The herd map the autotuner picks (busiest herds on wavefronts of
    their own) leaves the results unchanged."""
    m = C.sise3_sp_model(n=900, days=120, seed=13, coupling=0.002)
    ref = oracle.run(m, solver="ssm", seed=16)
    assert ref.err == 0
    with si.Plan(m) as p:
        p.autotune(seed=3)
        for solo, per_wave in ((None, None), (list(range(5)), 16)):
            if solo is not None:
                p.set_herd_map(solo=solo, herds_per_wavefront=per_wave)
            print("case: SISe3_sp 900 herds, herd map", p.herd_map(), flush=True)
            p.run(16)
            p.download()
            np.testing.assert_array_equal(m.U, ref.U)
            np.testing.assert_array_equal(m.V, ref.V)


def test_local_spread_linear_map_on_device(oracle):
    """phi under the local spread alone equals the matrix form of the map
    (to 1e-12) and the checker (bit for bit)."""
    n, days, D = 200, 40, 0.02
    dist = C.scattered(n, 7)
    rng = np.random.default_rng(8)
    N = rng.integers(1, 80, n)
    N[[3, 50]] = 0
    phi0 = rng.uniform(0, 2, n)
    m = C.frozen_sise_sp(N, phi0, dist, D, days)
    want = C.spread_by_matrix(N, phi0, dist, D, days)
    for solver in ("ssm", "aem"):
        _, b = _check(m, oracle, solver=solver, seed=1)
        np.testing.assert_allclose(b.V, want, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("coupling, code", [(-1.0, -10), (np.inf, -7)])
def test_spread_errors_on_device(oracle, coupling, code):
    dist = si.distance_matrix([0.0, 1.0], [0.0, 0.0], 2.0)
    m = C.frozen_sise_sp([10, 10], [0.0, 1.0], dist, coupling, 5)
    for solver in ("ssm", "aem"):
        ref, _ = _check(m, oracle, solver=solver, seed=1)
        assert ref.err == code


def test_bad_neighbor_lists_refused():
    """A neighbour index outside the nodes, or a list without its stop
    index, is refused when the plan is made ("Invalid model."), before
    any kernel could read outside the state arrays."""
    m = C.sise_sp_model(n=20, days=5, seed=2)
    bad_index = m.ldata.copy()
    bad_index[4, 3] = 20.0
    no_stop = m.ldata.copy()
    no_stop[4:, 5] = 1.0
    for ld in (bad_index, no_stop):
        m.ldata = np.asfortranarray(ld)
        with pytest.raises(si.SimInfError) as e:
            si.run(m, seed=1)
        assert "Invalid model" in str(e.value)


def test_pfilter_on_a_spatial_model(oracle):
    """Particles of SISe_sp (one workgroup per particle, each reading its
    own particle's neighbours) give the checker's filter bit for bit."""
    from tests.test_gpu_pfilter import _check as pf_check
    m = C.sise_sp_model(n=60, days=50, seed=9)
    times = np.arange(5, 51, 5.0)
    rng = np.random.default_rng(4)
    obs = rng.poisson(5.0, size=(len(times), 1))
    err, _ = pf_check(oracle, m, 1, times, obs, 48, seed=7, p=0.5, dist="poisson",
                      offset=1.0)
    assert err == 0


def test_sparse_output_of_a_spatial_model(oracle):
    """Sparse U and V (punchcard) of SISe_sp with two replicates equal
    the dense run's values and the checker's."""
    from siminf_amd.punchcard import punchcard
    from tests.test_punchcard import sp_values
    m = C.sise_sp_model(n=150, days=60, seed=10, replicates=2)
    print("case: SISe_sp 150 herds x 2 replicates, punchcard", flush=True)
    si.run(m, seed=6)
    U, V = m.U.copy(), m.V.copy()
    rng = np.random.default_rng(2)
    punchcard(m, {"node": rng.integers(1, 151, 400), "time": rng.integers(1, 61, 400),
                  "S": rng.random(400) < 0.5, "I": rng.random(400) < 0.7,
                  "phi": rng.random(400) < 0.6})
    si.run(m, seed=6)
    cu, cv = m.U_sparse.tocoo(), m.V_sparse.tocoo()
    np.testing.assert_array_equal(cu.data, U[cu.row, cu.col].astype(float))
    np.testing.assert_array_equal(cv.data, V[cv.row, cv.col])
    ref = oracle.run(m, seed=6)
    assert ref.err == 0
    np.testing.assert_array_equal(cu.data, sp_values(m.U_sparse, ref.prU))
    np.testing.assert_array_equal(cv.data, sp_values(m.V_sparse, ref.prV))


def _rep_ldata(n, R, seed):
    """Per-replicate ldata of a spatial model: each replicate its own
    interval ends and its own neighbour lists (cutoffs 1.0, 0.9, ...),
    padded to the longest list."""
    from siminf_amd.spatial import ldata_sp
    rng = np.random.default_rng(seed)
    k = int(np.ceil(np.sqrt(n)))
    gx, gy = np.meshgrid(np.arange(k), np.arange(k))
    x = 0.8 * gx.ravel()[:n] + rng.uniform(-0.3, 0.3, n)
    y = 0.8 * gy.ravel()[:n] + rng.uniform(-0.3, 0.3, n)
    cols = []
    for r in range(R):
        ends = np.vstack([np.full(n, e) for e in rng.permutation([91, 182, 273, 365])])
        cols.append(ldata_sp(ends, si.distance_matrix(x, y, 1.0 - 0.1 * r), 1))
    rows = max(c.shape[0] for c in cols)
    return np.hstack([np.vstack([c, np.zeros((rows - c.shape[0], n))]) for c in cols])


@pytest.mark.parametrize("n", [200, 600])
def test_spatial_ldata_per_replicate(oracle, n):
    """Replicates with their own neighbour lists (one workgroup per
    replicate at 200 herds, grid launches at 600) equal the checker given
    the same per-replicate ldata; a list naming a node outside the model
    is refused."""
    R = 3
    ld = _rep_ldata(n, R, seed=n)
    m = C.sise_sp_model(n=n, days=80, seed=11, replicates=R)
    # The model's ldata and the per-replicate lists padded to one height.
    rows = max(ld.shape[0], m.ldata.shape[0])
    ld = np.vstack([ld, np.zeros((rows - ld.shape[0], ld.shape[1]))])
    m.ldata = np.asfortranarray(np.vstack([m.ldata, np.zeros((rows - m.ldata.shape[0], n))]))
    m.Nld = m.ldata.shape[0]
    m.ldata_names = ["L%d" % (i + 1) for i in range(m.Nld)]
    print("case: SISe_sp %d herds x %d replicates, per-replicate ldata" % (n, R), flush=True)
    ref = oracle.run(m, seed=5, rep_ldata=ld)
    assert ref.err == 0
    with si.Plan(m) as p:
        p.set_replicate_data(ldata=ld)
        tr, _, _ = p.run(5)
        out = p.download()
        np.testing.assert_array_equal(out.U, ref.U)
        np.testing.assert_array_equal(out.V, ref.V)
        assert tr == ref.transitions
        bad = ld.copy()
        bad[4, n + 7] = float(n)
        with pytest.raises(si.SimInfError, match="Invalid model."):
            p.set_replicate_data(ldata=bad)
