"""GPU parity on edge cases: one node and one time point, fractional
tspan with events before its start, empty nodes, common random numbers
over replicates, a hub herd whose external transfers chain through many
levels a day, many replicates of a coupled model, and the solver-level
errors (tests/SIR.R, tests/solver_aem.R, tests/SimInf.R cases)."""
import numpy as np
import pytest

import siminf_amd as si
from siminf_amd import models
from tests.test_gpu_parity import _check

pytestmark = pytest.mark.gpu


def _events(rows):
    rows = sorted(rows, key=lambda r: r[1])
    return {k: np.array([r[i] for r in rows]) for i, k in enumerate(
        ("event", "time", "node", "dest", "n", "proportion", "select", "shift"))}


@pytest.mark.parametrize("solver", ["ssm", "aem"])
def test_single_node_single_time_point(oracle, solver):
    m = si.SIR(u0={"S": [90], "I": [10], "R": [0]}, tspan=[5.0], beta=0.3, gamma=0.1)
    _check(m, oracle, solver=solver, seed=7)
    m = si.SEIR(u0={"S": [50], "E": [2], "I": [3], "R": [0]}, tspan=[0.0, 40.0],
                beta=0.4, epsilon=0.2, gamma=0.1)
    _check(m, oracle, solver=solver, seed=8)


def test_fractional_tspan_and_events_before_start(oracle):
    rows = [(1, t, nd, 0, 2, 0.0, 2, 0) for t in (0, 1, 3, 10) for nd in (1, 4, 7)]
    rows += [(0, 2, 2, 0, 0, 0.5, 4, 0), (0, 6, 5, 0, 3, 0.0, 4, 0)]
    m = si.SIR(u0={"S": np.arange(20, 30), "I": np.arange(10) % 3, "R": 0},
               tspan=[0.5, 1.7, 3.2, 10.9], beta=0.3, gamma=0.1, events=_events(rows))
    _check(m, oracle, seed=3)


def test_empty_nodes(oracle):
    m = si.SIR(u0={"S": [0, 50, 0, 30], "I": [0, 3, 0, 0], "R": [0, 0, 0, 4]},
               tspan=np.arange(1, 31.0), beta=0.3, gamma=0.1)
    _check(m, oracle, seed=2)
    # The parsed model divides by N: an empty node gives a NaN rate (-11).
    mp = si.SIR_mparse(u0={"S": [0, 50], "I": [0, 3], "R": [0, 0]},
                       tspan=np.arange(1, 11.0), beta=0.16, gamma=0.077)
    ref, _ = _check(mp, oracle, seed=2)
    assert ref.err == -11


def test_common_random_numbers_over_replicates(oracle):
    m = si.SIR(u0={"S": np.arange(40, 60), "I": np.arange(20) % 4, "R": 0},
               tspan=np.arange(1, 41.0), beta=0.25, gamma=0.1, replicates=4)
    _check(m, oracle, seed=11, crn=True)


def _sise3(nodes, days, events, replicates=1):
    rng = np.random.default_rng(1)
    S1 = rng.integers(50, 120, nodes)
    S1[0] = 5000
    u0 = {"S_1": S1, "I_1": rng.integers(0, 5, nodes), "S_2": rng.integers(10, 40, nodes),
          "I_2": 0, "S_3": rng.integers(50, 200, nodes), "I_3": rng.integers(0, 3, nodes)}
    return si.SISe3(u0=u0, tspan=np.arange(1, days + 1.0), events=events, phi=0.5,
                    end_t1=91, end_t2=182, end_t3=273, end_t4=365, replicates=replicates,
                    upsilon_1=0.02, upsilon_2=0.02, upsilon_3=0.02, gamma_1=0.1, gamma_2=0.1,
                    gamma_3=0.1, alpha=1.0, beta_t1=0.1, beta_t2=0.1, beta_t3=0.125,
                    beta_t4=0.125, epsilon=0.0)


@pytest.mark.parametrize("solver", ["ssm", "aem"])
def test_hub_herd_transfer_chains(oracle, solver):
    """Herd 1 sends to and receives from the others 40 times a day: the
    day's external transfers form chains of about 40 levels."""
    rng = np.random.default_rng(4)
    rows = []
    for t in range(1, 21):
        for k in range(40):
            other = int(rng.integers(2, 31))
            if k % 2:
                rows.append((3, t, 1, other, int(rng.integers(1, 4)), 0.0, 4, 0))
            else:
                rows.append((3, t, other, 1, 1, 0.0, 4, 0))
    m = _sise3(30, 25, _events(rows))
    _check(m, oracle, solver=solver, seed=5)


def test_many_replicates_coupled(oracle):
    rng = np.random.default_rng(6)
    rows = [(3, int(t), int(a), int(b), 1, 0.0, 4 + int(s), 0)
            for t, a, b, s in zip(rng.integers(1, 30, 120), rng.integers(1, 41, 120),
                                  rng.integers(1, 41, 120), rng.integers(0, 3, 120)) if a != b]
    m = _sise3(40, 30, _events(rows), replicates=150)
    _check(m, oracle, seed=9)


def test_solver_level_errors():
    m = si.SIR(u0={"S": [99], "I": [1], "R": [0]}, tspan=np.arange(1, 11.0), beta=0.2,
               gamma=0.1, replicates=2)
    with pytest.raises(si.SimInfError, match="Cannot run the 'aem' solver on a replicated model."):
        si.run(m, solver="aem", seed=1)
    with pytest.raises(si.SimInfError, match="Invalid 'solver' value."):
        si.run(m, solver="direct", seed=1)


def test_grid_coupled_replicates(oracle):
    """More herds than one workgroup holds with replicates: one
    cooperative launch per replicate (mssm), each with the E2 workgroup."""
    from siminf_amd import workloads as W
    m = W.sise3_herds(Nn=600, days=40, seed=12, replicates=2)
    _check(m, oracle, seed=13)


def test_grid_coupled_particle_filter(oracle):
    """A particle filter over 600 herds with external transfers runs each
    particle as a grid launch."""
    from siminf_amd import workloads as W
    from tests.test_gpu_pfilter import _check as pf_check
    h = W.sise3_herds(Nn=600, days=30, seed=14)
    rng = np.random.default_rng(15)
    times = np.arange(5, 31, 5.0)
    obs = rng.poisson(40.0, size=(len(times), 1))
    err, _ = pf_check(oracle, h, 1, times, obs, 6, seed=16, p=0.5, dist="poisson",
                      offset=1.0)
    assert err == 0


@pytest.mark.parametrize("coupled", [False, True])
@pytest.mark.parametrize("reps", [1, 4])
def test_rates_falling_to_zero_with_enter_events(oracle, coupled, reps):
    """Nodes whose last infected recover (total rate to zero, or a rounding
    residue at or below it) stop drawing that day, as the reference's loop
    does; later enter events then sample from the same stream positions.
    Run through the node kernel and, with one external transfer, the
    coupled kernel."""
    rng = np.random.default_rng(3)
    r0 = np.random.default_rng(5)
    S, I = r0.integers(40, 200, 20), r0.integers(1, 8, 20)
    rows = [(1, int(t), int(nd), 0, int(rng.integers(1, 4)), 0.0, 4, 0)
            for t, nd in zip(rng.integers(1, 40, 60), rng.integers(1, 21, 60))]
    if coupled:
        rows.append((3, 39, 1, 2, 1, 0.0, 1, 0))
    m = si.SIR_mparse(u0={"S": S, "I": I, "R": np.zeros(20, int)},
                      tspan=np.arange(1, 41.0), beta=0.25, gamma=0.1, events=_events(rows),
                      E=si.models.select_matrix_SIR(), replicates=reps)
    _check(m, oracle, seed=21)
