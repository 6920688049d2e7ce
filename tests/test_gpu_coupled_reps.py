"""GPU parity of coupled models too large for one workgroup run with many
replicates: the replicates share cooperative launches (each with its own
barrier words and error key), bit for bit against the checker's mssm,
which runs them one after another (SimInf_solver_mssm.c:88-105)."""
import numpy as np
import pytest

import siminf_amd as si
from siminf_amd import workloads as W
from tests.test_gpu_parity import _check

pytestmark = pytest.mark.gpu


def test_sise3_1600_herds_replicates(oracle):
    """1600 herds with scheduled events (external transfers) and 6
    replicates in one launch."""
    m = W.sise3_herds(Nn=1600, days=150, seed=21, replicates=6)
    ref, _ = _check(m, oracle, seed=31)
    assert ref.err == 0 and ref.transitions > 0


def test_replicates_over_several_launches(oracle):
    """More replicates than one cooperative launch holds: the plan runs
    them in several launches."""
    m = W.sise3_herds(Nn=600, days=20, seed=22, replicates=400)
    with si.Plan(m) as p:
        # The default map: 10 herd workgroups and the E2 workgroup per
        # replicate.
        assert p.herd_map() == (0, 64)
    ref, _ = _check(m, oracle, seed=32)
    assert ref.err == 0


def test_error_in_one_replicate_only(oracle):
    """Replicate 2's herd 5 cannot supply an exit event the others can:
    the run reports that replicate's error (the first in replicate order),
    and the other replicates' workgroups, which share the launch, finish
    their days without waiting on it."""
    base = W.sise3_herds(Nn=700, days=30, seed=23, events=False)
    R, Nn = 5, base.Nn
    ev = {"event": [0, 3], "time": [10, 12], "node": [5, 8], "dest": [0, 9],
          "n": [20, 1], "proportion": [0.0, 0.0], "select": [6, 6], "shift": [0, 0]}
    names = ("S_1", "I_1", "S_2", "I_2", "S_3", "I_3")
    m = si.SISe3(u0={c: base.u0[k] for k, c in enumerate(names)},
                 tspan=np.arange(1, 31.0), events=ev, phi=0.1, end_t1=91, end_t2=182,
                 end_t3=273, end_t4=365, replicates=R, **W.SISE3_GDATA)
    # Herd 5 (zero-based 4): 40 individuals in every compartment, except
    # S_3 and I_3 of replicate 2, which hold 3 and 2 (the exit event takes
    # 20 of S_3 + I_3, a total the SSA leaves unchanged).
    for r in range(R):
        m._u0[:, r * Nn + 4] = 40
    m._u0[4:6, 2 * Nn + 4] = (3, 2)
    ref, _ = _check(m, oracle, seed=33)
    assert ref.err == -8


def test_coupled_size_bound(oracle):
    """One coupled replicate's herd workgroups must be resident together:
    40,000 herds run (bit for bit against the checker), 200,000 are
    refused at plan creation with the reference's allocation message."""
    m = W.sise3_herds(Nn=40_000, days=8, seed=24)
    ref, _ = _check(m, oracle, seed=34)
    assert ref.err == 0
    big = W.sise3_herds(Nn=200_000, days=3, seed=25)
    print("case: SISe3 200000 herds refused", flush=True)
    with pytest.raises(si.SimInfError, match="Unable to allocate memory buffer."):
        si.Plan(big)
