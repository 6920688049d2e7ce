"""The node kernels' launch shape (lanes per wavefront of the active pass)
changes time, never results; the autotuner picks a shape and the plan's
trajectories stay bit-identical."""
import numpy as np
import pytest

import siminf_amd as si
from siminf_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_lanes_per_wave_do_not_change_results():
    m = W.seir_million(Nn=20000, days=60, seed=3, seeded_frac=0.02)
    print("case: SEIR 20000 nodes x 60 days, lanes per wave", flush=True)
    with si.Plan(m) as plan:
        plan.run(5)
        ref = plan.download().U.copy()
        for shift in (0, 3, 6):
            plan.set_lanes_per_wave(shift)
            plan.run(5)
            np.testing.assert_array_equal(plan.download().U, ref)
        best, ms = plan.autotune(seed=5, reps=1)
        assert 0 <= best <= 6 and all(t > 0 for t in ms)
        assert ms[best] == min(ms)
        plan.run(5)
        np.testing.assert_array_equal(plan.download().U, ref)


def test_fill_blocks_do_not_change_results():
    """The fill kernel's columns per block (the trajectory-flush schedule)
    change time, never results: blocks of 1, 3, 8, 64 and more columns
    than the run has, then the autotuner's choice."""
    m = W.seir_million(Nn=20000, days=60, seed=4, seeded_frac=0.02)
    print("case: SEIR 20000 nodes x 60 days, fill blocks", flush=True)
    with si.Plan(m) as plan:
        assert plan.fill_block() == 8
        plan.run(6)
        ref = plan.download().U.copy()
        for cols in (1, 3, 64, 4096):
            plan.set_fill_block(cols)
            assert plan.fill_block() == cols
            plan.run(6)
            np.testing.assert_array_equal(plan.download().U, ref)
        plan.autotune(seed=5, reps=1)
        assert plan.fill_block() in (4, 8, 16, 32, 64)
        plan.run(6)
        np.testing.assert_array_equal(plan.download().U, ref)


def test_herd_maps_do_not_change_results(oracle):
    """A grid-wide coupled plan (one replicate with external transfers):
    herds moved to wavefronts of their own, or packed fewer to a
    wavefront, give the same trajectory, and the autotuner's choice too."""
    m = W.sise3_herds(Nn=300, days=60, seed=2)
    print("case: SISe3 300 herds x 60 days, herd maps", flush=True)
    ref = oracle.run(m, seed=4)
    assert ref.err == 0
    with si.Plan(m) as plan:
        assert plan.herd_map() == (0, 64)
        for solo, pack in (([], 64), ([7], 64), ([299, 0, 150], 16), (list(range(16)), 8),
                           ([], 8), ([5, 6], 33)):
            plan.set_herd_map(solo, pack)
            assert plan.herd_map() == (len(solo), pack)
            plan.run(4)
            np.testing.assert_array_equal(plan.download().U, ref.U)
        best, ms = plan.autotune(seed=4, reps=1)
        assert len(ms) == 8 and all(t > 0 for t in ms)
        solo, pack = plan.herd_map()
        assert best == solo and solo in (0, 2, 4, 8, 16) and pack in (64, 32, 16, 8)
        plan.run(4)
        np.testing.assert_array_equal(plan.download().U, ref.U)
        for bad in (([3, 3], 64), (list(range(17)), 64), ([], 7), ([], 65), ([300], 64)):
            with pytest.raises(si.SimInfError):
                plan.set_herd_map(*bad)


def test_herd_maps_aem(oracle):
    """The all-events method in the coupled kernel under other maps."""
    m = W.sise3_herds(Nn=300, days=60, seed=2)
    print("case: SISe3 300 herds x 60 days, aem, herd maps", flush=True)
    ref = oracle.run(m, solver="aem", seed=6)
    assert ref.err == 0 and ref.aem_ties == 0
    with si.Plan(m, solver="aem") as plan:
        for solo, pack in (([], 64), ([1, 2, 3, 250], 8), (list(range(16)), 32)):
            plan.set_herd_map(solo, pack)
            plan.run(6)
            np.testing.assert_array_equal(plan.download().U, ref.U)


def test_autotune_replicated_coupled_plan_has_nothing_to_tune():
    m = W.sise3_herds(Nn=100, days=20, seed=2, replicates=3)
    print("case: SISe3 replicates autotune", flush=True)
    with si.Plan(m) as plan:
        best, _ = plan.autotune(reps=1)
        assert best == -1
