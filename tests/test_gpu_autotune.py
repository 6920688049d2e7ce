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


def test_autotune_coupled_plan_has_nothing_to_tune():
    m = W.sise3_herds(Nn=100, days=20, seed=2)
    print("case: SISe3 autotune", flush=True)
    with si.Plan(m) as plan:
        best, _ = plan.autotune(reps=1)
        assert best == -1
