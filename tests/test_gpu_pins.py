"""GPU tests of the SEIR and SISe3 plugin bodies and the reference's event
and post-time-step error cases: each case runs through the C ABI and must
equal the checker bit for bit (error code, U, V and transition count),
and the device's own output must follow the exact laws of
tests/pin_cases.py."""
import numpy as np
import pytest

from tests import pin_cases as C
from tests.test_gpu_parity import _check, _gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("solver", ["ssm", "aem"])
def test_seir_brute_force_on_device(oracle, solver):
    n = 30000
    m = C.seir_brute_force_model(n)
    _, b = _check(m, oracle, solver=solver, seed=21)
    assert C.seir_brute_force_p(b.U, n) > 1e-4


def test_seir_pure_death_on_device(oracle):
    n = 8000
    m = C.seir_pure_death_model(n)
    for solver in ("ssm", "aem"):
        _, b = _check(m, oracle, solver=solver, seed=5)
        assert min(C.seir_pure_death_ps(b.U, n)) > 1e-4, solver


@pytest.mark.parametrize("param", [c[0] for c in C.SISE3_SINGLE])
def test_sise3_one_transition_at_a_time_on_device(oracle, param):
    n = 4000
    m = C.sise3_single_model(param, n)
    for solver in ("ssm", "aem"):
        _, b = _check(m, oracle, solver=solver, seed=7)
        assert min(C.sise3_single_ps(b.U, b.V, n, param)) > 1e-4, (param, solver)


def test_sise3_age_classes_on_device(oracle):
    n = 20000
    m = C.sise3_chain_model(n)
    for solver in ("ssm", "aem"):
        _, b = _check(m, oracle, solver=solver, seed=9)
        assert min(C.sise3_chain_ps(b.U, n)) > 1e-4, solver


@pytest.mark.parametrize("ends", C.PHI_CASES)
def test_sise3_phi_recurrence_on_device(oracle, ends):
    m = C.phi_model(ends)
    exp = C.phi_expected(ends)
    for solver in ("ssm", "aem"):
        _, b = _check(m, oracle, solver=solver, seed=1)
        np.testing.assert_allclose(b.V, exp, rtol=1e-13, atol=0)


@pytest.mark.parametrize("case", C.error_cases(), ids=lambda c: c[0])
def test_reference_event_and_pts_errors_on_device(oracle, case):
    name, build, code = case
    for solver in ("ssm", "aem"):
        ref, _ = _check(build(), oracle, solver=solver, seed=3)
        assert ref.err == code, (name, solver)


def test_enter_shift_known_answer_on_device():
    m, U = C.enter_shift_known_answer()
    for solver in ("ssm", "aem"):
        err, b, _ = _gpu(m, solver=solver, seed=3)
        assert err == 0
        np.testing.assert_array_equal(b.U, U)


def test_where_a_plan_failed():
    """SimInf_hip_plan_last_error names what the reference prints on
    failure: the event (its index in the sorted event list) for event
    errors, the node for a post-time-step error; and nothing after a run
    that succeeded."""
    import siminf_amd as si
    cases = {c[0]: c for c in C.error_cases()}
    for name, phase, key, value in (("dest_out_of_bounds", "external transfers", "event", 0),
                                    ("exit_empty_select", "scheduled events", "event", 0),
                                    ("sise3_v_negative", "post time step", "node", None)):
        m = cases[name][1]()
        print("case: last error of %s" % name, flush=True)
        with si.Plan(m) as p:
            with pytest.raises(si.SimInfError) as e:
                p.run(3)
            d = e.value.detail
            assert d is not None and d["code"] == cases[name][2] and d["phase"] == phase, d
            assert key in d and d["replicate"] == 0
            if value is not None:
                assert d[key] == value
            else:
                assert 0 <= d[key] < m.Nn
            assert p.last_error() == d
    m = C.seir_pure_death_model(50)
    with si.Plan(m) as p:
        p.run(1)
        assert p.last_error() is None
