"""CPU tests that pin the checker's SEIR and SISe3 plugin bodies and its
event and post-time-step error branches independently of the plugin code
they share with the device (siminf_amd/csrc/siminf_models.h): exact laws
of small chains (matrix exponentials, binomial laws), the documented
interval table of the seasonal decay, a hand-computed phi recurrence, and
the reference's error tests (tests/SimInf_events.R:537-565,
tests/sample_select.R:857-941, tests/SISe3.R:1395-1445)."""
import numpy as np
import pytest

from tests import pin_cases as C


def test_seir_ctmc_brute_force(oracle):
    n = 30000
    m = C.seir_brute_force_model(n)
    for solver in ("ssm", "aem"):
        r = oracle.run(m, solver=solver, seed=21)
        assert r.err == 0
        assert C.seir_brute_force_p(r.U, n) > 1e-4, solver


def test_seir_brute_force_catches_a_wrong_population():
    """The brute-force law separates a plausible slip: infection divided
    by S+I+R instead of S+E+I+R."""
    import scipy.linalg
    states, index, P = C.seir_one_day_law()
    Q = np.zeros((len(states), len(states)))
    for (s, e, i, r), k in index.items():
        for dst, rate in (((s - 1, e + 1, i, r), 0.9 * s * i / max(s + i + r, 1)),
                          ((s, e - 1, i + 1, r), 0.5 * e), ((s, e, i - 1, r + 1), 0.4 * i)):
            if rate > 0:
                Q[k, index[dst]] += rate
                Q[k, k] -= rate
    Pw = scipy.linalg.expm(Q)[index[C.SEIR_START]]
    rng = np.random.default_rng(0)
    draws = rng.choice(len(states), size=30000, p=Pw / Pw.sum())
    obs = np.bincount(draws, minlength=len(states))
    assert C.chisq_p(obs, P, 30000) < 1e-6


def test_seir_pure_death(oracle):
    n = 8000
    m = C.seir_pure_death_model(n)
    for solver in ("ssm", "aem"):
        r = oracle.run(m, solver=solver, seed=5)
        assert r.err == 0
        assert min(C.seir_pure_death_ps(r.U, n)) > 1e-4, solver


@pytest.mark.parametrize("param", [c[0] for c in C.SISE3_SINGLE])
def test_sise3_one_transition_at_a_time(oracle, param):
    n = 4000
    m = C.sise3_single_model(param, n)
    for solver in ("ssm", "aem"):
        r = oracle.run(m, solver=solver, seed=7)
        assert r.err == 0
        assert min(C.sise3_single_ps(r.U, r.V, n, param)) > 1e-4, (param, solver)


def test_sise3_age_classes_are_sis_chains(oracle):
    n = 20000
    m = C.sise3_chain_model(n)
    for solver in ("ssm", "aem"):
        r = oracle.run(m, solver=solver, seed=9)
        assert r.err == 0
        assert min(C.sise3_chain_ps(r.U, n)) > 1e-4, solver


def test_decay_interval_table_is_a_partition():
    for ends in C.PHI_CASES:
        ks = [C.decay_interval(d, ends) for d in range(365)]
        assert set(ks) == {1, 2, 3, 4}
        # Interval k ends at end_tk (non-inclusive).
        for k in range(4):
            e = ends[k]
            assert ks[e - 1] == k + 1 and ks[e % 365] != k + 1


@pytest.mark.parametrize("ends", C.PHI_CASES)
def test_sise3_phi_recurrence(oracle, ends):
    m = C.phi_model(ends)
    exp = C.phi_expected(ends)
    for solver in ("ssm", "aem"):
        r = oracle.run(m, solver=solver, seed=1)
        assert r.err == 0
        np.testing.assert_allclose(r.V, exp, rtol=1e-13, atol=0)
        # u never changes.
        assert (r.U == m.u0[:, :].ravel(order="F")[:, None]).all()


def test_phi_recurrence_catches_a_dropped_class():
    """A recurrence that leaves I_3 out of the infected sum differs."""
    ends = C.PHI_CASES[0]
    exp = C.phi_expected(ends)
    u = C.PHI_U[0]
    I, n = u[1] + u[3], sum(u)
    phi, wrong = 0.5, []
    beta = [None, 0.1, 0.2, 0.3, 0.4]
    for j in range(400):
        phi = phi * (1.0 - beta[C.decay_interval(j % 365, ends)]) + (0.7 * I / n + 0.01)
        wrong.append(phi)
    assert not np.allclose(exp[0], wrong, rtol=1e-6)


@pytest.mark.parametrize("case", C.error_cases(), ids=lambda c: c[0])
def test_reference_event_and_pts_errors(oracle, case):
    name, build, code = case
    for solver in ("ssm", "aem"):
        r = oracle.run(build(), solver=solver, seed=3)
        assert r.err == code, (name, solver, r.err)


def test_enter_shift_known_answer(oracle):
    m, U = C.enter_shift_known_answer()
    for solver in ("ssm", "aem"):
        r = oracle.run(m, solver=solver, seed=3)
        assert r.err == 0
        np.testing.assert_array_equal(r.U, U)
