"""GPU parity: the HIP solver core (through the C ABI) against the CPU
checker on identical SimInf_solver_args and seeds.  Integer trajectories
and transition counts must match bit for bit; V (double) must match bit
for bit too, since both sides run the same IEEE operation sequence."""
import ctypes
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import siminf_amd as si
from siminf_amd import _native, workloads as W
from siminf_amd.model import device_plugin

pytestmark = pytest.mark.gpu


def _gpu(model, solver="ssm", seed=0, crn=False):
    lib = _native.load()
    dev = device_plugin(model, lib)
    assert dev
    b = _native.ArgsBundle(model, solver=0 if solver == "ssm" else 1,
                           crn=int(crn), seed=seed)
    plan = ctypes.c_void_p()
    err = lib.SimInf_hip_plan_create(ctypes.byref(plan), ctypes.byref(b.args),
                                     ctypes.c_void_p(dev))
    assert err == 0, err
    try:
        err = lib.SimInf_hip_plan_run(plan, seed)
        cp = lambda a, t: _native.ptr(a, t)
        if err == 0:
            assert lib.SimInf_hip_plan_download(
                plan, cp(b.U, ctypes.c_int), cp(b.V, ctypes.c_double),
                cp(b.prU, ctypes.c_double), cp(b.prV, ctypes.c_double)) == 0
        tr = ctypes.c_longlong()
        nd = ctypes.c_longlong()
        ms = ctypes.c_double()
        lib.SimInf_hip_plan_stats(plan, ctypes.byref(tr), ctypes.byref(nd),
                                  ctypes.byref(ms))
    finally:
        lib.SimInf_hip_plan_destroy(plan)
    return err, b, tr.value


def _check(model, oracle, solver="ssm", seed=1, crn=False):
    print("case: %s %s Nn=%d Nrep=%d tlen=%d events=%d seed=%d" % (
        getattr(model.plugin, "name", model.plugin), solver, model.Nn, model.replicates, len(model.tspan),
        len(model.events["event"]), seed), flush=True)
    ref = oracle.run(model, solver=solver, seed=seed, crn=crn)
    err, b, transitions = _gpu(model, solver=solver, seed=seed, crn=crn)
    assert err == ref.err
    if err == 0:
        if b.U is not None:
            np.testing.assert_array_equal(b.U, ref.U)
        else:
            np.testing.assert_array_equal(b.prU, ref.prU)
        if b.V is not None:
            np.testing.assert_array_equal(b.V, ref.V)
        else:
            np.testing.assert_array_equal(b.prV, ref.prV)
        assert transitions == ref.transitions
    return ref, b


def test_mparse_sir_readme_full(oracle):
    """The README's mparse SIR at full size: 1000 nodes x 150 days."""
    m = W.mparse_sir_readme()
    ref, b = _check(m, oracle, seed=123)
    assert ref.transitions > 100000


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_seir_ssm(oracle, seed):
    m = W.seir_million(Nn=20000, days=120, seed=seed, seeded_frac=0.05)
    _check(m, oracle, solver="ssm", seed=seed)


@pytest.mark.parametrize("seed", [1, 2])
def test_seir_aem(oracle, seed):
    m = W.seir_million(Nn=5000, days=100, seed=seed, seeded_frac=0.05)
    ref, _ = _check(m, oracle, solver="aem", seed=seed)
    assert ref.aem_ties == 0


def test_sir_builtin_aem_and_ssm(oracle):
    m = si.SIR(u0={"S": np.arange(50, 250), "I": 3, "R": 0},
               tspan=np.linspace(0, 60, 31), beta=0.3, gamma=0.1)
    _check(m, oracle, solver="ssm", seed=9)
    _check(m, oracle, solver="aem", seed=9)


def test_mssm_replicates(oracle):
    m = W.mparse_sir_readme(Nn=30, days=60, replicates=200)
    _check(m, oracle, solver="ssm", seed=5)
    _check(m, oracle, solver="ssm", seed=5, crn=True)


def test_sise3_e1_events(oracle):
    m = W.sise3_herds(Nn=300, days=200, seed=3, external_rate=0.0)
    assert (m.events["event"] == 3).sum() == 0
    _check(m, oracle, seed=4)


def test_sise3_e2_events(oracle):
    m = W.sise3_herds(Nn=300, days=200, seed=5)
    assert (m.events["event"] == 3).sum() > 0
    _check(m, oracle, seed=6)


@pytest.mark.parametrize("seed", [6, 11])
def test_sise3_e2_events_aem(oracle, seed):
    """aem with external transfers: the all-events method in the coupled
    kernel (E1, E2, PTS refresh through calcTimes at the next day)."""
    m = W.sise3_herds(Nn=300, days=200, seed=5)
    ref, _ = _check(m, oracle, solver="aem", seed=seed)
    assert ref.err == 0 and ref.transitions > 0
    assert ref.aem_ties == 0


def test_sise3_1600_herds_full_aem(oracle):
    """SISe3, 1600 herds x 1460 days with ~1.2M scheduled events, under
    aem."""
    m = W.sise3_herds()
    ref, _ = _check(m, oracle, solver="aem", seed=5)
    assert ref.transitions > 0
    assert ref.aem_ties == 0


def test_sise3_e2_events_replicates(oracle):
    m = W.sise3_herds(Nn=120, days=100, seed=8, replicates=6)
    _check(m, oracle, seed=7)


def test_sise3_1600_herds_full(oracle):
    """SISe3 at full size: 1600 herds, 1460 days, ~1.2M events."""
    m = W.sise3_herds()
    ref, _ = _check(m, oracle, seed=5)
    assert ref.transitions > 0


def test_sparse_output(oracle):
    m = W.mparse_sir_readme(Nn=50, days=40)
    rows = np.array([0, 1, 5, 7, 8, 149])
    cols = np.array([0, 3, 3, 10, 39, 39])
    m.U_sparse = sp.csc_matrix((np.ones(len(rows)), (rows, cols)),
                               shape=(m.Nn * m.Nc, len(m.tspan)))
    _check(m, oracle, seed=2)


def test_sise_phi_sparse_v(oracle):
    m = si.SISe(u0={"S": [100], "I": [0]}, tspan=np.arange(1, 701, 7.0), phi=1.0,
                upsilon=0.0, gamma=0.0, alpha=0.0, beta_t1=0.19, beta_t2=0.085,
                beta_t3=0.075, beta_t4=0.185, end_t1=91, end_t2=182, end_t3=273,
                end_t4=365, epsilon=0.000011)
    m.V_sparse = sp.csc_matrix(np.ones((1, len(m.tspan))))
    _check(m, oracle, seed=1)


def test_seir_enter_events_known_answer():
    """tests/SEIR.R:231-295: one individual enters every compartment of
    every node every day; zero rates."""
    from tests.golden_data import seir_enter_model, seir_enter_expected
    m = seir_enter_model()
    si.run(m, seed=1)
    np.testing.assert_array_equal(m.U, seir_enter_expected())


def test_errors_match(oracle):
    # Invalid rate at initialisation (tests/SIR.R:483-492).
    m = si.SIR(u0={"S": [99, 99], "I": 1, "R": 0}, tspan=np.arange(1, 101.0),
               beta=0.16, gamma=-0.077)
    for solver in ("ssm", "aem"):
        ref, _ = _check(m, oracle, solver=solver, seed=1)
        assert ref.err == -11
    # Negative state from a corrupted S (tests/SIR.R:554-567).
    m = si.SIR(u0={"S": [99], "I": 1, "R": 0}, tspan=np.arange(1, 101.0),
               beta=0.16, gamma=0.0)
    S = m.S.tocsc()
    S.data[0] = -1000
    m.S = S
    for solver in ("ssm", "aem"):
        ref, _ = _check(m, oracle, solver=solver, seed=1)
        assert ref.err == -1
    # Sampling more individuals than exist (tests/sample_select.R:42-80).
    m = si.SISe3(u0={"S_1": [0, 1], "I_1": 0, "S_2": 0, "I_2": 0, "S_3": 0, "I_3": 0},
                 tspan=np.arange(0, 11.0),
                 events={"event": [3], "time": [1], "node": [2], "dest": [1],
                         "n": [2], "proportion": [1.0], "select": [4], "shift": [0]},
                 phi=0.0, upsilon_1=0, upsilon_2=0, upsilon_3=0, gamma_1=1,
                 gamma_2=1, gamma_3=1, alpha=0, beta_t1=1, beta_t2=1, beta_t3=1,
                 beta_t4=1, end_t1=91, end_t2=182, end_t3=273, end_t4=365,
                 epsilon=0)
    ref, _ = _check(m, oracle, seed=1)
    assert ref.err == -8


def test_run_api_raises_reference_message():
    m = si.SIR(u0={"S": [99, 99], "I": 1, "R": 0}, tspan=np.arange(1, 101.0),
               beta=0.16, gamma=-0.077)
    with pytest.raises(si.SimInfError, match=r"Invalid rate detected \(non-finite or < 0\.0\)\."):
        si.run(m, solver="ssm", seed=1)


def test_noncanonical_tables(oracle):
    """S with an explicitly stored zero and G with extra (harmless)
    dependencies select the kernels that read S and G from the arguments;
    the trajectory must equal both the checker's and the canonical run's."""
    m = W.sise3_herds(Nn=200, days=120, seed=9)
    ref_canon, b_canon = _check(m, oracle, seed=3)
    S = m.S.toarray()
    G = m.G.toarray()
    G[0, 2] = 1.0          # rate 1 recomputed after transition 3 fires
    G[5, 0] = 1.0
    S_sp = sp.csc_matrix(S)
    # store an explicit zero in column 0, row 2
    S_sp = sp.csc_matrix((np.append(S_sp.data, 0.0),
                          (np.append(S_sp.indices, 2),
                           np.append(np.repeat(np.arange(6), np.diff(S_sp.indptr)), 0))),
                         shape=S.shape)
    S_sp.sum_duplicates()
    S_sp.sort_indices()
    m.S = S_sp
    m.G = sp.csc_matrix(G)
    ref, b = _check(m, oracle, seed=3)
    np.testing.assert_array_equal(b.U, b_canon.U)
    np.testing.assert_array_equal(b.V, b_canon.V)


def test_generated_sir_plugin(oracle):
    """The README model through the parser: the device plugin compiled
    from its generated C gives the checker's trajectories (which run the
    same C on the host) and those of the hand-written plugin."""
    gen = W.mparse_sir_generated(Nn=300, days=80, replicates=2)
    _, b_gen = _check(gen, oracle, seed=21)
    hand = W.mparse_sir_readme(Nn=300, days=80, replicates=2)
    _, b_hand = _check(hand, oracle, seed=21)
    np.testing.assert_array_equal(b_gen.U, b_hand.U)


@pytest.mark.parametrize("solver", ["ssm", "aem"])
def test_generated_pts_plugin(oracle, solver):
    """Generated model with v0, ldata, enum constants and a post time step
    that writes v_new: U and V bit-exact against the checker."""
    m = W.mparse_sis_env(Nn=400, days=150)
    ref, _ = _check(m, oracle, solver=solver, seed=8)
    assert ref.transitions > 0 and ref.V.max() > 0
