"""GPU parity of plans whose replicates have parameters of their own
(Plan.set_replicate_data, the layout batched ABC-SMC proposals run in):
node kernels, the one-workgroup-per-replicate coupled kernel and the
grid-wide coupled kernel, bit for bit against the checker given the same
per-replicate ldata / gdata; and a batch whose plan gets them after the
graph was captured."""
import numpy as np
import pytest

import siminf_amd as si
from siminf_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run(model, seed, ldata=None, gdata=None, solver="ssm"):
    with si.Plan(model, solver=solver) as p:
        p.set_replicate_data(ldata=ldata, gdata=gdata)
        tr, _, _ = p.run(seed)
        m = p.download()
        return m.U.copy(), None if m.V is None else m.V.copy(), tr


def _check(oracle, model, seed, ldata=None, gdata=None):
    print("case: %s Nn=%d Nrep=%d per-replicate %s" % (
        getattr(model.plugin, "name", model.plugin), model.Nn, model.replicates,
        "ldata" if ldata is not None else "gdata"), flush=True)
    U, V, tr = _run(model, seed, ldata, gdata)
    ref = oracle.run(model, seed=seed, rep_ldata=ldata, rep_gdata=gdata)
    assert ref.err == 0
    np.testing.assert_array_equal(U, ref.U)
    if V is not None and V.size:
        np.testing.assert_array_equal(V, ref.V)
    assert tr == ref.transitions
    return U


def test_mparse_sir_gdata_per_replicate(oracle):
    m = W.mparse_sir_readme(Nn=30, days=60, replicates=50)
    rng = np.random.default_rng(1)
    g = np.vstack([rng.uniform(0.08, 0.24, 50), rng.uniform(0.04, 0.12, 50)])
    U = _check(oracle, m, 4, gdata=g)
    # Replicates differ in their parameters, so in their outbreaks.
    # U rows node * 3 + compartment, columns replicate * 60 + day.
    R_end = U.reshape(3, 30, 60, 50, order="F")[2, :, -1, :].sum(axis=0)
    assert len(set(R_end.tolist())) > 10


def test_seir_ldata_per_replicate(oracle):
    m = W.seir_million(Nn=2000, days=90, seed=3, seeded_frac=0.05)
    base = m.ldata
    R = 6
    rep = si.SEIR(u0={c: m.u0[k] for k, c in enumerate("SEIR")}, tspan=m.tspan,
                  beta=base[0], epsilon=base[1], gamma=base[2], replicates=R)
    scale = np.repeat(np.linspace(0.5, 1.5, R), m.Nn)
    ld = np.tile(base, (1, R)) * scale[None, :]
    U = _check(oracle, rep, 7, ldata=ld)
    # The same values as one column per replicate.
    U1 = _check(oracle, rep, 7, ldata=ld[:, ::m.Nn].copy())
    np.testing.assert_array_equal(U, U1)


def test_sise3_gdata_per_replicate_coupled(oracle):
    rng = np.random.default_rng(2)
    for Nn, R in ((120, 6), (600, 3)):
        m = W.sise3_herds(Nn=Nn, days=60, seed=8, replicates=R)
        g = np.tile(m.gdata.reshape(-1, 1), (1, R))
        g[0:3] *= rng.uniform(0.5, 1.5, (3, R))
        g[3:6] *= rng.uniform(0.5, 1.5, (3, R))
        _check(oracle, m, 9, gdata=g)


def test_batch_recaptures_after_parameters_move():
    """A batch captured before its plan's parameters became per-replicate
    launches with them afterwards (the graph is captured again)."""
    m = W.mparse_sir_readme(Nn=50, days=40, replicates=20)
    rng = np.random.default_rng(5)
    g = np.vstack([rng.uniform(0.08, 0.24, 20), rng.uniform(0.04, 0.12, 20)])
    want = _run(m, 11, gdata=g)[0]
    with si.Plan(m) as p:
        with si.Batch([p]) as b:
            b.set_graph(True)
            b.run([11])
            p.set_replicate_data(gdata=g)
            b.run([11])
            np.testing.assert_array_equal(p.download().U, want)
