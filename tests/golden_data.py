"""Models and expected outputs built from the fixtures in tests/golden/."""
import json
import os

import numpy as np

import siminf_amd as si

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def seir_enter_model():
    """tests/SEIR.R:231-254: SEIR, 6 nodes, tspan 0:9, zero rates, one
    enter event per (time 1..9, node, compartment) with E = diag(4) and
    select cycling 1..4."""
    rows = []
    for time in range(1, 10):
        for node in range(1, 7):
            for _ in range(4):
                rows.append((1, time, node, 0, 1, 0.0, 1, 0))
    ev = {k: np.array([r[i] for r in rows]) for i, k in enumerate(
        ("event", "time", "node", "dest", "n", "proportion", "select", "shift"))}
    ev["select"] = np.resize(np.arange(1, 5), len(rows))
    u0 = {"S": [110, 210, 310, 410, 510, 610], "E": [120, 220, 320, 420, 520, 620],
          "I": [130, 230, 330, 430, 530, 630], "R": [140, 240, 340, 440, 540, 640]}
    m = si.SEIR(u0=u0, tspan=np.arange(0, 10.0), events=None, beta=0, epsilon=0,
                gamma=0)
    return si.SimInfModel("SEIR", m.G, m.S, m.tspan, m.u0, ldata=m.ldata,
                          E=np.eye(4), events=ev, compartments=m.compartments)


def seir_enter_expected():
    """U (Nn*Nc x tlen) from the data.frame of tests/SEIR.R:256-292."""
    d = load("seir_enter_events.json")
    U = np.zeros((6 * 4, 10), dtype=np.int32)
    for k in range(len(d["node"])):
        node, t = d["node"][k] - 1, d["time"][k]
        for c, name in enumerate("SEIR"):
            U[node * 4 + c, t] = d[name][k]
    return U


def seir_zero_rates_model():
    """tests/SEIR.R:134-141 (u0 of tests/SEIR.R:27-33 has S = 0..5)."""
    u0 = {"S": [0, 1, 2, 3, 4, 5], "E": [0] * 6, "I": [0] * 6, "R": [0] * 6}
    return si.SEIR(u0=u0, tspan=np.arange(0, 10.0), beta=0, epsilon=0, gamma=0)


def sise_phi_model(sparse_v=False):
    """tests/phi.R:86-101."""
    m = si.SISe(u0={"S": [100], "I": [0]}, tspan=np.arange(1, 701, 7.0),
                phi=1.0, upsilon=0.0, gamma=0.0, alpha=0.0, beta_t1=0.19,
                beta_t2=0.085, beta_t3=0.075, beta_t4=0.185, end_t1=91,
                end_t2=182, end_t3=273, end_t4=365, epsilon=0.000011)
    if sparse_v:
        import scipy.sparse as sp
        m.V_sparse = sp.csc_matrix(np.ones((1, len(m.tspan))))
    return m
