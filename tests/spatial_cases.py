"""Spatial-model (SISe_sp, SISe3_sp) cases shared by the CPU pins
(tests/test_spatial.py) and the GPU parity tests
(tests/test_gpu_spatial.py)."""
import numpy as np
import scipy.sparse as sp

import siminf_amd as si
from siminf_amd import workloads

# The parameters of the reference's SISe_sp tests (tests/SISe_sp.R:990-1007).
SISE_SP = dict(upsilon=0.0357, gamma=0.1, alpha=1.0, beta_t1=0.19,
               beta_t2=0.085, beta_t3=0.075, beta_t4=0.185, end_t1=91,
               end_t2=182, end_t3=273, end_t4=365)


def scattered(n, seed):
    """Points on a grid of spacing 0.8 moved by up to 0.3 along each axis
    (no two closer than 0.2), and their distance matrix with cutoff 1:
    two to eight neighbours per point."""
    rng = np.random.default_rng(seed)
    k = int(np.ceil(np.sqrt(n)))
    gx, gy = np.meshgrid(np.arange(k), np.arange(k))
    x = 0.8 * gx.ravel()[:n] + rng.uniform(-0.3, 0.3, n)
    y = 0.8 * gy.ravel()[:n] + rng.uniform(-0.3, 0.3, n)
    return si.distance_matrix(x, y, 1.0)


def sise_sp_model(n=120, days=200, seed=3, coupling=0.01, replicates=1,
                  events=False, **over):
    """SISe_sp on scattered herds of 0-59 animals, a fifth of them with
    infected animals, phi drawn in [0, 0.5); with events, enter events
    and external transfers between random herds."""
    rng = np.random.default_rng(seed)
    S = rng.integers(0, 60, n)
    I = np.where(rng.random(n) < 0.2, rng.integers(1, 6, n), 0)
    S[0] = 0
    I[0] = 0           # one empty herd: no local spread into or out of it
    kw = dict(SISE_SP)
    kw.update(over)
    ev = None
    if events:
        ev = _events(rng, n, days, select_all=2)
    return si.SISe_sp({"S": S, "I": I}, np.arange(1, days + 1, dtype=float),
                      events=ev, phi=rng.uniform(0, 0.5, n), coupling=coupling,
                      distance=scattered(n, seed + 1), replicates=replicates, **kw)


def _events(rng, n, days, select_all):
    """Enter events of 3 animals into compartment 1 and external
    transfers of a binomial fifth of a herd's animals (E column
    select_all), always feasible."""
    ev = {"event": [], "time": [], "node": [], "dest": [], "n": [],
          "proportion": [], "select": [], "shift": []}
    for d in range(1, days + 1):
        for node in rng.choice(n, max(1, n // 20), replace=False):
            ev["event"].append(1)
            ev["time"].append(d)
            ev["node"].append(node + 1)
            ev["dest"].append(0)
            ev["n"].append(3)
            ev["proportion"].append(0.0)
            ev["select"].append(1)
            ev["shift"].append(0)
        if d > 1:
            for src in rng.choice(n, max(1, n // 25), replace=False):
                dst = int(rng.integers(0, n))
                if dst == src:
                    continue
                ev["event"].append(3)
                ev["time"].append(d)
                ev["node"].append(src + 1)
                ev["dest"].append(dst + 1)
                ev["n"].append(0)
                ev["proportion"].append(0.2)
                ev["select"].append(select_all)
                ev["shift"].append(0)
    return {k: np.asarray(v) for k, v in ev.items()}


def sise3_sp_model(n=300, days=365, seed=11, events=True, coupling=0.0005,
                   replicates=1):
    """The SISe3 herd workload with local spread (workloads.sise3_sp_herds)."""
    return workloads.sise3_sp_herds(Nn=n, days=days, seed=seed, events=events,
                                    coupling=coupling, replicates=replicates)


def frozen_sise_sp(N, phi0, distance, coupling, days, I=None):
    """SISe_sp whose transitions never fire (upsilon = gamma = 0) and whose
    phi neither decays (beta = 0) nor grows from infected animals
    (alpha = 0): phi then moves by the local spread alone."""
    n = len(N)
    I = np.zeros(n, int) if I is None else np.asarray(I)
    kw = dict(SISE_SP, upsilon=0.0, gamma=0.0, alpha=0.0, beta_t1=0.0,
              beta_t2=0.0, beta_t3=0.0, beta_t4=0.0)
    return si.SISe_sp({"S": np.asarray(N) - I, "I": I},
                      np.arange(1, days + 1, dtype=float), phi=phi0,
                      coupling=coupling, distance=distance, **kw)


def spread_by_matrix(N, phi0, distance, coupling, days):
    """phi of frozen_sise_sp after each day, from the linear map in matrix
    form: phi_i <- phi_i + D / N_i * sum_j W_ij (phi_j N_j - phi_i N_i),
    W_ij = 1 / d_ij over node i's neighbours with N_j > 0, nodes with
    N_i = 0 unchanged.  (Nn x days)"""
    N = np.asarray(N, dtype=float)
    W = sp.csc_matrix(distance).T.toarray()        # row i: node i's neighbours
    W = np.where(W > 0, 1.0 / np.where(W > 0, W, 1.0), 0.0)
    W[:, N <= 0] = 0.0
    live = N > 0
    phi = np.asarray(phi0, dtype=float).copy()
    out = np.empty((len(N), days))
    for t in range(days):
        x = phi * N
        flow = W @ x - W.sum(axis=1) * x
        step = np.zeros_like(phi)
        step[live] = coupling * flow[live] / N[live]
        phi = phi + step
        out[:, t] = phi
    return out
