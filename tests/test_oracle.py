"""CPU tests that pin the checker (oracle/siminf_oracle.c) before it is
trusted: known-answer vectors, the reference's RNG-free golden outputs,
closed-form distributions, brute force on tiny chains and invariants."""
import json
import os

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp
from scipy import stats

import siminf_amd as si
from siminf_amd import workloads as W
from tests import golden_data as G

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _chisq_p(obs, pmf, n, min_exp=10):
    """Chi-square p-value with the bins of expected count < min_exp
    pooled into one (dropped when the pool is empty)."""
    obs = np.asarray(obs, float)
    e = np.asarray(pmf, float) * n
    keep = e >= min_exp
    o = list(obs[keep])
    ex = list(e[keep])
    if e[~keep].sum() > 0 or obs[~keep].sum() > 0:
        o.append(obs[~keep].sum())
        ex.append(e[~keep].sum())
    o, ex = np.array(o), np.array(ex)
    if ex[-1] < 5 and len(ex) > 1:
        o[-2] += o[-1]; ex[-2] += ex[-1]; o, ex = o[:-1], ex[:-1]
    ex *= o.sum() / ex.sum()
    return stats.chisquare(o, ex).pvalue


# ---------------------------------------------------------------- RNG ---

def test_philox_known_answers(oracle):
    kat = json.load(open(os.path.join(GOLDEN, "philox_kat.json")))["vectors"]
    assert len(kat) >= 64
    for v in kat:
        assert oracle.philox(v["ctr"], v["key"]) == v["out"]
    # The Random123 paper's published vectors are among them.
    assert kat[0]["out"] == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert kat[2]["out"] == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_stream_words_map_to_open_unit_interval(oracle):
    """Uniform 2b and 2b+1 of a stream come from words (0,1) and (2,3) of
    Philox block b with counter (b, 0, id, domain << 28)."""
    seed, domain, sid = 0x123456789ABCDEF, 1, 77
    u = oracle.uniforms(seed, domain, sid, 6)
    for b in range(3):
        w = oracle.philox([b, 0, sid, domain << 28],
                          [seed & 0xFFFFFFFF, seed >> 32])
        for h in range(2):
            x = (w[2 * h + 1] << 32) | w[2 * h]
            assert u[2 * b + h] == ((x >> 12) + 0.5) * 2.0 ** -52
    big = oracle.uniforms(5, 0, 3, 200000)
    assert big.min() > 0.0 and big.max() < 1.0
    assert stats.kstest(big, "uniform").pvalue > 1e-3


def test_streams_are_distinct(oracle):
    a = oracle.uniforms(1, 0, 0, 1000)
    b = oracle.uniforms(1, 0, 1, 1000)
    c = oracle.uniforms(1, 1, 0, 1000)
    d = oracle.uniforms(2, 0, 0, 1000)
    assert len({a[0], b[0], c[0], d[0]}) == 4
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.15


def test_log_accuracy(oracle):
    import mpmath
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.random(500000), np.exp(rng.uniform(-740, 709, 200000)),
                        1 + rng.uniform(-1e-7, 1e-7, 50000),
                        [1.0, 2.0, 0.5, 5e-324, 1e-310, np.nextafter(1, 0)]])
    y = oracle.log(x)
    z = np.log(x)
    assert np.max(np.abs(y - z) / np.spacing(np.abs(z) + 1e-300)) <= 2.0
    mpmath.mp.prec = 120
    for xi, yi in zip(x[::997], y[::997]):
        t = mpmath.log(mpmath.mpf(float(xi)))
        if t != 0:
            err = abs((mpmath.mpf(float(yi)) - t) / mpmath.mpf(float(np.spacing(abs(float(t))))))
            assert err <= 2.0
    assert oracle.log(np.array([1.0]))[0] == 0.0
    special = oracle.log(np.array([0.0, np.inf, -1.0, np.nan]))
    assert special[0] == -np.inf and special[1] == np.inf
    assert np.isnan(special[2]) and np.isnan(special[3])


def test_sqrt_accuracy(oracle):
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.random(200000) * 1e6, np.exp(rng.uniform(-740, 709, 100000))])
    s = oracle.sqrt(x)
    assert np.max(np.abs(s - np.sqrt(x)) / np.spacing(np.sqrt(x))) <= 1.0


@pytest.mark.parametrize("n,p", [(10, 0.3), (40, 0.05), (100, 0.45), (1000, 0.4),
                                 (5000, 0.7), (60, 0.98), (200000, 0.5), (3, 0.5)])
def test_binomial_distribution(oracle, n, p):
    x = oracle.binomial(17 + n, p, n, 60000)
    assert x.min() >= 0 and x.max() <= n
    mean, var = n * p, n * p * (1 - p)
    assert abs(x.mean() - mean) < 5 * np.sqrt(var / len(x)) + 1e-9
    k = np.arange(n + 1)
    pmf = stats.binom.pmf(k, n, p)
    obs = np.bincount(x, minlength=n + 1)
    assert _chisq_p(obs, pmf, len(x), 20) > 1e-4


def test_binomial_edges(oracle):
    assert (oracle.binomial(1, 0.0, 50, 100) == 0).all()
    assert (oracle.binomial(1, 1.0, 50, 100) == 50).all()
    assert (oracle.binomial(1, 0.3, 0, 100) == 0).all()


@pytest.mark.parametrize("n1,n2,t", [(5, 7, 4), (30, 70, 10), (30, 70, 80),
                                     (200, 50, 120), (1, 99, 50), (10, 0, 5)])
def test_hypergeometric_distribution(oracle, n1, n2, t):
    x = oracle.hypergeometric(3 + t, n1, n2, t, 50000)
    k = np.arange(0, min(n1, t) + 1)
    pmf = stats.hypergeom.pmf(k, n1 + n2, n1, t)
    obs = np.bincount(x, minlength=len(k))[:len(k)]
    assert obs.sum() == len(x)
    if (pmf * len(x) >= 10).sum() <= 1:
        assert np.all(x == k[np.argmax(pmf)])
        return
    assert _chisq_p(obs, pmf, len(x)) > 1e-4


# ------------------------------------------------- sample_select ---

def _E(cols, Nc):
    M = np.zeros((Nc, len(cols)))
    for j, c in enumerate(cols):
        for r, w in c.items():
            M[r, j] = w
    return sp.csc_matrix(M)


def test_sample_select_deterministic_branches(oracle):
    E = _E([{0: 1}, {0: 1, 1: 1, 2: 1}, {1: 1, 2: 2}], 3)
    u = np.array([4, 0, 6])
    # n == Nind: everything (SimInf_solver.c:111-115).
    err, out = oracle.sample(E, u, 1, 10, 0.0, 1, 3)
    assert err == 0 and (out == [4, 0, 6]).all()
    # One state in the column (:116-119).
    err, out = oracle.sample(E, u, 0, 3, 0.0, 1, 3)
    assert err == 0 and (out == [3, 0, 0]).all()
    # One state with individuals (:120-128).
    err, out = oracle.sample(E, np.array([0, 0, 9]), 1, 5, 0.0, 1, 3)
    assert err == 0 and (out == [0, 0, 5]).all()
    # Proportion 0 and 1 (tests/sample_select.R:321-464).
    err, out = oracle.sample(E, u, 1, 0, 0.0, 1, 3)
    assert err == 0 and (out == 0).all()
    err, out = oracle.sample(E, u, 1, 0, 1.0, 1, 3)
    assert err == 0 and (out == [4, 0, 6]).all()
    # Errors (:20-160): too many, negative n, bad proportion.
    assert oracle.sample(E, u, 1, 11, 0.0, 1, 1)[0] == -8
    assert oracle.sample(E, u, 1, -1, 0.0, 1, 1)[0] == -8
    assert oracle.sample(E, u, 1, 0, 1.5, 1, 1)[0] == -18
    assert oracle.sample(E, u, 1, 0, -0.1, 1, 1)[0] == -18


@pytest.mark.parametrize("npre", [0, 2, 8, 12, 32])
def test_precomputed_uniforms_change_nothing(oracle, npre):
    """The device computes an external transfer's first uniforms ahead in
    parallel and the urn then runs over them without the stream's
    branches; variates and stream positions must be those of a stream
    that computes every uniform itself."""
    assert oracle.prefetch_mismatches(77, npre, 20000) == 0


def test_sample_select_equal_weights_is_hypergeometric(oracle):
    E = _E([{0: 1, 1: 1, 2: 1}], 3)
    u = np.array([5, 8, 7])
    err, out = oracle.sample(E, u, 0, 6, 0.0, 9, 40000)
    assert err == 0 and (out.sum(1) == 6).all()
    # Marginal of each compartment is hypergeometric(u_k, N - u_k, n).
    for c in range(3):
        k = np.arange(0, 7)
        pmf = stats.hypergeom.pmf(k, 20, u[c], 6)
        obs = np.bincount(out[:, c], minlength=7)
        assert _chisq_p(obs, pmf, len(out)) > 1e-4


def _wallenius_exact(u, w, n):
    """Exact distribution of the biased urn by dynamic programming."""
    from functools import lru_cache

    @lru_cache(None)
    def rec(taken, left):
        if left == 0:
            return {taken: 1.0}
        weights = [w[i] * (u[i] - taken[i]) for i in range(len(u))]
        tot = sum(weights)
        out = {}
        for i, wi in enumerate(weights):
            if wi > 0:
                t = list(taken); t[i] += 1
                for k, v in rec(tuple(t), left - 1).items():
                    out[k] = out.get(k, 0.0) + v * wi / tot
        return out
    return rec(tuple([0] * len(u)), n)


def test_sample_select_unequal_weights_is_wallenius(oracle):
    E = _E([{0: 1.0, 1: 3.0, 2: 0.5}], 3)
    u = np.array([4, 3, 5])
    n = 4
    err, out = oracle.sample(E, u, 0, n, 0.0, 21, 40000)
    assert err == 0
    exact = _wallenius_exact(tuple(u), (1.0, 3.0, 0.5), n)
    keys = sorted(exact)
    counts = {}
    for row in map(tuple, out):
        counts[row] = counts.get(row, 0) + 1
    assert set(counts) <= set(keys)
    o = np.array([counts.get(k, 0) for k in keys], float)
    assert _chisq_p(o, [exact[k] for k in keys], len(out)) > 1e-4


def test_sample_select_enter_is_multinomial(oracle):
    E = _E([{0: 1.0, 2: 3.0}], 3)
    err, out = oracle.sample(E, np.zeros(3, int), 0, 5, 0.0, 4, 30000, enter=True)
    assert err == 0 and (out[:, 1] == 0).all() and (out.sum(1) == 5).all()
    k = np.arange(6)
    pmf = stats.binom.pmf(k, 5, 0.25)
    obs = np.bincount(out[:, 0], minlength=6)
    assert _chisq_p(obs, pmf, len(out)) > 1e-4
    # A single row takes all; proportion draws from the selected total.
    err, out = oracle.sample(_E([{1: 1.0}], 3), np.array([0, 40, 0]), 0, 0, 0.5,
                             4, 2000, enter=True)
    assert err == 0 and abs(out[:, 1].mean() - 20) < 1.0


# --------------------------------------------- reference golden runs ---

def test_seir_zero_rates_golden(oracle):
    """tests/SEIR.R:134-171."""
    m = G.seir_zero_rates_model()
    r = oracle.run(m, seed=1)
    assert r.err == 0
    z = G.load("seir_zero_rates.json")
    for c, name in ((0, "S"), (2, "I"), (3, "R")):
        got = r.U[c::4, :]
        exp = np.array(z[name]).reshape(10, 6).T
        np.testing.assert_array_equal(got, exp)


def test_seir_enter_events_golden(oracle):
    """tests/SEIR.R:231-295."""
    r = oracle.run(G.seir_enter_model(), seed=1)
    assert r.err == 0
    np.testing.assert_array_equal(r.U, G.seir_enter_expected())


@pytest.mark.parametrize("sparse_v", [False, True])
def test_phi_golden(oracle, sparse_v):
    """tests/phi.R:86-117: phi of SISe with upsilon = gamma = alpha = 0."""
    m = G.sise_phi_model(sparse_v)
    r = oracle.run(m, seed=1)
    assert r.err == 0
    d = G.load("phi_sise.json")
    phi = r.V[0, :] if r.V is not None else r.prV
    assert np.all(np.abs(phi - np.array(d["phi"])) < d["tol"])


# ------------------------------------------------ closed forms ---

def test_pure_death_is_binomial(oracle):
    """With beta = 0, I(t) of SIR is Binomial(I0, exp(-gamma (t - t0)))."""
    Nn, I0, gamma = 4000, 20, 0.1
    m = si.SIR(u0={"S": np.full(Nn, 10), "I": np.full(Nn, I0), "R": 0},
               tspan=np.array([0.0, 5.0, 10.0]), beta=0.0, gamma=gamma)
    for solver in ("ssm", "aem"):
        r = oracle.run(m, solver=solver, seed=3)
        assert r.err == 0
        for j, t in enumerate([5.0, 10.0]):
            # Column j of U holds the state after day floor(tspan[j]).
            I = r.U[1::3, j + 1]
            p = np.exp(-gamma * (np.floor(t) + 1))
            k = np.arange(I0 + 1)
            pmf = stats.binom.pmf(k, I0, p)
            obs = np.bincount(I, minlength=I0 + 1)
            assert _chisq_p(obs, pmf, Nn) > 1e-4, (solver, t)


def test_sir_ctmc_brute_force(oracle):
    """One SIR node with N = 6: the law of (S, I) after one day from the
    matrix exponential of the generator, against 30000 replicates."""
    beta, gamma, N = 0.9, 0.4, 6
    states = [(s, i) for s in range(N + 1) for i in range(N + 1 - s)]
    index = {x: k for k, x in enumerate(states)}
    Q = np.zeros((len(states), len(states)))
    for (s, i), k in index.items():
        r1 = beta * s * i / N
        r2 = gamma * i
        if r1 > 0:
            Q[k, index[(s - 1, i + 1)]] += r1
        if r2 > 0:
            Q[k, index[(s, i - 1)]] += r2
        Q[k, k] = -(r1 + r2)
    P = scipy.linalg.expm(Q)[index[(4, 2)]]
    Nrep = 30000
    m = si.SIR(u0={"S": [4], "I": [2], "R": [0]}, tspan=np.array([0.0]),
               beta=beta, gamma=gamma, replicates=Nrep)
    r = oracle.run(m, seed=11)
    assert r.err == 0
    U = r.U.reshape(3, Nrep, order="F")
    obs = np.zeros(len(states))
    for s, i in zip(U[0], U[1]):
        obs[index[(s, i)]] += 1
    assert _chisq_p(obs, P, Nrep) > 1e-4
    # aem on the same chain (one replicate per node instead).
    m2 = si.SIR(u0={"S": np.full(Nrep, 4), "I": 2, "R": 0}, tspan=np.array([0.0]),
                beta=beta, gamma=gamma)
    r2 = oracle.run(m2, solver="aem", seed=12)
    U2 = r2.U[:, 0].reshape(-1, 3)
    obs2 = np.zeros(len(states))
    for s, i, _ in U2:
        obs2[index[(s, i)]] += 1
    assert _chisq_p(obs2, P, Nrep) > 1e-4


def _sir_final_size_pmf(S0, I0, beta, gamma):
    """Exact law of S at extinction for the Markov SIR with frequency
    dependent infection beta*S*I/N: along the embedded jump chain an
    infection happens with probability beta*S / (beta*S + gamma*N)."""
    N = S0 + I0
    P = {(S0, I0): 1.0}
    final = np.zeros(S0 + 1)
    # Process states in order of decreasing S + I... iterate by events.
    frontier = dict(P)
    while frontier:
        nxt = {}
        for (s, i), pr in frontier.items():
            if i == 0:
                final[s] += pr
                continue
            pi = beta * s / (beta * s + gamma * N)
            if s > 0:
                nxt[(s - 1, i + 1)] = nxt.get((s - 1, i + 1), 0.0) + pr * pi
            nxt[(s, i - 1)] = nxt.get((s, i - 1), 0.0) + pr * (1 - pi)
        frontier = nxt
    return final


def test_final_size_distribution_readme_nodes(oracle):
    """The README model's nodes (99, 5, 0), beta 0.16, gamma 0.077, run long enough
    for every outbreak to end: the law of the final S against the exact
    final-size distribution of the chain."""
    Nn = 6000
    m = W.mparse_sir_readme(Nn=Nn, days=700)
    pmf = _sir_final_size_pmf(99, 5, 0.16, 0.077)
    for solver in ("ssm", "aem"):
        r = oracle.run(m, solver=solver, seed=123)
        assert r.err == 0
        U = r.U[:, -1].reshape(Nn, 3)
        assert (U[:, 1] == 0).all()
        obs = np.bincount(U[:, 0], minlength=100)
        assert _chisq_p(obs, pmf, Nn) > 1e-4, solver


def test_conservation_and_nonnegativity(oracle):
    m = W.seir_million(Nn=3000, days=60, seed=4, seeded_frac=0.1)
    for solver in ("ssm", "aem"):
        r = oracle.run(m, solver=solver, seed=2)
        assert r.err == 0 and (r.U >= 0).all()
        tot = r.U.reshape(3000, 4, -1).sum(1)
        assert (tot == m.u0.sum(0)[:, None]).all()


def test_external_transfers_conserve_population(oracle):
    m = W.sise3_herds(Nn=200, days=120, seed=2, enter_rate=0.0, exit_rate=0.0,
                      internal_every=0)
    assert (m.events["event"] == 3).all() and len(m.events["event"]) > 1000
    r = oracle.run(m, seed=3)
    assert r.err == 0
    tot = r.U.sum(0)
    assert (tot == m.u0.sum()).all()


def test_inert_nodes_never_change(oracle):
    m = W.seir_million(Nn=2000, days=50, seed=1, seeded_frac=0.01)
    r = oracle.run(m, seed=5)
    inert = m.u0[2, :] == 0
    assert (r.U.reshape(2000, 4, -1)[inert] == m.u0.T[inert][:, :, None]).all()


def test_crn_equals_mssm_and_replicates_independent(oracle):
    m = W.mparse_sir_readme(Nn=10, days=40, replicates=20)
    a = oracle.run(m, seed=9)
    b = oracle.run(m, seed=9, crn=True)
    np.testing.assert_array_equal(a.U, b.U)
    U = a.U.reshape(30, 20, 40, order="F")
    assert not all((U[:, 0] == U[:, r]).all() for r in range(1, 20))


def test_aem_has_no_finite_ties(oracle):
    m = W.seir_million(Nn=3000, days=80, seed=6, seeded_frac=0.2)
    r = oracle.run(m, solver="aem", seed=1)
    assert r.err == 0 and r.aem_ties == 0


def test_reference_error_cases(oracle):
    m = si.SIR(u0={"S": [99, 99], "I": 1, "R": 0}, tspan=np.arange(1, 101.0),
               beta=0.16, gamma=-0.077)
    assert oracle.run(m, solver="ssm", seed=1).err == -11
    assert oracle.run(m, solver="aem", seed=1).err == -11
    m = si.SIR(u0={"S": [99], "I": 1, "R": 0}, tspan=np.arange(1, 101.0),
               beta=0.16, gamma=0.0)
    S = m.S.tocsc(); S.data[0] = -1000; m.S = S
    assert oracle.run(m, solver="ssm", seed=1).err == -1
    assert oracle.run(m, solver="aem", seed=1).err == -1
    m = si.SIR(u0={"S": [9, 9], "I": 1, "R": 0}, tspan=np.arange(1, 5.0),
               beta=0.16, gamma=0.1, replicates=2)
    assert oracle.run(m, solver="aem", seed=1).err == -19


@pytest.mark.parametrize("solver", ["ssm", "aem"])
def test_threads_give_the_same_results(oracle, solver):
    """The checker's node loops over several OpenMP threads (the
    reference's static node-block partition) give the one-thread results:
    every node keeps its own stream."""
    for m in (W.seir_million(Nn=4000, days=60, seed=3, seeded_frac=0.05),
              W.sise3_herds(Nn=200, days=80, seed=2)):
        a = oracle.run(m, solver=solver, seed=4)
        b = oracle.run(m, solver=solver, seed=4, threads=4)
        assert a.err == b.err == 0
        np.testing.assert_array_equal(a.U, b.U)
        if a.V is not None:
            np.testing.assert_array_equal(a.V, b.V)
        assert a.transitions == b.transitions
    m = si.SIR(u0={"S": [99] * 50, "I": 1, "R": 0}, tspan=np.arange(1, 11.0),
               beta=0.16, gamma=-0.077)
    assert oracle.run(m, solver=solver, seed=1, threads=4).err == -11


def test_replicate_parameters(oracle):
    """Parameters per replicate: the same values in every replicate give
    the shared-parameter run bit for bit; a replicate given zero rates
    keeps its initial state."""
    m = W.mparse_sir_readme(Nn=20, days=40, replicates=5)
    a = oracle.run(m, seed=3)
    g = np.tile(m.gdata.reshape(-1, 1), (1, 5))
    b = oracle.run(m, seed=3, rep_gdata=g)
    np.testing.assert_array_equal(a.U, b.U)
    g[:, 2] = 0.0
    c = oracle.run(m, seed=3, rep_gdata=g)
    # U column r * tlen + t.
    U = c.U.reshape(60, 40, 5, order="F")
    assert (U[:, :, 2] == m.u0[:, :20].ravel(order="F")[:, None]).all()
    np.testing.assert_array_equal(U[:, :, 1], a.U.reshape(60, 40, 5, order="F")[:, :, 1])
    s = si.SEIR(u0={"S": np.arange(100, 130), "E": 0, "I": 3, "R": 0},
                tspan=np.arange(1, 31.0), beta=0.3, epsilon=0.2, gamma=0.1, replicates=4)
    a = oracle.run(s, seed=8)
    ld = np.tile(s.ldata, (1, 4))
    b = oracle.run(s, seed=8, rep_ldata=ld)
    np.testing.assert_array_equal(a.U, b.U)
    ld[:, 30:60] = 0.0
    c = oracle.run(s, seed=8, rep_ldata=ld)
    U = c.U.reshape(30 * 4, 30, 4, order="F")
    assert (U[:, :, 1] == s.u0[:, :30].ravel(order="F")[:, None]).all()
    # One column per replicate for all its nodes: the same as spelling it
    # out per node.
    per = np.array([[0.3, 0.2, 0.1], [0.0, 0.0, 0.0], [0.5, 0.1, 0.2], [0.2, 0.3, 0.05]]).T
    d = oracle.run(s, seed=8, rep_ldata=per)
    e = oracle.run(s, seed=8, rep_ldata=np.repeat(per, 30, axis=1))
    np.testing.assert_array_equal(d.U, e.U)
