"""The particle filter's arithmetic and the checker's filter.

Pins: exp, log n! and the binomial and Poisson log-probabilities against
numpy/scipy;
systematic resampling against its defining property (every particle gets
floor(n w) or floor(n w) + 1 offspring, ancestors in order) and the error
cases of SimInf_particle_filter.c:42-48; the interval split of
pfilter_tspan (R/pfilter.R:177-193); and the whole filter against the
exact likelihood of a one-node SIR (the filter's likelihood estimate is
unbiased: the mean of exp(loglik) over independent runs must match the
forward algorithm on the CTMC's transition matrices)."""
import math

import numpy as np
import pytest
from scipy import linalg, special, stats

import siminf_amd as si
from siminf_amd.particle_filter import pfilter_tspan

INVALID_OBS = -101


def test_exp_accuracy(oracle):
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-745, 709, 200000), rng.uniform(-1, 1, 100000),
                        [0.0, -0.0, 1.0, -1.0, 1e-300, -1e-300, 709.7, -708.4, -745.1]])
    y = oracle.exp(x)
    ref = np.exp(x)
    normal = ref > 2.3e-308
    ulp = np.abs(y[normal] - ref[normal]) / np.spacing(ref[normal])
    assert ulp.max() <= 2.0
    assert np.all(np.abs(y[~normal] - ref[~normal]) <= 2 * 5e-324)
    assert oracle.exp([1000.0])[0] == np.inf and oracle.exp([-1000.0])[0] == 0.0
    assert np.isnan(oracle.exp([np.nan])[0])


def test_lfact_and_log_dbinom(oracle):
    n = np.concatenate([np.arange(0, 200), np.random.default_rng(2).integers(200, 10**6, 5000)])
    lf = oracle.lfact(n)
    assert np.allclose(lf, special.gammaln(n + 1.0), rtol=1e-14, atol=1e-12)
    rng = np.random.default_rng(3)
    N = rng.integers(0, 3000, 20000)
    x = (rng.random(20000) * (N + 1)).astype(int)
    for p in (0.5, 0.1, 0.93):
        y = oracle.log_dbinom(x, N, p)
        assert np.allclose(y, stats.binom.logpmf(x, N, p), rtol=1e-12, atol=1e-9)
    assert oracle.log_dbinom([5], [3], 0.5)[0] == -np.inf
    assert oracle.log_dbinom([0], [7], 0.0)[0] == 0.0
    assert oracle.log_dbinom([7], [7], 1.0)[0] == 0.0


def test_log_dpois(oracle):
    rng = np.random.default_rng(6)
    lam = np.concatenate([rng.uniform(1e-6, 5, 10000), rng.uniform(5, 2e5, 10000)])
    x = rng.poisson(lam)
    x[::7] = rng.integers(0, 400, len(x[::7]))
    y = oracle.log_dpois(x, lam)
    assert np.allclose(y, stats.poisson.logpmf(x, lam), rtol=1e-12, atol=1e-8)
    assert oracle.log_dpois([0], 0.0)[0] == 0.0
    assert oracle.log_dpois([3], 0.0)[0] == -np.inf
    assert oracle.log_dpois([-1], 2.0)[0] == -np.inf


def test_systematic_resampling(oracle):
    rng = np.random.default_rng(4)
    for trial in range(200):
        n = int(rng.integers(1, 300))
        w = rng.gamma(0.3, size=n)
        w /= w.sum()
        err, idx = oracle.resample(w, float(rng.random()))
        assert err == 0
        assert np.all(np.diff(idx) >= 0) and idx.min() >= 0 and idx.max() < n
        counts = np.bincount(idx, minlength=n)
        lo = np.floor(n * w - 1e-9)
        assert np.all(counts >= lo) and np.all(counts <= lo + 1 + 1e-9)
    # equal weights: every particle once
    err, idx = oracle.resample(np.full(50, 0.02), 0.37)
    assert err == 0 and np.array_equal(idx, np.arange(50))
    assert oracle.resample([0.5, -0.1, 0.6], 0.5)[0] == -1
    assert oracle.resample([0.5, np.inf], 0.5)[0] == -1
    assert oracle.resample([0.0, 0.0], 0.5)[0] == -2


def test_pfilter_tspan():
    """R/pfilter.R:177-193."""
    t = pfilter_tspan(1.0, [1, 2, 3])
    assert np.isnan(t[:, 0]).all() and t[:, 1].tolist() == [1, 2, 3]
    t = pfilter_tspan(1.0, [5, 6, 10])
    assert t[0].tolist() == [1, 5] and np.isnan(t[1, 0]) and t[2].tolist() == [6, 10]
    with pytest.raises(ValueError):
        pfilter_tspan(3.0, [2, 4])
    with pytest.raises(ValueError):
        pfilter_tspan(1.0, [2, 2])


def _sir(n_nodes, S, I, beta, gamma, t0=1.0):
    return si.SIR_mparse(u0={"S": [S] * n_nodes, "I": [I] * n_nodes, "R": [0] * n_nodes},
                         tspan=np.array([t0, t0 + 1.0]), beta=beta, gamma=gamma)


def test_pfilter_all_missing(oracle):
    """No observation: every weight is 0, the increments are exactly 0,
    every particle keeps itself."""
    m = _sir(4, 40, 3, 0.3, 0.1)
    obs = -np.ones((5, 4), dtype=int)
    err, ll, ess, anc = oracle.pfilter(m, 1, [1, 2, 3, 5, 9], obs, 37, seed=4)
    assert err == 0 and ll == 0.0
    assert np.allclose(ess, 37.0, rtol=1e-14) and np.all(anc == np.arange(37))


def test_pfilter_frozen_model(oracle):
    """Zero rates: every particle keeps u0, so the log-likelihood is the
    sum of the observations' binomial log-probabilities."""
    m = _sir(6, 50, 8, 0.0, 0.0)
    rng = np.random.default_rng(5)
    obs = rng.binomial(8, 0.5, size=(4, 6))
    err, ll, ess, anc = oracle.pfilter(m, 1, [1, 2, 4, 8], obs, 20, obs_prob=0.5, seed=2)
    assert err == 0
    assert ll == pytest.approx(stats.binom.logpmf(obs, 8, 0.5).sum(), rel=1e-12)
    assert np.allclose(ess, 20.0)
    # the sum over the nodes as one observation
    err, ll, _, _ = oracle.pfilter(m, 1, [1, 2, 4, 8], obs.sum(axis=1), 20, seed=2)
    assert ll == pytest.approx(stats.binom.logpmf(obs.sum(axis=1), 48, 0.5).sum(), rel=1e-12)
    # Poisson(lambda = 0.3 X + 0.2)
    err, ll, _, _ = oracle.pfilter(m, 1, [1, 2, 4, 8], obs, 20, obs_prob=0.3, seed=2,
                                   obs_dist="poisson", obs_offset=0.2)
    assert err == 0
    assert ll == pytest.approx(stats.poisson.logpmf(obs, 0.3 * 8 + 0.2).sum(), rel=1e-12)


def test_pfilter_impossible_observation(oracle):
    """A particle with zero likelihood is an invalid observation process
    (pfilter.R:455-458)."""
    m = _sir(1, 50, 2, 0.0, 0.0)
    err, _, _, _ = oracle.pfilter(m, 1, [1, 2], [[1], [9]], 10, seed=1)
    assert err == INVALID_OBS


def _sir_generator(N, beta, gamma):
    """CTMC generator of a one-node SIR with S + I + R = N over the
    states (S, I)."""
    states = [(s, i) for s in range(N + 1) for i in range(N + 1 - s)]
    index = {st: k for k, st in enumerate(states)}
    Q = np.zeros((len(states), len(states)))
    for (s, i), k in index.items():
        inf = beta * s * i / N
        rec = gamma * i
        if inf > 0:
            Q[k, index[(s - 1, i + 1)]] += inf
        if rec > 0:
            Q[k, index[(s, i - 1)]] += rec
        Q[k, k] = -(inf + rec)
    return states, index, Q


def _exact_likelihood(N, S0, I0, beta, gamma, times, emission):
    """Forward algorithm: alpha <- alpha exp(Q elapsed) diag(emission(k, I))."""
    states, index, Q = _sir_generator(N, beta, gamma)
    tsp = pfilter_tspan(1.0, times)
    alpha = np.zeros(len(states))
    alpha[index[(S0, I0)]] = 1.0
    for k, (tb, te) in enumerate(tsp):
        elapsed = 1.0 if math.isnan(tb) else te + 1.0 - tb
        alpha = alpha @ linalg.expm(Q * elapsed)
        alpha = alpha * np.array([emission(k, i) for (s, i) in states])
    return alpha.sum()


def test_pfilter_unbiased_likelihood_poisson(oracle):
    """As below with Poisson(lambda = p I + c) emissions, which give every
    particle a positive weight for any count."""
    N, S0, I0, beta, gamma, p, c = 8, 6, 2, 0.9, 0.3, 0.5, 0.3
    times = [2.0, 4.0, 5.0]
    obs = np.array([[1], [3], [0]])
    m = _sir(1, S0, I0, beta, gamma, t0=1.0)
    exact = _exact_likelihood(N, S0, I0, beta, gamma, times,
                              lambda k, i: stats.poisson.pmf(obs[k, 0], p * i + c))
    est = []
    for seed in range(400):
        err, ll, _, _ = oracle.pfilter(m, 1, times, obs, 50, obs_prob=p, seed=seed,
                                       obs_dist="poisson", obs_offset=c)
        assert err == 0
        est.append(math.exp(ll))
    est = np.array(est)
    se = est.std(ddof=1) / math.sqrt(len(est))
    assert abs(est.mean() - exact) < 4 * se + 1e-12, (est.mean(), exact, se)


def test_pfilter_unbiased_likelihood(oracle):
    """E[exp(loglik)] equals the exact likelihood (forward algorithm on
    exp(Q * elapsed) with binomial emissions).  Each interval simulates
    the days from its first time point to one day past its last, as the
    reference's runs do: elapsed 1 for a one-point interval, t_i + 1 - t_(i-1)
    for a two-point one."""
    # Zero counts are possible for every particle, so no weight is -inf
    # (which the reference treats as an invalid observation process);
    # they still weigh particles by (1 - p)^I.
    N, S0, I0, beta, gamma, p = 8, 6, 2, 0.9, 0.3, 0.4
    times = [2.0, 3.0, 6.0]
    obs = np.array([[0], [0], [0]])
    m = _sir(1, S0, I0, beta, gamma, t0=1.0)
    states, index, Q = _sir_generator(N, beta, gamma)
    tsp = pfilter_tspan(1.0, times)
    alpha = np.zeros(len(states))
    alpha[index[(S0, I0)]] = 1.0
    for k, (tb, te) in enumerate(tsp):
        elapsed = 1.0 if math.isnan(tb) else te + 1.0 - tb
        alpha = alpha @ linalg.expm(Q * elapsed)
        emis = np.array([stats.binom.pmf(obs[k, 0], i, p) for (s, i) in states])
        alpha = alpha * emis
    exact = alpha.sum()
    est = []
    for seed in range(400):
        err, ll, _, _ = oracle.pfilter(m, 1, times, obs, 50, obs_prob=p, seed=seed)
        assert err == 0
        est.append(math.exp(ll))
    est = np.array(est)
    se = est.std(ddof=1) / math.sqrt(len(est))
    assert abs(est.mean() - exact) < 4 * se + 1e-12, (est.mean(), exact, se)


def _enter_events(node_times, select, n=1, event=1, dest=None):
    rows = [(event, t, nd, 0 if dest is None else dest, n, 0.0, select, 0)
            for nd, t in node_times]
    rows.sort(key=lambda r: r[1])
    return {k: np.array([r[i] for r in rows]) for i, k in enumerate(
        ("event", "time", "node", "dest", "n", "proportion", "select", "shift"))}


def test_pfilter_interval_events_closed_form(oracle):
    """Zero rates and one individual entering I at node 1 on days 2..20:
    at an observation time t the particle holds I0 + #(events due at or
    before t), so the log-likelihood is a sum of Poisson log-probabilities.
    An interval that ran an earlier interval's events again, or dropped
    its own, changes I (pfilter_events, R/pfilter.R:200-220)."""
    times = [3.0, 4.0, 9.0, 15.0, 16.0]
    ev = _enter_events([(1, t) for t in range(2, 21)], select=2)
    m = si.SIR(u0={"S": [10, 10], "I": [2, 0], "R": [0, 0]}, tspan=[1.0, 2.0],
               beta=0.0, gamma=0.0, events=ev)
    obs = np.array([[3, 0], [4, 1], [7, 0], [15, 2], [16, 0]])
    err, ll, ess, _ = oracle.pfilter(m, 1, times, obs, 8, obs_prob=0.5, seed=3,
                                     obs_dist="poisson", obs_offset=0.25)
    assert err == 0
    I1 = [2 + sum(1 for d in range(2, 21) if d <= t) for t in times]
    exact = sum(stats.poisson.logpmf(obs[k, 0], 0.5 * I1[k] + 0.25) +
                stats.poisson.logpmf(obs[k, 1], 0.25) for k in range(len(times)))
    assert ll == pytest.approx(exact, rel=1e-12)
    assert np.allclose(ess, 8.0)


def test_pfilter_interval_external_transfers_closed_form(oracle):
    """External transfers of one infected individual from node 1 to node 2
    on days 2, 5, 6 and 12, zero rates: per-node I is known at every
    observation time."""
    days = [2, 5, 6, 12]
    ev = _enter_events([(1, t) for t in days], select=2, event=3, dest=2)
    m = si.SIR(u0={"S": [5, 5], "I": [6, 0], "R": [0, 0]}, tspan=[1.0, 2.0],
               beta=0.0, gamma=0.0, events=ev)
    times = [2.0, 5.0, 6.0, 10.0, 13.0]
    moved = [sum(1 for d in days if d <= t) for t in times]
    obs = np.array([[6 - k, k] for k in moved])
    err, ll, _, _ = oracle.pfilter(m, 1, times, obs, 5, obs_prob=1.0, seed=1,
                                   obs_dist="poisson", obs_offset=0.5)
    assert err == 0
    exact = sum(stats.poisson.logpmf(6 - k, 6 - k + 0.5) + stats.poisson.logpmf(k, k + 0.5)
                for k in moved)
    assert ll == pytest.approx(exact, rel=1e-12)
