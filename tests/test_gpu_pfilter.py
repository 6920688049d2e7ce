"""GPU parity of the device particle filter: log-likelihood, effective
sample sizes and ancestors must equal the checker's bit for bit (both run
the same SSA streams and the same weight, log-sum-exp and resampling
arithmetic)."""
import numpy as np
import pytest

import siminf_amd as si

pytestmark = pytest.mark.gpu


def _sir(n_nodes, seed=0):
    rng = np.random.default_rng(seed)
    S = rng.integers(40, 200, n_nodes)
    I = rng.integers(1, 8, n_nodes)
    return si.SIR_mparse(u0={"S": S, "I": I, "R": np.zeros(n_nodes, int)},
                         tspan=np.array([1.0, 2.0]), beta=0.25, gamma=0.1)


def _check(oracle, model, c, times, obs, P, seed, p=0.5, dist="binomial", offset=0.0):
    print("case: pfilter %s Nn=%d P=%d intervals=%d obs=%s %s seed=%d" % (
        getattr(model.plugin, "name", model.plugin), model.Nn, P, len(times),
        np.shape(obs), dist, seed), flush=True)
    kw = dict(obs_prob=p, seed=seed, obs_dist=dist, obs_offset=offset)
    err, ll, ess, anc = oracle.pfilter(model, c, times, obs, P, **kw)
    if err:
        with pytest.raises(si.SimInfError) as e:
            si.pfilter(model, c, times, obs, P, **kw)
        assert e.value.code == err
        return err, None
    r = si.pfilter(model, c, times, obs, P, **kw)
    assert r.loglik == ll
    np.testing.assert_array_equal(r.ess, ess)
    np.testing.assert_array_equal(r.ancestors, anc)
    return 0, r


def test_pfilter_per_node_observations(oracle):
    m = _sir(40)
    times = [1, 2, 3, 4, 8, 9, 15]
    rng = np.random.default_rng(1)
    obs = np.where(rng.random((len(times), 40)) < 0.5, 0, -1)   # zeros and missing
    err, r = _check(oracle, m, 1, times, obs, 64, seed=3)
    assert err == 0 and np.isfinite(r.loglik) and r.loglik < 0


def test_pfilter_aggregate_observations(oracle):
    m = _sir(120, seed=2)
    times = np.arange(1, 31, 7.0)
    obs = np.array([[20], [25], [30], [20], [10]])
    err, r = _check(oracle, m, 1, times, obs, 256, seed=9)
    assert err == 0


def test_pfilter_poisson_observations(oracle):
    m = _sir(50, seed=4)
    times = np.arange(1, 43, 7.0)
    rng = np.random.default_rng(8)
    obs = rng.poisson(2.0, size=(len(times), 50))
    err, r = _check(oracle, m, 1, times, obs, 128, seed=11, p=0.5, dist="poisson",
                    offset=0.1)
    assert err == 0 and np.isfinite(r.loglik)
    agg = obs.sum(axis=1, keepdims=True)
    err, r = _check(oracle, m, 1, times, agg, 128, seed=12, p=0.5, dist="poisson",
                    offset=1e-6)
    assert err == 0


def test_pfilter_seir_builtin(oracle):
    m = si.SEIR(u0={"S": [300] * 16, "E": [0] * 16, "I": [4] * 16, "R": [0] * 16},
                tspan=np.array([1.0, 2.0]), beta=0.3, epsilon=0.2, gamma=0.1)
    times = np.arange(3, 40, 4.0)
    obs = np.zeros((len(times), 1), dtype=int)
    _check(oracle, m, 2, times, obs, 100, seed=5, p=0.1)


@pytest.mark.parametrize("P", [1, 7168, 7169, 9000])
def test_pfilter_particle_counts(oracle, P):
    """Weights staged in LDS up to 7168 particles, in a global buffer
    above; one particle resamples itself."""
    m = _sir(3, seed=6)
    times = np.arange(1, 20, 3.0)
    obs = np.random.default_rng(P).poisson(20.0, size=(len(times), 1))
    err, r = _check(oracle, m, 1, times, obs, P, seed=P, p=0.3, dist="poisson", offset=1.0)
    assert err == 0
    if P == 1:
        assert np.all(r.ancestors == 0)


def test_pfilter_invalid_observation(oracle):
    m = _sir(3)
    err, _ = _check(oracle, m, 1, [1, 2], [[0, 0, 0], [10**6, 0, 0]], 16, seed=1)
    assert err == -101


def test_pfilter_interval_events(oracle):
    """Scheduled events split per interval: enter events (E1, k_ssm_nodes)
    and external transfers (E2, one workgroup per particle), with
    individual sampling drawing from the streams."""
    from siminf_amd import workloads as W
    rng = np.random.default_rng(3)
    m = _sir(20, seed=5)
    rows = [(1, int(t), int(nd), 0, int(rng.integers(1, 4)), 0.0, 4, 0)
            for t, nd in zip(rng.integers(1, 40, 60), rng.integers(1, 21, 60))]
    rows.sort(key=lambda r: r[1])
    ev = {k: np.array([r[i] for r in rows]) for i, k in enumerate(
        ("event", "time", "node", "dest", "n", "proportion", "select", "shift"))}
    me = si.SIR_mparse(u0={"S": m.u0[0], "I": m.u0[1], "R": m.u0[2]},
                       tspan=np.array([1.0, 2.0]), beta=0.25, gamma=0.1, events=ev,
                       E=si.models.select_matrix_SIR())
    times = np.array([3, 4, 8, 15, 22, 29, 36, 40.0])
    obs = rng.poisson(20.0, size=(len(times), 1))
    err, _ = _check(oracle, me, 1, times, obs, 96, seed=21, p=0.5, dist="poisson",
                    offset=1.0)
    assert err == 0
    h = W.sise3_herds(Nn=40, days=50, seed=6)      # the filter reads tspan[0] only
    times = np.arange(5, 51, 5.0)
    obs = rng.poisson(5.0, size=(len(times), 1))
    err, _ = _check(oracle, h, 1, times, obs, 48, seed=7, p=0.5, dist="poisson",
                    offset=1.0)
    assert err == 0
