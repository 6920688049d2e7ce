"""ABC-SMC on the GPU: the device weights equal the checker's bit for bit,
and the whole loop (one node per proposal, the simulations and the weights
on the device) recovers a known infection rate."""
import numpy as np
import pytest

import siminf_amd as si
from siminf_amd import abc_smc as A
from tests.test_abc import _case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,priors", [
    (40, ["beta ~ uniform(0, 1)"]),
    (300, ["beta ~ uniform(0, 1)", "gamma ~ gamma(2, 3)"]),
    (3000, ["a ~ normal(0.5, 0.2)", "b ~ lognormal(0, 1)", "c ~ uniform(-1, 2)"]),
])
def test_device_weights_equal_checker(oracle, n, priors):
    print("case: abc weights n=%d d=%d" % (n, len(priors)), flush=True)
    pr, x, xx, w, sigma = _case(n, n, priors)
    dist = "".join(p.kind for p in pr)
    p1, p2 = [p.p1 for p in pr], [p.p2 for p in pr]
    lib = si._native.load()
    err, ww = A.call_weights(lib.SimInf_hip_abc_weights, dist, p1, p2, x, xx, w, sigma)
    rerr, rww = oracle.abc_weights(dist, p1, p2, x, xx, w, sigma)
    assert err == rerr == 0
    np.testing.assert_array_equal(ww, rww)
    bad = xx.copy()
    bad[n // 2, 0] = 7.0 if dist[0] == "u" else bad[n // 2, 0]
    if dist[0] == "u":
        err, _ = A.call_weights(lib.SimInf_hip_abc_weights, dist, p1, p2, x, bad, w, sigma)
        assert err == oracle.abc_weights(dist, p1, p2, x, bad, w, sigma)[0] == -102


def _final_recovered(model):
    """Recovered individuals per node at the last time point."""
    return model.U[2::3, -1].astype(float)


def test_abc_ldata_recovers_beta():
    truth = si.SIR(u0={"S": [990], "I": [10], "R": [0]}, tspan=np.arange(1, 61.0),
                   beta=0.3, gamma=0.1)
    si.run(truth, seed=11)
    observed = _final_recovered(truth)[0]
    print("case: abc ldata, observed R=%d" % observed, flush=True)

    def distance(result, generation, data):
        return np.abs(_final_recovered(result) - data) / 1000.0

    model = si.SIR(u0={"S": [990], "I": [10], "R": [0]}, tspan=np.arange(1, 61.0),
                   beta=0.2, gamma=0.1)
    fit = si.abc(model, ["beta ~ uniform(0.05, 0.8)"], n_particles=200,
                 distance=distance, tolerance=[0.3, 0.1, 0.03], data=observed, seed=5)
    assert len(fit.x) == 3
    for g in range(3):
        assert fit.x[g].shape == (200, 1)
        assert np.all(fit.distance[g] <= fit.tolerance[g])
        assert abs(fit.weight[g].sum() - 1.0) < 1e-12 and fit.ess[g] > 10
    beta = fit.x[-1][:, 0]
    mean = float(np.sum(fit.weight[-1] * beta))
    assert 0.2 < mean < 0.45 and np.std(beta) < np.std(fit.x[0][:, 0])


def test_abc_gdata_one_run_per_proposal():
    model = si.SIR_mparse(u0={"S": [99], "I": [5], "R": [0]}, tspan=np.arange(1, 31.0),
                          beta=0.16, gamma=0.077)
    print("case: abc gdata", flush=True)

    def distance(result, generation, data):
        return [abs(float(result.U[1, -1]) - 10.0) / 100.0]

    fit = si.abc(model, ["beta ~ uniform(0.05, 0.5)"], n_particles=12, distance=distance,
                 tolerance=[0.5, 0.2], seed=2)
    assert [x.shape for x in fit.x] == [(12, 1), (12, 1)]
    assert all(n >= 12 for n in fit.nprop)


def test_abc_gdata_callbacks():
    """init_model sees every gdata proposal's model before its run (here
    it fixes gamma, which the distance then reads back); post_gen sees the
    result after each generation."""
    model = si.SIR_mparse(u0={"S": [99], "I": [5], "R": [0]}, tspan=np.arange(1, 31.0),
                          beta=0.16, gamma=0.077)
    print("case: abc gdata callbacks", flush=True)
    calls, gens = [], []

    def init_model(m):
        calls.append(float(m.gdata[0]))
        m.gdata[1] = 0.05
        return m

    def distance(result, generation, data):
        assert result.gdata[1] == 0.05
        return [abs(float(result.U[1, -1]) - 10.0) / 100.0]

    fit = si.abc(model, ["beta ~ uniform(0.05, 0.5)"], n_particles=8, distance=distance,
                 tolerance=[0.5, 0.3], seed=3, init_model=init_model,
                 post_gen=lambda r: gens.append(len(r.x)))
    assert gens == [1, 2]
    assert len(calls) == sum(fit.nprop)
    assert all(0.05 <= c <= 0.5 for c in calls)


class _Enough(Exception):
    pass


def test_abc_adaptive_ldata():
    """Adaptive tolerance (tolerance None, n_init proposals first): the
    first generation keeps the n_particles closest of n_init proposals and
    its tolerance is the largest kept distance; later tolerances come from
    the density ratio of successive generations and do not grow."""
    truth = si.SIR(u0={"S": [990], "I": [10], "R": [0]}, tspan=np.arange(1, 61.0),
                   beta=0.3, gamma=0.1)
    si.run(truth, seed=11)
    observed = _final_recovered(truth)[0]
    print("case: abc adaptive ldata, observed R=%d" % observed, flush=True)

    def distance(result, generation, data):
        return np.abs(_final_recovered(result) - data).reshape(-1, 1) / 1000.0

    def post_gen(r):
        if len(r.x) == 4:
            raise _Enough()

    model = si.SIR(u0={"S": [990], "I": [10], "R": [0]}, tspan=np.arange(1, 61.0),
                   beta=0.2, gamma=0.1)
    seen = []
    with pytest.raises(_Enough):
        si.abc(model, ["beta ~ uniform(0.05, 0.8)"], n_particles=150, distance=distance,
               data=observed, seed=7, n_init=450,
               post_gen=lambda r: (seen.append(r), post_gen(r)))
    fit = seen[-1]
    assert fit.nprop[0] == 450 and len(fit.x) == 4
    assert fit.distance[0].shape == (150, 1)
    assert fit.tolerance[0][0] == fit.distance[0].max()
    for g in range(4):
        assert np.all(fit.distance[g] <= fit.tolerance[g])
        assert abs(fit.weight[g].sum() - 1.0) < 1e-12
    tol = [t[0] for t in fit.tolerance]
    assert all(b <= a for a, b in zip(tol, tol[1:])) and tol[-1] < tol[0]
    mean = float(np.sum(fit.weight[-1] * fit.x[-1][:, 0]))
    assert 0.2 < mean < 0.45


def test_abc_adaptive_gdata():
    model = si.SIR_mparse(u0={"S": [99], "I": [5], "R": [0]}, tspan=np.arange(1, 31.0),
                          beta=0.16, gamma=0.077)
    print("case: abc adaptive gdata", flush=True)

    def distance(result, generation, data):
        return [abs(float(result.U[1, -1]) - 10.0) / 100.0]

    seen = []

    def post_gen(r):
        seen.append(r)
        if len(r.x) == 3:
            raise _Enough()

    with pytest.raises(_Enough):
        si.abc(model, ["beta ~ uniform(0.05, 0.5)"], n_particles=10, distance=distance,
               seed=4, n_init=25, post_gen=post_gen)
    fit = seen[-1]
    assert [x.shape for x in fit.x] == [(10, 1)] * 3
    assert fit.nprop[0] == 25 and all(n >= 10 for n in fit.nprop[1:])
    for g in range(3):
        assert np.all(fit.distance[g] <= fit.tolerance[g])
