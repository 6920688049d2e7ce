"""The serving loop on the GPU.

* Each request's result is the same alone and in the mix: requests run
  concurrently on their own streams give log-likelihoods, ESS and
  ancestors equal bit for bit to the same requests run one at a time.
* The device-resident filter agrees in distribution with the plain path
  (one run() per interval, weights and resampling on the host): two-sample
  KS on the log-likelihood estimates over independent seeds, p > 0.01."""
import numpy as np
import pytest
from scipy import stats

import siminf_amd as si
from siminf_amd import particle_filter, serving

pytestmark = pytest.mark.gpu


def _model(n, beta, gamma, seed):
    rng = np.random.default_rng(seed)
    S = rng.integers(50, 150, n)
    return si.SIR_mparse(u0={"S": S, "I": np.full(n, 5), "R": np.zeros(n, int)},
                         tspan=np.array([1.0, 2.0]), beta=beta, gamma=gamma)


def _requests():
    rng = np.random.default_rng(21)
    out = []
    for k, n in enumerate([30, 400, 75, 1200, 10, 250]):
        m = _model(n, float(rng.uniform(0.1, 0.4)), float(rng.uniform(0.05, 0.2)), k)
        times = 1.0 + 5 * np.arange(1, 11)
        obs = rng.poisson(3.0 * n, size=(len(times), 1))
        dist = "poisson" if k % 2 == 0 else "binomial"
        if dist == "binomial":
            obs = np.zeros((len(times), 1), dtype=int)
        out.append(dict(model=m, times=times, obs=obs, P=int(rng.integers(64, 600)),
                        seed=int(rng.integers(0, 2**40)), dist=dist))
    return out


def test_results_same_alone_and_in_the_mix():
    specs = _requests()
    alone = []
    for s in specs:
        print("case: alone Nn=%d P=%d %s" % (s["model"].Nn, s["P"], s["dist"]), flush=True)
        alone.append(particle_filter.pfilter(s["model"], 1, s["times"], s["obs"], s["P"],
                                             obs_prob=0.5, seed=s["seed"],
                                             obs_dist=s["dist"], obs_offset=0.5))
    print("case: mix of %d requests" % len(specs), flush=True)
    with serving.PFilterServer(hbm_budget=8 << 30, max_inflight=len(specs)) as srv:
        reqs = [srv.submit(serving.Request(s["model"], 1, s["times"], s["obs"], s["P"],
                                           obs_prob=0.5, seed=s["seed"],
                                           obs_dist=s["dist"], obs_offset=0.5))
                for s in specs]
        mixed = [r.wait(120) for r in reqs]
    assert srv.peak_inflight >= 2
    for a, b in zip(alone, mixed):
        assert a.loglik == b.loglik
        np.testing.assert_array_equal(a.ess, b.ess)
        np.testing.assert_array_equal(a.ancestors, b.ancestors)


def test_device_filter_agrees_with_plain_path():
    m = _model(12, 0.3, 0.1, 5)
    times = 1.0 + 4 * np.arange(1, 6)
    obs = np.array([[40], [70], [90], [80], [60]])
    dev, host = [], []
    for seed in range(60):
        if seed % 20 == 0:
            print("case: KS seeds %d.." % seed, flush=True)
        r = particle_filter.pfilter(m, 1, times, obs, 100, obs_prob=0.5, seed=seed,
                                    obs_dist="poisson", obs_offset=1.0)
        dev.append(r.loglik)
        ll, _ = serving.host_pfilter(m, 1, times, obs, 100, obs_prob=0.5, seed=10_000 + seed,
                                     obs_dist="poisson", obs_offset=1.0)
        host.append(ll)
    p = stats.ks_2samp(dev, host).pvalue
    assert p > 0.01, (p, np.mean(dev), np.mean(host))
