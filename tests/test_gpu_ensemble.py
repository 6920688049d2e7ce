"""Ensemble agreement on the GPU under the library's own streams.

ssm (direct method) and aem (all-events method) are different algorithms
for one Markov law, and the GPU runs each with its own draws, so their
ensembles over many independent nodes must agree: two-sample KS on final
sizes and on peak prevalence (p > 0.01), and every day's mean and
variance of the infected count inside a simultaneous 99 % band
(Bonferroni over the days).  Each solver's final sizes must also follow
the exact final-size law of the chain once every outbreak is over.

* SIR of the README model (nodes S = 99, I = 5, beta 0.16, gamma 0.077),
  10,000 independent nodes over 150 days; and over 700 days for the
  final-size law.
* SEIR on 1,000,000 nodes over 365 days, ssm and aem in one batch launch:
  the comparisons over the 5,000 seeded nodes (the same node sizes on
  both sides), gathered on the device.
"""
import numpy as np
import pytest
from scipy import stats

import siminf_amd as si
from siminf_amd import workloads as W
from tests.test_oracle import _chisq_p, _sir_final_size_pmf

pytestmark = pytest.mark.gpu


def _band(days):
    """Two-sided normal quantile of a simultaneous 99 % band over days."""
    return stats.norm.ppf(1.0 - 0.01 / (2.0 * days))


def _moments_agree(a, b):
    """a, b: samples x days.  Per-day means and variances of two
    independent ensembles agree within the simultaneous band."""
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    n1, n2 = a.shape[0], b.shape[0]
    z = _band(a.shape[1])
    m1, m2 = a.mean(0), b.mean(0)
    v1, v2 = a.var(0, ddof=1), b.var(0, ddof=1)
    se = np.sqrt(v1 / n1 + v2 / n2)
    live = se > 0
    zm = np.abs(m1 - m2)[live] / se[live]
    # Standard error of a sample variance from the fourth central moment.
    k1 = ((a - m1) ** 4).mean(0)
    k2 = ((b - m2) ** 4).mean(0)
    sv = np.sqrt(np.maximum(k1 - v1 ** 2, 0) / n1 + np.maximum(k2 - v2 ** 2, 0) / n2)
    live_v = sv > 0
    zv = np.abs(v1 - v2)[live_v] / sv[live_v]
    return float(zm.max(initial=0.0)), float(zv.max(initial=0.0)), z


def _run(model, solver, seed):
    with si.Plan(model, solver=solver) as p:
        p.run(seed)
        return p.download().U.copy()


def test_sir_ssm_and_aem_ensembles_agree():
    Nn, days = 10_000, 150
    m = W.mparse_sir_readme(Nn=Nn, days=days)
    print("case: SIR 10000 nodes x 150 days ssm vs aem", flush=True)
    X = {s: _run(m, s, seed).reshape(Nn, 3, days)
         for s, seed in (("ssm", 101), ("aem", 202))}
    I = {s: x[:, 1, :] for s, x in X.items()}
    R = {s: x[:, 2, -1] for s, x in X.items()}
    assert stats.ks_2samp(R["ssm"], R["aem"]).pvalue > 0.01
    assert stats.ks_2samp(I["ssm"].max(1), I["aem"].max(1)).pvalue > 0.01
    zm, zv, z = _moments_agree(I["ssm"], I["aem"])
    assert zm < z and zv < z, (zm, zv, z)
    # The two ensembles are not copies of each other.
    assert (R["ssm"] != R["aem"]).mean() > 0.5


def test_sir_final_size_law_on_device():
    Nn = 10_000
    m = W.mparse_sir_readme(Nn=Nn, days=700)
    pmf = _sir_final_size_pmf(99, 5, 0.16, 0.077)
    for solver, seed in (("ssm", 303), ("aem", 404)):
        print("case: SIR final-size law %s" % solver, flush=True)
        U = _run(m, solver, seed)[:, -1].reshape(Nn, 3)
        assert (U[:, 1] == 0).all()
        obs = np.bincount(U[:, 0], minlength=100)
        assert _chisq_p(obs, pmf, Nn) > 1e-4, solver


def test_seir_million_nodes_ssm_and_aem_ensembles_agree():
    m = W.seir_million()
    seeded = np.flatnonzero(m.u0[2] > 0)
    days = len(m.tspan)
    print("case: SEIR 1e6 nodes x 365 days, ssm+aem batch, %d seeded" % len(seeded),
          flush=True)
    out = {}
    with si.Plan(m, solver="ssm") as ps, si.Plan(m, solver="aem") as pa:
        with si.Batch([ps, pa]) as b:
            b.run([505, 606])
        for s, p in (("ssm", ps), ("aem", pa)):
            # Rows node-major then (I, R); index is one-based.
            T = p.trajectory(["I", "R"], index=seeded + 1)
            out[s] = T.reshape(len(seeded), 2, days)
    I = {s: x[:, 0, :] for s, x in out.items()}
    R = {s: x[:, 1, -1] for s, x in out.items()}
    assert stats.ks_2samp(R["ssm"], R["aem"]).pvalue > 0.01
    assert stats.ks_2samp(I["ssm"].max(1), I["aem"].max(1)).pvalue > 0.01
    zm, zv, z = _moments_agree(I["ssm"], I["aem"])
    assert zm < z and zv < z, (zm, zv, z)
