"""Host logic of the particle-filter serving loop: footprints, admission
by slots and HBM budget in arrival order, error hand-back, and the request
generator of the filter stream (no GPU: the filter is replaced by a stub
runner that records what ran at the same time)."""
import threading
import time

import numpy as np
import pytest

import siminf_amd as si
from siminf_amd import serving
from siminf_amd import workloads as W


def _sir(n):
    return si.SIR_mparse(u0={"S": [99] * n, "I": [5] * n, "R": [0] * n},
                         tspan=np.array([1.0, 2.0]), beta=0.2, gamma=0.1)


def test_footprint_scales_with_lanes():
    a = serving.footprint(_sir(100), 1000)
    b = serving.footprint(_sir(200), 1000)
    c = serving.footprint(_sir(100), 2000)
    per = 4 * 3 * 4 + 8 * 1 * 4
    assert b - a == 100 * 1000 * per
    assert c - a == 100 * 1000 * per + 16 * 1000
    assert a > 100 * 1000 * per


class _Stub:
    """Runner that holds each request for a while and records the
    requests running together."""

    def __init__(self, hold=0.03):
        self.hold = hold
        self.lock = threading.Lock()
        self.running = []
        self.max_running = 0
        self.max_bytes = 0

    def __call__(self, req, lib):
        with self.lock:
            self.running.append(req)
            self.max_running = max(self.max_running, len(self.running))
            self.max_bytes = max(self.max_bytes, sum(r.bytes for r in self.running))
        time.sleep(self.hold)
        with self.lock:
            self.running.remove(req)
        if req.tag == "bad":
            raise si.SimInfError(-101, "Invalid observation process.")
        return req.tag


def _req(bytes_, tag):
    r = serving.Request(_sir(1), 1, [2.0], [[0]], 1, tag=tag)
    r.bytes = bytes_
    return r


def test_admission_respects_slots_budget_and_order():
    stub = _Stub()
    srv = serving.PFilterServer(hbm_budget=100, max_inflight=3, runner=stub)
    sizes = [40, 30, 20, 60, 10, 10, 90, 5, 5, 5]
    with srv:
        reqs = [srv.submit(_req(b, k)) for k, b in enumerate(sizes)]
        results = [r.wait(10) for r in reqs]
    assert results == list(range(len(sizes)))
    assert stub.max_running <= 3 and stub.max_bytes <= 100
    assert srv.peak_inflight <= 3 and srv.peak_bytes <= 100
    admit = [r.t_admit for r in reqs]
    assert admit == sorted(admit)                        # first come, first served
    assert stub.max_running >= 2                         # requests did overlap
    assert srv.free == 100 and srv.inflight == 0


def test_errors_are_handed_back_and_free_their_memory():
    stub = _Stub(hold=0.01)
    with serving.PFilterServer(hbm_budget=50, max_inflight=2, runner=stub) as srv:
        bad = srv.submit(_req(50, "bad"))
        good = srv.submit(_req(50, "good"))
        with pytest.raises(si.SimInfError):
            bad.wait(10)
        assert good.wait(10) == "good"
    assert srv.free == 50


def test_over_budget_request_is_refused():
    with serving.PFilterServer(hbm_budget=10, runner=_Stub()) as srv:
        with pytest.raises(ValueError):
            srv.submit(_req(11, "big"))


def test_latency_covers_queueing():
    stub = _Stub(hold=0.05)
    with serving.PFilterServer(hbm_budget=10, max_inflight=1, runner=stub) as srv:
        reqs = [srv.submit(_req(10, k)) for k in range(3)]
        for r in reqs:
            r.wait(10)
    # the third waited for the first two
    assert reqs[2].latency >= 0.1 and reqs[2].t_start >= reqs[1].t_done


def test_request_stream_generator():
    specs = W.pfilter_requests(12, seed=3, n_particles=50)
    nodes = np.array([s.model.Nn for s in specs])
    assert nodes.min() >= 100 and nodes.max() <= 100000
    arr = np.array([s.arrival for s in specs])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    for s in specs:
        assert s.obs.shape == (30, 1) and (s.obs >= 0).all()
        assert np.array_equal(s.times, 1.0 + 7 * np.arange(1, 31))
        assert 0.05 <= s.beta <= 0.4 and 0.03 <= s.gamma <= 0.2
        assert list(s.model.gdata) == [s.beta, s.gamma]
        assert s.model.u0[:, 0].tolist() == [99, 5, 0]
