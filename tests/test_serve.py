"""Host logic of the served request stream (wsclean_amd/serve.py):
admission by slot and HBM budget, arrival order, latency statistics.  The
per-request work is a host stand-in here; the GPU tests run the gridder."""
import threading
import time

import numpy as np
import pytest

from wsclean_amd.serve import Server, latency_stats, major_cycle_model


class FakeRequest:
    def __init__(self, rid, arrival, nbytes, work_s, n_vis=1000, cycles=5):
        self.rid, self.arrival, self.nbytes, self.work_s = rid, arrival, nbytes, work_s
        self.n_vis, self.cycles = n_vis, cycles
        self.start = self.done = 0.0
        self.image = None

    def hbm_bytes(self):
        return self.nbytes


class Recorder:
    def __init__(self):
        self.lock = threading.Lock()
        self.running = 0
        self.bytes = 0
        self.max_running = 0
        self.max_bytes = 0
        self.starts = []

    def run(self, req):
        with self.lock:
            self.running += 1
            self.bytes += req.nbytes
            self.max_running = max(self.max_running, self.running)
            self.max_bytes = max(self.max_bytes, self.bytes)
            self.starts.append(req.rid)
        time.sleep(req.work_s)
        with self.lock:
            self.running -= 1
            self.bytes -= req.nbytes
        return req.rid


def test_slots_bound_concurrency_and_all_requests_finish():
    rec = Recorder()
    reqs = [FakeRequest(i, 0.002 * i, 10, 0.03) for i in range(12)]
    out = Server(slots=3, hbm_budget=1 << 40, run=rec.run).serve(reqs)
    assert [r.rid for r in out] == list(range(12))
    assert rec.max_running == 3
    assert all(r.image == r.rid and r.done > r.start >= r.arrival - 1e-3 for r in out)


def test_hbm_budget_is_respected_and_admission_is_in_arrival_order():
    rec = Recorder()
    sizes = [60, 50, 30, 70, 20, 90, 10, 40]
    reqs = [FakeRequest(i, 0.0, s, 0.02 + 0.005 * (i % 3)) for i, s in enumerate(sizes)]
    srv = Server(slots=4, hbm_budget=100, run=rec.run)
    srv.serve(reqs)
    assert rec.max_bytes <= 100
    assert srv.peak_bytes <= 100
    assert rec.starts == list(range(len(sizes)))   # FIFO: no overtaking


def test_request_larger_than_budget_runs_alone():
    rec = Recorder()
    reqs = [FakeRequest(0, 0.0, 500, 0.02), FakeRequest(1, 0.0, 10, 0.02)]
    Server(slots=4, hbm_budget=100, run=rec.run).serve(reqs)
    assert rec.max_running == 1


def test_arrivals_are_honoured():
    rec = Recorder()
    reqs = [FakeRequest(0, 0.0, 1, 0.0), FakeRequest(1, 0.15, 1, 0.0)]
    out = Server(slots=2, hbm_budget=10, run=rec.run).serve(reqs)
    assert out[1].start >= 0.15 - 1e-3


def test_error_in_a_request_is_raised_after_the_stream():
    def run(req):
        if req.rid == 1:
            raise RuntimeError("request 1 failed")
        time.sleep(0.01)
        return req.rid
    reqs = [FakeRequest(i, 0.0, 1, 0.0) for i in range(3)]
    with pytest.raises(RuntimeError, match="request 1 failed"):
        Server(slots=2, hbm_budget=10, run=run).serve(reqs)


def test_latency_stats():
    reqs = [FakeRequest(i, float(i), 1, 0.0, n_vis=100, cycles=5) for i in range(100)]
    for i, r in enumerate(reqs):
        r.done = r.arrival + 0.001 * (i + 1)       # latencies 1 .. 100 ms
    s = latency_stats(reqs)
    assert s["requests"] == 100
    assert s["p50_ms"] == pytest.approx(50.5)
    assert s["p99_ms"] == pytest.approx(99.01)
    assert s["max_ms"] == pytest.approx(100.0)
    assert s["throughput_vis_s"] == pytest.approx(100 * 100 * 5 / (99.1 - 0.0))


def test_major_cycle_model_keeps_bright_pixels():
    import torch
    img = torch.tensor([1.0, 0.4, 0.6, float("nan")], dtype=torch.float32)
    np.testing.assert_allclose(major_cycle_model(img).numpy(), [0.1, 0.0, 0.06, 0.0], rtol=1e-6)
    out = torch.empty_like(img)
    major_cycle_model(img, out=out)
    np.testing.assert_allclose(out.numpy(), [0.1, 0.0, 0.06, 0.0], rtol=1e-6)
