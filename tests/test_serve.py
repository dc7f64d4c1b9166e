"""Host logic of the served request stream (wsclean_amd/serve.py):
admission by slot and by a fixed or measured memory budget, arrival order,
the handle pool, latency statistics.  The per-request work and the device
are host stand-ins here; the GPU tests run the gridder."""
import threading
import time

import numpy as np
import pytest

from wsclean_amd.serve import HandlePool, Server, latency_stats, major_cycle_model


class FakeRequest:
    def __init__(self, rid, arrival, nbytes, work_s, n_vis=1000, cycles=5):
        self.rid, self.arrival, self.nbytes, self.work_s = rid, arrival, nbytes, work_s
        self.n_vis, self.cycles = n_vis, cycles
        self.start = self.done = 0.0
        self.image = None
        self.allocated = 0
        self.key = None

    def hbm_bytes(self):
        return self.nbytes

    def pool_key(self):
        return self.key


class Recorder:
    def __init__(self):
        self.lock = threading.Lock()
        self.running = 0
        self.bytes = 0
        self.max_running = 0
        self.max_bytes = 0
        self.starts = []

    def run(self, req, handle=None, on_alloc=None):
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
    def run(req, handle=None, on_alloc=None):
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


class FakeDevice:
    """Device memory stand-in: requests allocate their bytes in two halves
    (reporting each through on_alloc), idle pooled handles hold theirs."""

    def __init__(self, total):
        self.total = total
        self.lock = threading.Lock()
        self.used = 0
        self.max_used = 0

    def alloc(self, n):
        with self.lock:
            self.used += n
            self.max_used = max(self.max_used, self.used)
            assert self.used <= self.total, "device out of memory"

    def free(self, n):
        with self.lock:
            self.used -= n

    def free_bytes(self):
        with self.lock:
            return self.total - self.used


class FakeHandle:
    def __init__(self, dev, key):
        self.dev, self.key, self.bytes, self.closed = dev, key, 0, False

    def device_bytes(self):
        return self.bytes

    def grow(self, n):
        if n > self.bytes:
            self.dev.alloc(n - self.bytes)
            self.bytes = n

    def close(self):
        assert not self.closed
        self.dev.free(self.bytes)
        self.bytes = 0
        self.closed = True


def device_run(dev):
    def run(req, handle=None, on_alloc=None):
        # half now, half later, as a prepare and a first cycle would
        half = req.nbytes // 2
        if handle is not None:
            handle.grow(half)
        else:
            dev.alloc(half)
        req.allocated = half
        on_alloc()
        time.sleep(req.work_s / 2)
        if handle is not None:
            handle.grow(req.nbytes)
        else:
            dev.alloc(req.nbytes - half)
        req.allocated = req.nbytes
        on_alloc()
        time.sleep(req.work_s / 2)
        if handle is None:
            dev.free(req.nbytes)
        return req.rid
    return run


def test_measured_admission_keeps_the_device_within_its_memory():
    dev = FakeDevice(total=1000)
    sizes = [300, 200, 400, 100, 250, 350, 150, 300, 200, 100]
    reqs = [FakeRequest(i, 0.001 * i, s, 0.03) for i, s in enumerate(sizes)]
    srv = Server(slots=4, run=device_run(dev), free_bytes=dev.free_bytes, reserve=100)
    out = srv.serve(reqs)
    assert [r.rid for r in out] == list(range(len(sizes)))
    # admitted estimates + reserve never exceed the device, so no
    # allocation failed and at least two requests ran together
    assert dev.max_used <= 900
    assert srv.peak_running >= 2
    assert dev.used == 0
    # every admission saw the running requests' outstanding bytes
    assert all(free is not None and to_come is not None for _, free, to_come in srv.admissions)


def test_pool_reuses_handles_of_the_same_geometry_and_evicts_under_pressure():
    dev = FakeDevice(total=1000)
    made = []

    def make(key):
        h = FakeHandle(dev, key)
        made.append(h)
        return h

    pool = HandlePool(make=make, max_idle=8)
    # three requests of geometry A in turn, then one of geometry B that only
    # fits once A's idle handle is closed
    reqs = [FakeRequest(0, 0.00, 600, 0.01), FakeRequest(1, 0.05, 600, 0.01),
            FakeRequest(2, 0.10, 500, 0.01), FakeRequest(3, 0.15, 800, 0.01)]
    for r in reqs[:3]:
        r.key = "A"
    reqs[3].key = "B"
    srv = Server(slots=2, run=device_run(dev), pool=pool, free_bytes=dev.free_bytes, reserve=50)
    srv.serve(reqs)
    assert pool.hits == 2 and pool.misses == 2
    assert [h.key for h in made] == ["A", "B"]
    assert made[0].closed and not made[1].closed   # A evicted for B
    assert pool.evictions == 1 and pool.idle() == 1
    assert dev.max_used <= 1000
    pool.close()
    assert made[1].closed and dev.used == 0


def test_pool_keeps_at_most_max_idle_handles():
    dev = FakeDevice(total=10 ** 9)
    pool = HandlePool(make=lambda key: FakeHandle(dev, key), max_idle=2)
    hs = [pool.take(k) for k in "abc"]
    for k, h in zip("abc", hs):
        pool.put(k, h)
    assert pool.idle() == 2 and hs[0].closed and not hs[2].closed
    assert pool.take("c") is hs[2] and pool.take("b") is hs[1]
    pool.close()
