"""Served request stream on one GPU: continuous batching of w-gridding
requests with HBM admission.

A request is one imaging job: an image geometry, uvw / frequencies and
visibilities resident in HBM, and a number of major cycles.  Each cycle
inverts the residual visibilities into a dirty image, takes a model from
it (a stand-in for the deconvolution: a tenth of the pixels above half the
peak), predicts the model's visibilities and forms
the next residual (WSClean's major cycle, main/wsclean.cpp:1519-1632).  The
request keeps its sorted visibility set (wg_prepare_visibilities), its
grids and its images on the device across the cycles; only the last dirty
image goes to the host.

Requests arrive at their own times (a seeded Poisson stream in
``make_requests``).  A request is admitted, in arrival order, as soon as a
worker slot is free and its device memory fits: the free memory is measured
at each admission, less what the running requests have still to allocate.
Each admitted request drives a gridder handle of its own, taken from a pool
of idle handles of its geometry when there is one (so the grid-correction
table, FFT twiddles and device arena of earlier requests are reused), and
each handle owns a HIP stream and a stream-ordered arena, so the work of
admitted requests overlaps on the device and no handle's allocation or
release waits on another's work.  The library calls release the interpreter
lock while the GPU works, so one Python thread per running request keeps
the slots busy.

A request's result does not depend on what runs beside it: handles share
nothing but the device, and every kernel of the path is deterministic
(integer subgrid accumulation, fixed-order gathers and reductions).
"""
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib, synth
from .gridder import WGriddingGridder_Simple


@dataclass
class Request:
    rid: int
    npix: int
    pixel_rad: float
    eps: float
    uvw: object            # device tensor [n_rows, 3] float64, metres
    freq: np.ndarray       # host [n_chan] float64, Hz
    vis: object            # device tensor [n_rows, n_chan, 2] float32
    out: object            # device tensor, same shape: the last predicted visibilities
    cycles: int = 5
    arrival: float = 0.0   # seconds after the stream starts
    # filled in by the server
    start: float = 0.0
    done: float = 0.0
    image: np.ndarray = field(default=None, repr=False)
    # keep the last cycle's residual visibilities (device tensor) in
    # `residual`, so that the last dirty image can be checked
    keep_residual: bool = False
    residual: object = field(default=None, repr=False)
    # device bytes the running request holds (reported by run_request)
    allocated: int = 0

    @property
    def n_rows(self):
        return int(self.uvw.shape[0])

    @property
    def n_chan(self):
        return len(self.freq)

    @property
    def n_vis(self):
        return self.n_rows * self.n_chan

    def pool_key(self):
        """Geometry of the request's gridder handle (``HandlePool``)."""
        return (self.npix, self.pixel_rad, self.eps)

    def hbm_bytes(self, slab_bytes=8 << 30):
        """Upper estimate of the device memory the request allocates beyond
        its inputs: per visibility the sorted set (keys and permutation with
        their ping-pong copies, coordinates, phasors, sorted and accumulated
        visibilities, grown by the arena's 1/8 rounding) and the residual;
        per pixel the fp64 image, the dirty / correction / host-transfer
        images; the FFT slab (w-planes of at most (2 npix)^2 uv cells plus
        two line buffers, capped by the planner's slab budget); and 64 MiB
        for plan tables."""
        n2 = self.npix * self.npix
        slab = min(slab_bytes, 64 * n2 * 32)
        return int(72 * self.n_vis + 24 * self.n_rows + 32 * n2 + slab + (64 << 20))


def major_cycle_model(img, out=None):
    """The model a cycle predicts: 0.1 x the pixels at or above half the
    image's peak, zero elsewhere (NaN pixels are ignored).  ``img`` is a
    float32 torch tensor on any device; the model goes to ``out`` when
    given."""
    import torch
    peak = torch.nan_to_num(img, nan=float("-inf")).amax()
    if out is None:
        out = torch.empty_like(img)
    torch.mul(img, 0.1, out=out)
    return out.masked_fill_(~(img >= 0.5 * peak), 0.0)


def torch_bytes(req):
    """Device bytes ``run_request`` allocates through torch beside the
    handle: the residual visibilities and two float32 images."""
    return int(8 * req.n_vis + 8 * req.npix * req.npix)


def run_request(req, handle=None, on_alloc=None):
    """Runs one request to completion and returns the last cycle's dirty
    image (float32 [npix, npix], host); ``req.out`` holds the last cycle's
    predicted visibilities.

    ``handle``: a gridder handle of the request's geometry to run on (one
    taken from a ``HandlePool``); None creates one and closes it at the
    end.  After the visibility set is prepared and after the first cycle,
    ``req.allocated`` is set to the device bytes the request holds (the
    handle's arena plus ``torch_bytes``) and ``on_alloc()`` is called, so
    that an admission loop can see how much of its estimate is still to
    come."""
    import torch
    n = req.npix
    dev = req.vis.device
    g = handle
    own = g is None
    if own:
        g = WGriddingGridder_Simple(n, n, n, n, req.pixel_rad, req.pixel_rad, 0.0, 0.0, 1,
                                    req.eps)

    def report():
        req.allocated = g.device_bytes() + torch_bytes(req)
        if on_alloc is not None:
            on_alloc()

    try:
        g.PrepareVisibilities(req.uvw, req.freq, req.n_rows, req.n_chan)
        residual = req.vis.clone()
        img = torch.empty(n * n, dtype=torch.float32, device=dev)
        model = torch.empty_like(img)
        report()
        # torch work of this thread runs on its current stream; the library
        # calls synchronise their own stream before returning, so a host
        # synchronisation of this stream orders the two
        stream = torch.cuda.current_stream(dev)
        stream.synchronize()
        for c in range(req.cycles):
            g.InitializeInversion()
            g.AddInversionPrepared(residual)
            g.FinalizeImage(1.0 / req.n_vis)
            g.RealImageDevice(img)
            major_cycle_model(img, out=model)
            stream.synchronize()
            g.InitializePredictionDevice(model)
            g.PredictPrepared(req.out)
            if c == 0:
                report()
            if c + 1 < req.cycles:
                torch.sub(req.vis, req.out, out=residual)
                stream.synchronize()
        if req.keep_residual:
            req.residual = residual
        return img.view(n, n).cpu().numpy()
    finally:
        if own:
            g.close()


class HandlePool:
    """Idle gridder handles kept between requests, keyed by geometry
    (``Request.pool_key``: image size, pixel size, epsilon).

    A request of a pooled geometry runs on a handle that already holds the
    geometry's grid-correction table, FFT twiddles and a grown device arena,
    instead of building them again.  A handle serves one request at a time:
    ``take`` removes it from the pool and ``put`` returns it.  At most
    ``max_idle`` handles are kept; beyond that, and whenever admission is
    short of device memory (``evict``), the least recently returned handle
    is closed.  Handles are any objects with ``device_bytes()`` and
    ``close()``; ``make(key)`` creates one for a request that finds none."""

    def __init__(self, make=None, max_idle=8):
        self.make = make or _make_handle
        self.max_idle = int(max_idle)
        self._idle = []          # (key, handle), least recently returned first
        self._lock = threading.Lock()
        self.hits = 0
        self.misses = 0
        self.evictions = 0

    def take(self, key):
        """A pooled handle of ``key``, or a new one."""
        with self._lock:
            for i in range(len(self._idle) - 1, -1, -1):
                if self._idle[i][0] == key:
                    self.hits += 1
                    return self._idle.pop(i)[1]
            self.misses += 1
        return self.make(key)

    def put(self, key, handle):
        with self._lock:
            self._idle.append((key, handle))
            drop = self._idle[:-self.max_idle] if len(self._idle) > self.max_idle else []
            self._idle = self._idle[len(drop):]
            self.evictions += len(drop)
        for _, h in drop:
            h.close()

    def evict(self):
        """Closes the least recently returned idle handle; False when the
        pool is empty."""
        with self._lock:
            if not self._idle:
                return False
            _, h = self._idle.pop(0)
            self.evictions += 1
        h.close()
        return True

    def idle(self):
        with self._lock:
            return len(self._idle)

    def close(self):
        while self.evict():
            pass


def _make_handle(key):
    npix, pixel_rad, eps = key
    return WGriddingGridder_Simple(npix, npix, npix, npix, pixel_rad, pixel_rad, 0.0, 0.0, 1,
                                   eps)


def device_free_bytes():
    """Free device memory now: what the device reports free plus what
    torch's caching allocator holds without using (it hands that out
    first)."""
    import torch
    free, _ = torch.cuda.mem_get_info()
    return int(free + torch.cuda.memory_reserved() - torch.cuda.memory_allocated())


class Server:
    """Continuous batching over ``slots`` concurrent requests.

    Admission is in arrival order: the head of the queue starts as soon as
    a slot is free and its device memory fits; later arrivals wait behind
    the head, so large requests are not starved.  A request that would not
    fit even alone runs when nothing else does.

    Memory (``hbm_budget=None``, the default): measured at each admission.
    The head fits when its estimate (``hbm_bytes()``, less what its pooled
    handle already holds) plus what the running requests have still to
    allocate (their estimates less their reported ``allocated`` bytes) plus
    ``reserve`` is at most ``free_bytes()`` (default: the device's free
    memory plus torch's unused cache).  When it does not fit, idle pooled
    handles are closed first, least recently used first.  A fixed
    ``hbm_budget`` instead admits while the running requests' estimates sum
    to at most the budget (no measurement).

    ``pool``: a ``HandlePool``; each admitted request takes a handle of its
    geometry from it and returns it when done (None: a new handle per
    request).  ``run(req, handle, on_alloc)`` runs one request
    (``run_request``)."""

    def __init__(self, slots=4, hbm_budget=None, run=run_request, pool=None, free_bytes=None,
                 reserve=1 << 30):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        self.slots = int(slots)
        self.hbm_budget = hbm_budget
        self.run = run
        self.pool = pool
        self.free_bytes = free_bytes or device_free_bytes
        self.reserve = int(reserve)
        self._cv = threading.Condition()
        self._in_use = 0
        self._running = []
        self.peak_running = 0
        self.peak_bytes = 0
        self.admissions = []     # (rid, measured free bytes, bytes still to come) per admission

    def _notify(self):
        with self._cv:
            self._cv.notify_all()

    def _to_come(self):
        return sum(max(0, need - getattr(r, "allocated", 0)) for r, need in self._running)

    def _admit(self, req):
        need = req.hbm_bytes()
        key = req.pool_key() if self.pool is not None else None
        handle = None
        with self._cv:
            while True:
                if len(self._running) < self.slots:
                    if handle is None and self.pool is not None:
                        handle = self.pool.take(key)
                    have = handle.device_bytes() if handle is not None else 0
                    if self.hbm_budget is not None:
                        fits = self._in_use + need <= self.hbm_budget
                        free = to_come = None
                    else:
                        free, to_come = self.free_bytes(), self._to_come()
                        fits = max(0, need - have) + to_come + self.reserve <= free
                        if not fits and self.pool is not None and self.pool.evict():
                            continue   # measure again with an idle handle closed
                    if fits or not self._running:
                        req.allocated = have
                        self._in_use += need
                        self._running.append((req, need))
                        self.peak_running = max(self.peak_running, len(self._running))
                        self.peak_bytes = max(self.peak_bytes, self._in_use)
                        self.admissions.append((getattr(req, "rid", None), free, to_come))
                        return need, handle
                self._cv.wait(timeout=0.05)

    def _release(self, req, need, handle, ok):
        if handle is not None:
            if ok:
                self.pool.put(req.pool_key(), handle)
            else:
                handle.close()
        with self._cv:
            self._in_use -= need
            self._running = [(r, n) for r, n in self._running if r is not req]
            self._cv.notify_all()

    def serve(self, requests):
        """Replays ``requests`` at their arrival times and returns them in
        arrival order with start / done times (seconds from the stream
        start) and results.  The first error of any request is raised once
        every admitted request has ended."""
        order = sorted(requests, key=lambda r: r.arrival)
        errors = []
        threads = []
        t0 = time.perf_counter()

        def work(req, need, handle):
            ok = False
            try:
                req.start = time.perf_counter() - t0
                req.image = self.run(req, handle, self._notify)
                req.done = time.perf_counter() - t0
                ok = True
            except BaseException as e:
                errors.append(e)
            finally:
                self._release(req, need, handle, ok)

        for req in order:
            delay = req.arrival - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            if errors:
                break
            need, handle = self._admit(req)
            th = threading.Thread(target=work, args=(req, need, handle), daemon=True)
            th.start()
            threads.append(th)
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        return order


class NativeServer:
    """The library's own request loop (wg_server_*, csrc/server.hip): the
    same continuous batching as ``Server`` -- worker slots, admission in
    arrival order by measured free device memory, a per-geometry pool of
    idle gridder handles -- with the major cycles of ``run_request`` (model
    and residual by device kernels) run by C++ worker threads, each with a
    HIP stream of its own.

    ``serve(requests)`` submits each request at its arrival time and returns
    them in arrival order with start / done times (seconds from the stream
    start), ``image`` (the last dirty image), ``out`` filled and, with
    ``keep_residual``, ``residual``.  ``stats`` holds the server's counters
    after the stream."""

    def __init__(self, slots=4, reserve=1 << 30, max_idle=8):
        self.slots = int(slots)
        self.reserve = int(reserve)
        self.max_idle = int(max_idle)
        self.stats = None

    def serve(self, requests):
        import torch
        lib = _lib.load()
        c_void_p = _lib.ctypes.c_void_p
        srv = c_void_p()
        _lib.check(lib.wg_server_create(self.slots, self.reserve, self.max_idle,
                                        _lib.ctypes.byref(srv)))
        order = sorted(requests, key=lambda r: r.arrival)
        pending = []
        try:
            for r in order:
                n = r.npix
                r.image = np.empty((n, n), np.float32)
                if r.keep_residual:
                    r.residual = torch.empty_like(r.vis)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for r in order:
                delay = r.arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                freq = np.ascontiguousarray(r.freq, dtype=np.float64)
                q = _lib.WgRequest(
                    r.npix, r.pixel_rad, r.eps, r.cycles, r.n_rows, r.n_chan,
                    c_void_p(r.uvw.data_ptr()), freq.ctypes.data_as(c_void_p),
                    c_void_p(r.vis.data_ptr()), c_void_p(r.out.data_ptr()),
                    r.image.ctypes.data_as(c_void_p),
                    c_void_p(r.residual.data_ptr()) if r.keep_residual else None)
                rid = _lib.ctypes.c_uint64()
                _lib.check(lib.wg_server_submit(srv, _lib.ctypes.byref(q), _lib.ctypes.byref(rid)))
                pending.append((r, rid.value, time.perf_counter() - t0, freq))
            errors = []
            times = (_lib.ctypes.c_double * 3)()
            for r, rid, t_sub, _ in pending:
                st = lib.wg_server_wait(srv, rid, times)
                if st != 0:
                    errors.append(_lib.WgError(lib.wg_last_error().decode()))
                    continue
                # library times are seconds since the server started: shift
                # them by the submission time seen here
                off = t_sub - times[0]
                r.start, r.done = times[1] + off, times[2] + off
            stats = (_lib.ctypes.c_uint64 * 8)()
            _lib.check(lib.wg_server_stats(srv, stats))
            self.stats = dict(zip(("completed", "peak_running", "pool_hits", "pool_misses",
                                   "handles_closed", "min_free_bytes_at_admission",
                                   "admission_waits", "idle_handles"), list(stats)))
            if errors:
                raise errors[0]
        finally:
            lib.wg_server_destroy(srv)
        return order


def latency_stats(requests):
    """p50 / p99 / mean latency (arrival to completion) and the stream's
    throughput in visibilities per second (every cycle's invert and
    predict counted once, as one grid + degrid pass)."""
    lat = np.array([r.done - r.arrival for r in requests])
    span = max(r.done for r in requests) - min(r.arrival for r in requests)
    nvis = sum(r.n_vis * r.cycles for r in requests)
    return {"requests": len(requests),
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3),
            "max_ms": float(lat.max() * 1e3),
            "throughput_vis_s": float(nvis / span) if span > 0 else None,
            "span_s": float(span),
            "vis_total": int(sum(r.n_vis for r in requests))}


def make_requests(n_requests, rate_hz, seed=1005, device="cuda:0", npix_choices=None,
                  vis_range=(1e5, 5e7), cycles=5, eps=1e-5, specs=None):
    """Seeded served-stream workload: Poisson arrivals at ``rate_hz``; image
    size drawn from ``npix_choices`` (512^2 ... 8192^2), 1 arcsec pixels; a
    random array of 32-256 antennas uniform in a 3 km disc, 1-64 channels
    over 120-180 MHz, 10 s time steps at dec -30 deg until a visibility
    count drawn log-uniformly from ``vis_range`` is reached; 20 power-law
    point sources plus Gaussian noise (sigma 0.01), generated on the
    device.  ``specs`` (optional): one (npix, visibility count) pair per
    request in place of the random image size and count."""
    import torch
    rng = np.random.Generator(np.random.PCG64(seed))
    npix_choices = npix_choices or [512, 1024, 2048, 4096, 8192]
    lib = _lib.load()
    dev = torch.device(device)
    c_void_p = _lib.ctypes.c_void_p
    reqs = []
    t = 0.0
    for rid in range(n_requests):
        t += rng.exponential(1.0 / rate_hz)
        npix = int(rng.choice(npix_choices))
        n_ant = int(rng.integers(32, 257))
        n_chan = int(rng.integers(1, 65))
        target = float(np.exp(rng.uniform(np.log(vis_range[0]), np.log(vis_range[1]))))
        if specs is not None:
            npix, target = int(specs[rid][0]), float(specs[rid][1])
        n_bl = n_ant * (n_ant - 1) // 2
        n_times = max(1, int(round(target / (n_bl * n_chan))))
        enu = synth.antennas_disc(rng, n_ant, 1500.0)
        uvw_h = synth.earth_rotation_uvw(enu, n_times, 10.0, np.deg2rad(-30.0))
        freq = np.linspace(120e6, 180e6, n_chan) if n_chan > 1 else np.array([150e6])
        dx = synth.ARCSEC
        l, m, flux, _, _ = synth.point_sky(rng, 20, npix, dx)
        src = np.zeros((20, 5))
        src[:, 0], src[:, 1], src[:, 3] = l, m, flux
        src[:, 2] = np.sqrt(1.0 - l ** 2 - m ** 2) - 1.0
        uvw = torch.from_numpy(uvw_h).to(dev)
        n_rows = uvw_h.shape[0]
        vis = torch.empty((n_rows, n_chan, 2), dtype=torch.float32, device=dev)
        _lib.check(lib.wgx_synth_visibilities(
            c_void_p(uvw.data_ptr()), n_rows, n_chan, freq.ctypes.data_as(c_void_p), 20,
            src.ctypes.data_as(c_void_p), 0.01, int(seed) * 1000 + rid,
            c_void_p(vis.data_ptr())))
        out = torch.zeros_like(vis)
        reqs.append(Request(rid, npix, dx, eps, uvw, freq, vis, out, cycles, t))
    torch.cuda.synchronize()
    return reqs
