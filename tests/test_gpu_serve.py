"""Served request stream on the GPU: each request's result is the same
alone (on a fresh handle) and in the mix (on pooled handles reused across
requests), bitwise: every kernel of the path is deterministic; and the
requests of the mix do overlap on the device."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone_results():
    import torch
    from wsclean_amd.serve import HandlePool, Server, latency_stats, make_requests, run_request
    reqs = make_requests(6, rate_hz=200.0, seed=17, npix_choices=[256, 512],
                         vis_range=(3e4, 4e5), cycles=2)
    print("served mix:", [(r.rid, r.npix, r.n_rows, r.n_chan) for r in reqs], flush=True)
    alone = []
    for r in reqs:
        img = run_request(r)
        alone.append((img.copy(), r.out.cpu().numpy().copy()))
        r.out.zero_()
    torch.cuda.synchronize()
    # the mix runs on pooled handles (reused across requests of one
    # geometry) and must still equal the fresh-handle results bit for bit
    pool = HandlePool()
    srv = Server(slots=3, pool=pool)
    out = srv.serve(reqs)
    assert pool.hits + pool.misses == len(reqs)
    pool.close()
    for r, (img, vis) in zip(out, alone):
        assert np.isfinite(img).all()
        assert np.abs(img).max() > 0
        assert np.array_equal(r.image, img), "request %d image differs in the mix" % r.rid
        assert np.array_equal(r.out.cpu().numpy(), vis), "request %d prediction differs" % r.rid
    assert srv.peak_running >= 2
    assert any(a.start < b.done and b.start < a.done for a, b in itertools.combinations(out, 2))
    s = latency_stats(out)
    print("served stats:", s, flush=True)
    assert s["p99_ms"] >= s["p50_ms"] > 0


def test_pooled_handle_gives_the_fresh_handle_result():
    """A handle reused by a later request of its geometry (its correction
    table, twiddles and arena kept, its visibility set replaced) gives
    bitwise what a fresh handle gives, also when the later request has more
    visibilities (the arena grows) or fewer."""
    import torch
    from wsclean_amd.serve import HandlePool, make_requests, run_request
    reqs = make_requests(3, rate_hz=100.0, seed=23, npix_choices=[384],
                         vis_range=(5e4, 4e5), cycles=2)
    fresh = []
    for r in reqs:
        img = run_request(r)
        fresh.append((img.copy(), r.out.cpu().numpy().copy()))
        r.out.zero_()
    torch.cuda.synchronize()
    pool = HandlePool()
    for r, (img, vis) in zip(reqs, fresh):
        h = pool.take(r.pool_key())
        got = run_request(r, handle=h)
        pool.put(r.pool_key(), h)
        assert np.array_equal(got, img), "request %d image differs on a pooled handle" % r.rid
        assert np.array_equal(r.out.cpu().numpy(), vis), "request %d prediction differs" % r.rid
        assert r.allocated > 0
    assert (pool.hits, pool.misses) == (2, 1)
    pool.close()


def test_native_server_mix_equals_alone_and_the_python_loop():
    """The library's request loop (wg_server_*): each request's last dirty
    image, prediction and residual are bitwise the same when it runs alone
    on the native server, in the native server's mix, and in the Python
    loop (run_request: model and residual by torch); the mix overlaps."""
    import torch
    from wsclean_amd.serve import NativeServer, latency_stats, make_requests, run_request
    reqs = make_requests(6, rate_hz=200.0, seed=19, npix_choices=[256, 512],
                         vis_range=(3e4, 4e5), cycles=2)
    py = []
    for r in reqs:
        r.keep_residual = True
        img = run_request(r)
        py.append((img.copy(), r.out.cpu().numpy().copy(), r.residual.cpu().numpy().copy()))
        r.out.zero_()
    torch.cuda.synchronize()
    alone = []
    for r in reqs:
        srv = NativeServer(slots=1)
        r.arrival = 0.0
        srv.serve([r])
        alone.append((r.image.copy(), r.out.cpu().numpy().copy(), r.residual.cpu().numpy().copy()))
        r.out.zero_()
    torch.cuda.synchronize()
    for i, r in enumerate(reqs):
        r.arrival = 0.002 * i
    srv = NativeServer(slots=3)
    out = srv.serve(reqs)
    for r, a, p in zip(out, alone, py):
        mix = (r.image, r.out.cpu().numpy(), r.residual.cpu().numpy())
        for k, what in enumerate(("image", "prediction", "residual")):
            assert np.array_equal(mix[k], a[k], equal_nan=True), \
                "request %d %s differs in the native mix" % (r.rid, what)
            assert np.array_equal(a[k], p[k], equal_nan=True), \
                "request %d %s differs from the Python loop" % (r.rid, what)
    st = srv.stats
    print("native served stats:", st, latency_stats(out), flush=True)
    assert st["completed"] == len(reqs) and st["peak_running"] >= 2
    assert st["pool_hits"] + st["pool_misses"] == len(reqs)
    assert any(a.start < b.done and b.start < a.done for a, b in itertools.combinations(out, 2))


def test_native_server_reports_request_errors():
    import torch
    from wsclean_amd import WgError
    from wsclean_amd.serve import NativeServer, make_requests
    reqs = make_requests(2, rate_hz=100.0, seed=29, npix_choices=[256], vis_range=(3e4, 5e4),
                         cycles=1)
    reqs[1].eps = 1e-9   # below what the library serves: that request fails
    with pytest.raises(WgError, match="FP32"):
        NativeServer(slots=2).serve(reqs)
    torch.cuda.synchronize()
    assert reqs[0].done > 0   # the other request still ran to completion


def test_device_image_calls_match_host_calls():
    """RealImageDevice / InitializePredictionDevice give what RealImage /
    InitializePrediction give (trimmed image centred, NaN border)."""
    import torch
    from wsclean_amd import WGriddingGridder_Simple, synth
    rng = np.random.Generator(np.random.PCG64(5))
    enu = synth.antennas_disc(rng, 24, 1500.0)
    uvw_h = synth.earth_rotation_uvw(enu, 30, 10.0, np.deg2rad(-30.0))
    freq = np.array([140e6, 160e6])
    n_rows = uvw_h.shape[0]
    vis_h = (rng.standard_normal((n_rows, 2)) + 1j * rng.standard_normal((n_rows, 2))).astype(
        np.complex64)
    dev = torch.device("cuda:0")
    uvw = torch.from_numpy(uvw_h).to(dev)
    vis = torch.from_numpy(vis_h.view(np.float32).reshape(n_rows, 2, 2)).to(dev)
    w, h, wt, ht = 192, 160, 128, 128
    g = WGriddingGridder_Simple(w, h, wt, ht, 20 * synth.ARCSEC, 20 * synth.ARCSEC, 0.0, 0.0, 1,
                                1e-5)
    g.PrepareVisibilities(uvw, freq, n_rows, 2)
    g.InitializeInversion()
    g.AddInversionPrepared(vis)
    g.FinalizeImage(1.0 / (2 * n_rows))
    host = g.RealImage()
    img = torch.full((w * h,), 7.0, dtype=torch.float32, device=dev)
    # the fill runs on torch's stream, the library on its own
    torch.cuda.synchronize()
    g.RealImageDevice(img)
    assert np.array_equal(img.cpu().numpy(), host, equal_nan=True)
    model = rng.standard_normal(w * h).astype(np.float32)
    out_h = torch.empty_like(vis)
    out_d = torch.empty_like(vis)
    g.InitializePrediction(model)
    g.PredictPrepared(out_h)
    g.InitializePredictionDevice(torch.from_numpy(model).to(dev))
    g.PredictPrepared(out_d)
    assert torch.equal(out_h, out_d)
    with pytest.raises(Exception):
        g.RealImageDevice(torch.empty(w * h - 1, dtype=torch.float32, device=dev))
    g.close()
