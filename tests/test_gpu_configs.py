"""Accuracy of the benchmark workloads at their own sizes, against the FP64
direct FT (the DirectMSGridder restatement; epsilon means relative RMS
error against it, doc/source/wgridding.rst:19-20):

* the large w-gridding workload (8192^2 at 2 arcsec, 128 antennas, 2880 x
  10 s, 64 channels, 1.498e9 visibilities, epsilon 1e-5) with its default
  plan: invert on random pixels over all visibilities, predict on 1e5
  random visibilities, and the adjoint identity at full size;
* the batched-facets workload (36 facets of 1024^2 ... 3072^2, 16
  channels) with at most 2e5 rows per facet, every facet checked;
* a served stream of 8 requests from 512^2 to 8192^2 and 1e5 to 5e7
  visibilities: each result equals its result alone, and the last dirty
  image and prediction of every request match the direct FT.

The tolerance is epsilon = 1e-5 in relative L2 for every comparison.
"""
import numpy as np
import pytest

from wsclean_amd import WGriddingGridder_Simple, synth
from wsclean_amd.gridder import synth_visibilities

pytestmark = pytest.mark.gpu

EPS = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


# ------------------------------------------------------- large w-gridding

@pytest.fixture(scope="module")
def headline():
    import torch
    print("large w-gridding inputs: generating", flush=True)
    uvw_h, freq, dx, src = synth.large_wgridding_inputs(64, 8192)
    dev = torch.device("cuda:0")
    n_rows, n_chan = uvw_h.shape[0], len(freq)
    uvw = torch.from_numpy(uvw_h).to(dev)
    vis = torch.empty((n_rows, n_chan, 2), dtype=torch.float32, device=dev)
    synth_visibilities(uvw, freq, src, 0.05, 1002, vis)
    torch.cuda.synchronize()
    return dict(uvw_h=uvw_h, uvw=uvw, vis=vis, freq=freq, dx=dx, n_rows=n_rows,
                n_chan=n_chan)


def test_large_wgridding_matches_direct_ft(oracle, headline):
    import torch
    h = headline
    n, dx, freq = 8192, h["dx"], h["freq"]
    n_rows, n_chan = h["n_rows"], h["n_chan"]
    g = WGriddingGridder_Simple(n, n, n, n, dx, dx, 0.0, 0.0, 1, EPS)
    g.PrepareVisibilities(h["uvw"], freq, n_rows, n_chan)
    plan = g.plan()
    print("plan", {k: plan[k] for k in ("W", "sigma", "nu", "P", "Sg", "T")}, flush=True)
    # the benchmark's plan: matrix-core kernels, one sub-slab, fused gather
    assert (plan["W"], plan["nu"], plan["P"]) == (8, 13122, 26)
    assert abs(plan["sigma"] - 1.6) < 1e-9 and plan["P"] <= plan["Sg"]

    # ---- invert: all 1.498e9 visibilities, 64 random pixels
    g.InitializeInversion()
    g.AddInversionPrepared(h["vis"])
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(n, n)
    rng = np.random.Generator(np.random.PCG64(8192))
    xs, ys = rng.integers(0, n, 64), rng.integers(0, n, 64)
    vis_h = h["vis"].cpu().numpy().view(np.complex64).reshape(n_rows, n_chan)
    print("invert done; direct FT on 64 pixels", flush=True)
    ref = oracle.dft_invert_pixels_fast(n, n, dx, dx, 0.0, 0.0, h["uvw_h"], freq, vis_h, xs, ys)
    del vis_h
    e_inv = rel_l2(img[ys, xs], ref)
    print("invert rel_l2 %.3e on 64 pixels" % e_inv, flush=True)

    # ---- predict: a 500-source pixel model, 1e5 random visibilities
    model = np.zeros((n, n))
    _, _, flux, mx, my = synth.point_sky(rng, 500, n, dx, on_pixel=True)
    model[my, mx] = flux
    out = torch.empty_like(h["vis"])
    g.InitializePrediction(model.astype(np.float32).ravel())
    g.PredictPrepared(out)
    sel = rng.choice(n_rows * n_chan, 100000, replace=False)
    got = out.view(-1, 2)[torch.from_numpy(sel).to(out.device)].cpu().numpy()
    got = got[:, 0] + 1j * got[:, 1]
    refv = oracle.dft_predict(n, n, dx, dx, 0.0, 0.0, model, h["uvw_h"], freq, sel=sel)
    e_pred = rel_l2(got, refv)
    print("predict rel_l2 %.3e on 1e5 visibilities" % e_pred, flush=True)

    # ---- adjoint identity at full size: Re<A x, y> = <x, A^H y> with a
    # dense random image x and the workload's visibilities y
    x = rng.standard_normal((n, n)).astype(np.float32)
    g.InitializePrediction(x.ravel())
    g.PredictPrepared(out)
    a = out.view(-1, 2).double()
    y = h["vis"].view(-1, 2).double()
    rhs = float((a[:, 0] * y[:, 0] + a[:, 1] * y[:, 1]).sum())
    lhs = float(np.dot(x.ravel().astype(np.float64), img.ravel().astype(np.float64)))
    scale = float(np.linalg.norm(x) * np.linalg.norm(img))
    e_adj = abs(lhs - rhs) / scale
    print("adjoint |lhs - rhs| / (|x| |A^H y|) = %.3e" % e_adj, flush=True)
    del out, a, y
    g.close()
    torch.cuda.empty_cache()
    assert e_inv < EPS
    assert e_pred < EPS
    assert e_adj < EPS


# --------------------------------------------------------- batched facets

def test_batched_facets_match_direct_ft(oracle):
    import torch
    from wsclean_amd.facets import FacetBatch
    uvw_h, freq, facets, row_off, n_rows, src = synth.batched_facets_workload(
        share=1, seed=1004, max_rows=200000)
    dev = torch.device("cuda:0")
    n_chan = len(freq)
    R = uvw_h.shape[0]
    uvw = torch.from_numpy(uvw_h).to(dev)
    vis = torch.empty((R, n_chan, 2), dtype=torch.float32, device=dev)
    synth_visibilities(uvw, freq, src, 0.05, 1004, vis)
    print("facets: %d facets, %d rows in total" % (len(facets), sum(n_rows)), flush=True)
    fb = FacetBatch(facets)
    fb.Prepare(uvw, freq, row_off, n_rows)
    plan = fb.plan()
    print("atlas", {k: plan["atlas"][k] for k in ("W", "sigma", "ntu", "ntv", "P", "Sg", "Sf")},
          flush=True)
    fb.InitializeInversion()
    fb.AddInversion(vis)
    rng = np.random.Generator(np.random.PCG64(36))
    # the dirty images first: InitializePrediction replaces a facet's image
    imgs = []
    for f, fac in enumerate(facets):
        fb.FinalizeImage(f, 1.0)
        imgs.append(fb.RealImage(f).reshape(fac["height"], fac["width"]))
    models = []
    for f, fac in enumerate(facets):
        nx, ny = fac["width"], fac["height"]
        model = np.zeros((ny, nx))
        k = 16
        model[rng.integers(2, ny - 2, k), rng.integers(2, nx - 2, k)] = rng.uniform(0.1, 1.0, k)
        models.append(model)
        fb.InitializePrediction(f, model.astype(np.float32).ravel())
    n_virtual = int(sum(n_rows)) * n_chan
    out = torch.empty((n_virtual, 2), dtype=torch.float32, device=dev)
    fb.Predict(out)
    base = np.concatenate([[0], np.cumsum(np.asarray(n_rows, np.int64) * n_chan)])
    worst_i = worst_p = 0.0
    for f, fac in enumerate(facets):
        nx, ny = fac["width"], fac["height"]
        kw = (nx, ny, fac["pixel_size_x"], fac["pixel_size_y"], fac["l_shift"], fac["m_shift"])
        r0, nr = row_off[f], n_rows[f]
        img = imgs[f]
        xs, ys = rng.integers(0, nx, 24), rng.integers(0, ny, 24)
        u_f = uvw_h[r0:r0 + nr]
        v_f = vis[r0:r0 + nr].cpu().numpy().view(np.complex64).reshape(nr, n_chan)
        ref = oracle.dft_invert_pixels_fast(*kw, u_f, freq, v_f, xs, ys)
        e_i = rel_l2(img[ys, xs], ref)
        sel = rng.choice(nr * n_chan, 2000, replace=False)
        got = out[base[f]:base[f + 1]].cpu().numpy()[sel]
        refv = oracle.dft_predict(*kw, models[f], u_f, freq, sel=sel)
        e_p = rel_l2(got[:, 0] + 1j * got[:, 1], refv)
        print("facet %2d %4dx%4d rows %6d: invert %.2e predict %.2e" % (f, nx, ny, nr, e_i, e_p),
              flush=True)
        worst_i, worst_p = max(worst_i, e_i), max(worst_p, e_p)
    fb.close()
    del out, vis, uvw
    torch.cuda.empty_cache()
    assert worst_i < EPS
    assert worst_p < EPS


# ------------------------------------------------------ served stream

SPECS = [(512, 1e5), (1024, 4e5), (8192, 2e6), (2048, 5e7), (4096, 8e6), (512, 2e7),
         (8192, 3e7), (1024, 1e6)]


def test_served_stream_sizes(oracle):
    import torch
    from wsclean_amd.serve import HandlePool, Server, latency_stats, make_requests, run_request
    reqs = make_requests(len(SPECS), rate_hz=20.0, seed=1005, specs=SPECS, cycles=5, eps=EPS)
    print("served stream:", [(r.rid, r.npix, r.n_rows, r.n_chan, r.n_vis) for r in reqs],
          flush=True)
    alone = []
    for r in reqs:
        img = run_request(r)
        alone.append((img.copy(), r.out.cpu().numpy().copy()))
        r.out.zero_()
    torch.cuda.synchronize()
    for r in reqs:
        r.keep_residual = True
    pool = HandlePool()
    srv = Server(slots=4, pool=pool)
    out = srv.serve(reqs)
    pool.close()
    s = latency_stats(out)
    print("served stats: p50 %.1f ms p99 %.1f ms, peak running %d"
          % (s["p50_ms"], s["p99_ms"], srv.peak_running), flush=True)
    rng = np.random.Generator(np.random.PCG64(55))
    worst_i = worst_p = 0.0
    for r, (img, vis_alone) in zip(out, alone):
        assert np.array_equal(r.image, img), "request %d image differs in the mix" % r.rid
        pred = r.out.cpu().numpy()
        assert np.array_equal(pred, vis_alone), "request %d prediction differs" % r.rid
        n = r.npix
        uvw_h = r.uvw.cpu().numpy()
        res = r.residual.cpu().numpy().view(np.complex64).reshape(r.n_rows, r.n_chan)
        xs, ys = rng.integers(0, n, 32), rng.integers(0, n, 32)
        ref = oracle.dft_invert_pixels_fast(n, n, r.pixel_rad, r.pixel_rad, 0.0, 0.0, uvw_h,
                                            r.freq, res, xs, ys) / r.n_vis
        e_i = rel_l2(r.image[ys, xs], ref)
        # the last cycle's model: 0.1 x the image pixels at or above half its peak
        model = np.where(r.image >= 0.5 * np.nanmax(r.image), 0.1 * r.image, 0.0)
        sel = rng.choice(r.n_vis, min(r.n_vis, 5000), replace=False)
        refv = oracle.dft_predict(n, n, r.pixel_rad, r.pixel_rad, 0.0, 0.0, model, uvw_h, r.freq,
                                  sel=sel)
        got = pred.reshape(-1, 2)[sel]
        e_p = rel_l2(got[:, 0] + 1j * got[:, 1], refv)
        print("request %d %d^2 %d vis: invert %.2e predict %.2e (model %d pixels)"
              % (r.rid, n, r.n_vis, e_i, e_p, int((model != 0).sum())), flush=True)
        worst_i, worst_p = max(worst_i, e_i), max(worst_p, e_p)
    assert srv.peak_running >= 2
    assert worst_i < EPS
    assert worst_p < EPS
