"""Drop-in capacity of the C-ABI: calls of any size (the library splits them
into row chunks below 2^31 - 1 visibilities and within the free device
memory, as WSClean's chunked AddInversionData / PredictVisibilities calls
expect, wgridder/wgriddingmsgridder.cpp:56-95, :176-178), and the device
memory the accounting reports.  Results are compared with the FP64 direct
FT (tolerance epsilon, doc/source/wgridding.rst:19-20)."""
import numpy as np
import pytest

from wsclean_amd import WGriddingGridder_Simple, synth
from wsclean_amd.gridder import synth_visibilities

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def problem():
    return synth.small_problem(seed=71, n_ant=24, n_times=40, n_chan=4, npix=256,
                               pixel_arcsec=15.0, radius_m=3000.0, n_src=10, noise=0.01)


@pytest.mark.parametrize("chunk", [0, 7000, 1001])
def test_chunked_calls_match_direct_ft(oracle, chunk):
    # host-buffer and device calls split into chunks of at most `chunk`
    # visibilities (0: one chunk) give the direct FT to epsilon
    import torch
    p = problem()
    n, eps = 256, 1e-5
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    g = WGriddingGridder_Simple(n, n, n, n, p["dx"], p["dx"], 0.0, 0.0, 1, eps)
    g.set_max_chunk_vis(chunk)
    g.InitializeInversion()
    g.AddInversionData(n_rows, n_chan, p["uvw"], p["freq"], p["vis"])
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(n, n)
    rng = np.random.default_rng(chunk)
    xs, ys = rng.integers(0, n, 500), rng.integers(0, n, 500)
    ref = oracle.dft_invert_pixels(n, n, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"],
                                   p["vis"], xs, ys)
    e_inv = rel_l2(img[ys, xs], ref)
    # the device variant accumulates the same chunks
    dev = torch.device("cuda:0")
    uvw_d = torch.from_numpy(p["uvw"]).to(dev)
    vis_d = torch.from_numpy(p["vis"].view(np.float32).reshape(n_rows, n_chan, 2).copy()).to(dev)
    g.InitializeInversion()
    g.AddInversionDataDevice(uvw_d, p["freq"], vis_d)
    g.FinalizeImage(1.0)
    img_d = g.RealImage().reshape(n, n)
    model = np.zeros((n, n))
    model[rng.integers(4, n - 4, 20), rng.integers(4, n - 4, 20)] = rng.uniform(0.1, 1, 20)
    g.InitializePrediction(model.astype(np.float32).ravel())
    out = np.zeros((n_rows, n_chan), np.complex64)
    g.PredictVisibilities(n_rows, n_chan, p["uvw"], p["freq"], out)
    out_d = torch.empty_like(vis_d)
    g.PredictVisibilitiesDevice(uvw_d, p["freq"], out_d)
    sel = rng.choice(out.size, 3000, replace=False)
    refv = oracle.dft_predict(n, n, p["dx"], p["dx"], 0.0, 0.0, model, p["uvw"], p["freq"],
                              sel=sel)
    e_pred = rel_l2(out.ravel()[sel], refv)
    got_d = out_d.cpu().numpy().view(np.complex64).reshape(n_rows, n_chan)
    print("chunk %d: invert %.2e predict %.2e" % (chunk, e_inv, e_pred), flush=True)
    assert e_inv < eps and e_pred < eps
    assert rel_l2(img_d, img) < 1e-6
    assert rel_l2(got_d, out) < 1e-6
    g.close()


def test_chunked_correction_callback_equals_one_chunk(oracle):
    # the correction-callback call split into row chunks (its gains stay
    # whole) equals the same call in one chunk (tolerance 1e-6)
    p = problem()
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    rng = np.random.default_rng(3)
    vis = np.repeat(p["vis"][:, :, None], 4, axis=2).astype(np.complex64)
    gains = (1.0 + 0.1 * (rng.normal(size=(3, n_chan, 24, 4))
                          + 1j * rng.normal(size=(3, n_chan, 24, 4)))).astype(np.complex64)
    gains[..., 1:3] -= 1.0
    ant = rng.integers(0, 24, (n_rows, 2)).astype(np.uint32)
    toff = rng.integers(0, 3, n_rows).astype(np.uint64)
    imgs = []
    for chunk in (0, 3000):
        g = WGriddingGridder_Simple(256, 256, 256, 256, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-5)
        g.set_max_chunk_vis(chunk)
        g.InitializeInversion()
        g.AddInversionDataWithCorrectionCallback(4, 4, n_rows, p["uvw"], p["freq"], n_chan, ant,
                                                 vis, toff, gains, polarization=1)
        g.FinalizeImage(1.0)
        imgs.append(g.RealImage())
        g.close()
    assert rel_l2(imgs[1], imgs[0]) < 1e-6


def test_host_calls_equal_device_calls_bitwise():
    # host-buffer calls stage transfers above 4 MiB through page-locked
    # 64 MiB chunks: 1.09e7 visibilities (83 MiB, two chunks, the
    # second partial; uvw 2.5 MiB, direct) give the device calls' image and
    # predictions bit for bit, twice in a row on one handle
    import torch
    rng = np.random.Generator(np.random.PCG64(77))
    enu = synth.antennas_disc(rng, 40, 3000.0)
    uvw = synth.earth_rotation_uvw(enu, 140, 30.0, np.deg2rad(-30.0))
    freq = np.linspace(140e6, 160e6, 100)
    n_rows, n_chan = uvw.shape[0], len(freq)
    assert n_rows * n_chan * 8 > (64 << 20) and n_rows * 24 < (4 << 20)
    dev = torch.device("cuda:0")
    duvw = torch.from_numpy(uvw).to(dev)
    src = np.array([[0.002, -0.001, 0.0, 1.0, 0.0], [-0.003, 0.002, 0.0, 0.5, 0.0]])
    src[:, 2] = np.sqrt(1.0 - src[:, 0] ** 2 - src[:, 1] ** 2) - 1.0
    dvis = torch.empty((n_rows, n_chan, 2), dtype=torch.float32, device=dev)
    synth_visibilities(duvw, freq, src, 0.05, 77, dvis)
    torch.cuda.synchronize()
    vis = dvis.cpu().numpy().view(np.complex64).reshape(n_rows, n_chan).copy()
    n, dx = 512, 10.0 * synth.ARCSEC
    model = np.zeros((n, n), np.float32)
    model[rng.integers(8, n - 8, 12), rng.integers(8, n - 8, 12)] = 1.0
    g = WGriddingGridder_Simple(n, n, n, n, dx, dx, 0.0, 0.0, 1, 1e-5)
    for _ in range(2):
        g.InitializeInversion()
        g.AddInversionData(n_rows, n_chan, uvw, freq, vis)
        g.FinalizeImage(1.0)
        img_h = g.RealImage()
        g.InitializeInversion()
        g.AddInversionDataDevice(duvw, freq, dvis)
        g.FinalizeImage(1.0)
        np.testing.assert_array_equal(img_h, g.RealImage())
        g.InitializePrediction(model.ravel())
        out_h = np.zeros((n_rows, n_chan), np.complex64)
        g.PredictVisibilities(n_rows, n_chan, uvw, freq, out_h)
        out_d = torch.empty_like(dvis)
        g.PredictVisibilitiesDevice(duvw, freq, out_d)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out_h.view(np.float32).reshape(n_rows, n_chan, 2),
                                      out_d.cpu().numpy())
    g.close()


def test_device_memory_within_accounting():
    # the device memory a handle holds after a call stays within what
    # ConstantMemoryUsage + PerVisibilityMemoryUsage x visibilities report
    p = problem()
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    g = WGriddingGridder_Simple(256, 256, 256, 256, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-5)
    g.set_tuning(slab_bytes=32 << 20)   # a small FFT-slab budget makes the bound tight
    const, per_vis = g.ConstantMemoryUsage(), g.PerVisibilityMemoryUsage()
    g.InitializeInversion()
    g.AddInversionData(n_rows, n_chan, p["uvw"], p["freq"], p["vis"])
    out = np.zeros((n_rows, n_chan), np.complex64)
    g.PredictVisibilities(n_rows, n_chan, p["uvw"], p["freq"], out)
    held = g.device_bytes()
    print("device bytes %d, accounting %d + %d x %d" % (held, const, per_vis, n_rows * n_chan),
          flush=True)
    assert per_vis >= 40
    assert held <= const + per_vis * n_rows * n_chan
    g.close()


def test_more_than_2_31_visibilities(oracle):
    # 34.1e6 rows x 64 channels = 2.18e9 visibilities in one device call: the
    # library splits it below 2^31 - 1; the image matches the direct FT on
    # 16 pixels over all visibilities, and predictions at flat indices above
    # 2^31 match it too
    import torch
    rng = np.random.Generator(np.random.PCG64(2031))
    enu = synth.antennas_logradial(rng, 128, 10000.0)
    uvw_h = synth.earth_rotation_uvw(enu, 4200, 7.0, np.deg2rad(-30.0))
    freq = np.linspace(120e6, 180e6, 64)
    n_rows, n_chan = uvw_h.shape[0], 64
    assert n_rows * n_chan > 2 ** 31
    dx, n = 2.0 * synth.ARCSEC, 2048
    l, m, flux, _, _ = synth.point_sky(rng, 20, n, dx)
    src = np.zeros((20, 5))
    src[:, 0], src[:, 1], src[:, 3] = l, m, flux
    src[:, 2] = np.sqrt(1.0 - l ** 2 - m ** 2) - 1.0
    dev = torch.device("cuda:0")
    uvw = torch.from_numpy(uvw_h).to(dev)
    vis = torch.empty((n_rows, n_chan, 2), dtype=torch.float32, device=dev)
    synth_visibilities(uvw, freq, src, 0.05, 2031, vis)
    print("2^31+ case: %d visibilities" % (n_rows * n_chan), flush=True)
    g = WGriddingGridder_Simple(n, n, n, n, dx, dx, 0.0, 0.0, 1, 1e-5)
    g.InitializeInversion()
    g.AddInversionDataDevice(uvw, freq, vis)
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(n, n)
    xs, ys = rng.integers(0, n, 16), rng.integers(0, n, 16)
    vis_h = vis.cpu().numpy().view(np.complex64).reshape(n_rows, n_chan)
    ref = oracle.dft_invert_pixels_fast(n, n, dx, dx, 0.0, 0.0, uvw_h, freq, vis_h, xs, ys)
    del vis_h
    e_inv = rel_l2(img[ys, xs], ref)
    model = np.zeros((n, n))
    model[rng.integers(8, n - 8, 30), rng.integers(8, n - 8, 30)] = rng.uniform(0.1, 1.0, 30)
    g.InitializePrediction(model.astype(np.float32).ravel())
    g.PredictVisibilitiesDevice(uvw, freq, vis)   # overwrites the input buffer
    sel = np.concatenate([rng.choice(2 ** 31, 5000, replace=False),
                          2 ** 31 + rng.choice(n_rows * n_chan - 2 ** 31, 5000, replace=False)])
    got = vis.view(-1, 2)[torch.from_numpy(sel).to(dev)].cpu().numpy()
    refv = oracle.dft_predict(n, n, dx, dx, 0.0, 0.0, model, uvw_h, freq, sel=sel)
    e_pred = rel_l2(got[:, 0] + 1j * got[:, 1], refv)
    print("2^31+ case: invert %.2e predict %.2e" % (e_inv, e_pred), flush=True)
    g.close()
    del vis, uvw
    torch.cuda.empty_cache()
    assert e_inv < 1e-5
    assert e_pred < 1e-5


def test_grid_sides_above_the_lds_fft(oracle):
    # a 12288^2 image at eps 1e-5 pads to grid sides above one LDS FFT line
    # (16384 points): the plane transforms run in two passes over R
    # interleaved subsequences; invert and predict against the direct FT
    p = synth.small_problem(seed=73, n_ant=24, n_times=30, n_chan=2, npix=12288,
                            pixel_arcsec=1.0, radius_m=5000.0, n_src=12, noise=0.01)
    n, eps = 12288, 1e-5
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    g = WGriddingGridder_Simple(n, n, n, n, p["dx"], p["dx"], 0.0, 0.0, 1, eps)
    g.InitializeInversion()
    g.AddInversionData(n_rows, n_chan, p["uvw"], p["freq"], p["vis"])
    g.FinalizeImage(1.0)
    plan = g.plan()
    print("long-line plan", {k: plan[k] for k in ("W", "sigma", "nu", "nv", "P", "Sf")},
          flush=True)
    assert plan["nu"] > 16384 and plan["nv"] > 16384
    img = g.RealImage().reshape(n, n)
    rng = np.random.default_rng(73)
    xs, ys = rng.integers(0, n, 300), rng.integers(0, n, 300)
    ref = oracle.dft_invert_pixels(n, n, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"],
                                   p["vis"], xs, ys)
    e_inv = rel_l2(img[ys, xs], ref)
    del img
    model = np.zeros((n, n), np.float32)
    k = 24
    model[rng.integers(8, n - 8, k), rng.integers(8, n - 8, k)] = rng.uniform(-1.0, 1.0, k)
    g.InitializePrediction(model.ravel())
    out = np.zeros((n_rows, n_chan), np.complex64)
    g.PredictVisibilities(n_rows, n_chan, p["uvw"], p["freq"], out)
    g.close()
    sel = rng.choice(out.size, 2000, replace=False)
    refv = oracle.dft_predict(n, n, p["dx"], p["dx"], 0.0, 0.0, model.astype(np.float64),
                              p["uvw"], p["freq"], sel=sel)
    e_pred = rel_l2(out.ravel()[sel], refv)
    print("long-line invert %.3e predict %.3e" % (e_inv, e_pred), flush=True)
    assert e_inv < eps and e_pred < eps
