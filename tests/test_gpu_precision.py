"""Precision study on the 4096^2, 8-channel configuration: the 64-antenna
array in a 3 km disc, 300 x 10 s at dec -30 deg, 8 channels over
130-170 MHz, 1 arcsec pixels, 20 power-law point sources on pixel
positions plus noise.  Per channel and per epsilon in {1e-3, 1e-4, 1e-5,
1e-6}, the HIP gridder's dirty image is compared with the FP64 direct FT
(the DirectMSGridder restatement) on 1e4 random pixels, and its predicted
visibilities of the source model with the direct FT on 1e5 random rows.
The tolerance is epsilon in relative L2 (doc/source/wgridding.rst:19-20),
for every channel.  At epsilon 1e-5 and 1e-6 the first and last channels'
whole dirty images, and the first channel's predictions, are also compared
with the FP64 CPU w-gridder of the oracle on the same kernel and grid.
"""
import numpy as np
import pytest

from wsclean_amd import WGriddingGridder_Simple, synth

pytestmark = pytest.mark.gpu

N = 4096
DX = synth.ARCSEC
N_PIX = 10000
N_VIS = 100000
EPS = [1e-3, 1e-4, 1e-5, 1e-6]


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.fixture(scope="module")
def study(oracle):
    rng = np.random.Generator(np.random.PCG64(4096))
    enu = synth.antennas_disc(rng, 64, 3000.0)
    uvw = synth.earth_rotation_uvw(enu, 300, 10.0, np.deg2rad(-30.0))
    freq = np.linspace(130e6, 170e6, 8)
    l, m, flux, xs, ys = synth.point_sky(rng, 20, N, DX, on_pixel=True)
    vis = synth.add_noise(rng, synth.predict_sources(uvw, freq, l, m, flux), 0.01)
    vis = vis.astype(np.complex64)
    px = rng.integers(0, N, N_PIX)
    py = rng.integers(0, N, N_PIX)
    model = np.zeros((N, N))
    model[ys, xs] = flux
    sel = rng.choice(uvw.shape[0], N_VIS, replace=False)
    # every channel's sums in one pass over the visibilities
    per_chan = oracle.dft_invert_pixels_per_channel(N, N, DX, DX, 0.0, 0.0, uvw, freq, vis, px,
                                                    py)
    dirty, pred = [], []
    for c in range(len(freq)):
        f1 = freq[c:c + 1]
        dirty.append(per_chan[:, c])
        pred.append(oracle.dft_predict(N, N, DX, DX, 0.0, 0.0, model, uvw, f1, sel=sel))
    print("precision study: %d rows, %d channels" % (uvw.shape[0], len(freq)), flush=True)
    return dict(uvw=uvw, freq=freq, vis=vis, px=px, py=py, model=model, sel=sel, dirty=dirty,
                pred=pred)


@pytest.mark.parametrize("eps", EPS)
def test_precision_study_per_channel(study, eps):
    uvw, freq, vis = study["uvw"], study["freq"], study["vis"]
    n_rows = uvw.shape[0]
    g = WGriddingGridder_Simple(N, N, N, N, DX, DX, 0.0, 0.0, 1, eps)
    model32 = study["model"].astype(np.float32).ravel()
    err_i, err_p = [], []
    for c in range(len(freq)):
        f1 = freq[c:c + 1]
        g.InitializeInversion()
        g.AddInversionData(n_rows, 1, uvw, f1, np.ascontiguousarray(vis[:, c:c + 1]))
        g.FinalizeImage(1.0)
        img = g.RealImage().reshape(N, N)
        err_i.append(rel_l2(img[study["py"], study["px"]], study["dirty"][c]))
        g.InitializePrediction(model32)
        out = np.zeros((n_rows, 1), np.complex64)
        g.PredictVisibilities(n_rows, 1, uvw, f1, out)
        err_p.append(rel_l2(out[study["sel"], 0], study["pred"][c]))
    plan = g.plan()
    print("eps %.0e W %d sigma %.2f P %d: invert max %.2e mean %.2e, predict max %.2e mean %.2e"
          % (eps, plan["W"], plan["sigma"], plan["P"], max(err_i), np.mean(err_i), max(err_p),
             np.mean(err_p)), flush=True)
    assert max(err_i) < eps
    assert max(err_p) < eps


@pytest.mark.parametrize("eps", [1e-5, 1e-6])
def test_precision_study_against_fp64_wgridder(study, oracle, eps):
    # the oracle's FP64 CPU w-gridder on the same kernel and grid (W, beta,
    # sigma, nu, nv): whole dirty images of the first and last channel, and
    # the first channel's predictions on the 1e5 selected rows
    uvw, freq, vis = study["uvw"], study["freq"], study["vis"]
    n_rows = uvw.shape[0]
    g = WGriddingGridder_Simple(N, N, N, N, DX, DX, 0.0, 0.0, 1, eps)
    errs = []
    for c in (0, len(freq) - 1):
        f1, v1 = freq[c:c + 1], np.ascontiguousarray(vis[:, c:c + 1])
        g.InitializeInversion()
        g.AddInversionData(n_rows, 1, uvw, f1, v1)
        g.FinalizeImage(1.0)
        img = g.RealImage().reshape(N, N)
        pl = g.plan()
        kw = dict(W=pl["W"], beta=pl["beta"], sigma=pl["sigma"], nu=pl["nu"], nv=pl["nv"])
        ref = oracle.cpu_wgrid_invert(N, N, DX, DX, 0.0, 0.0, uvw, f1, v1, **kw)
        errs.append(rel_l2(img, ref))
    f1 = freq[:1]
    g.InitializePrediction(study["model"].astype(np.float32).ravel())
    out = np.zeros((n_rows, 1), np.complex64)
    g.PredictVisibilities(n_rows, 1, uvw, f1, out)
    refv = oracle.cpu_wgrid_predict(N, N, DX, DX, 0.0, 0.0, study["model"], uvw, f1, **kw)
    err_p = rel_l2(out[study["sel"], 0], refv[study["sel"], 0])
    print("eps %.0e W %d sigma %.2f: invert vs FP64 w-gridder (channels 0, %d) %.2e %.2e, "
          "predict %.2e" % (eps, kw["W"], kw["sigma"], len(freq) - 1, errs[0], errs[1], err_p),
          flush=True)
    assert max(errs) < eps
    assert err_p < eps


def test_plain_invert_predict_2048(oracle):
    # The plain CPU-runnable configuration: 2048^2 at 1 arcsec, one channel
    # at 150 MHz, 64 antennas in a 3 km disc (2016 baselines), 300 x 10 s at
    # dec -30 deg, 20 power-law point sources plus noise 0.01, eps 1e-4.
    # Invert on 2000 random pixels and predict on 2e4 random rows against
    # the FP64 direct FT.
    n, eps = 2048, 1e-4
    rng = np.random.Generator(np.random.PCG64(2048))
    enu = synth.antennas_disc(rng, 64, 3000.0)
    uvw = synth.earth_rotation_uvw(enu, 300, 10.0, np.deg2rad(-30.0))
    assert uvw.shape[0] == 2016 * 300
    freq = np.array([150e6])
    l, m, flux, xs, ys = synth.point_sky(rng, 20, n, DX, on_pixel=True)
    vis = synth.add_noise(rng, synth.predict_sources(uvw, freq, l, m, flux), 0.01)
    vis = vis.astype(np.complex64)
    g = WGriddingGridder_Simple(n, n, n, n, DX, DX, 0.0, 0.0, 1, eps)
    g.InitializeInversion()
    g.AddInversionData(uvw.shape[0], 1, uvw, freq, vis)
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(n, n)
    px, py = rng.integers(0, n, 2000), rng.integers(0, n, 2000)
    ref = oracle.dft_invert_pixels(n, n, DX, DX, 0.0, 0.0, uvw, freq, vis, px, py)
    err_i = rel_l2(img[py, px], ref)
    model = np.zeros((n, n))
    model[ys, xs] = flux
    g.InitializePrediction(model.astype(np.float32).ravel())
    out = np.zeros((uvw.shape[0], 1), np.complex64)
    g.PredictVisibilities(uvw.shape[0], 1, uvw, freq, out)
    sel = rng.choice(uvw.shape[0], 20000, replace=False)
    err_p = rel_l2(out[sel, 0], oracle.dft_predict(n, n, DX, DX, 0.0, 0.0, model, uvw, freq,
                                                    sel=sel))
    print("2048^2 plain: invert %.2e predict %.2e plan %s" % (err_i, err_p, g.plan()), flush=True)
    assert err_i < eps
    assert err_p < eps
