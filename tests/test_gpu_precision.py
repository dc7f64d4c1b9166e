"""Precision study on the 4096^2, 8-channel configuration: the 64-antenna
array in a 3 km disc, 300 x 10 s at dec -30 deg, 8 channels over
130-170 MHz, 1 arcsec pixels, 20 power-law point sources on pixel
positions plus noise.  Per channel and per epsilon in {1e-3, 1e-4, 1e-5,
1e-6}, the HIP gridder's dirty image is compared with the FP64 direct FT
(the DirectMSGridder restatement) on 1000 random pixels, and its predicted
visibilities of the source model with the direct FT on 1e5 random rows.
The tolerance is epsilon in relative L2 (doc/source/wgridding.rst:19-20),
for every channel.
"""
import numpy as np
import pytest

from wsclean_amd import WGriddingGridder_Simple, synth

pytestmark = pytest.mark.gpu

N = 4096
DX = synth.ARCSEC
N_PIX = 1000
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
    dirty, pred = [], []
    for c in range(len(freq)):
        f1 = freq[c:c + 1]
        v1 = np.ascontiguousarray(vis[:, c:c + 1])
        dirty.append(oracle.dft_invert_pixels(N, N, DX, DX, 0.0, 0.0, uvw, f1, v1, px, py))
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
