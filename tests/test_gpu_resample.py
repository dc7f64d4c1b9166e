"""Resampling on the GPU: Fourier interpolation against the oracle
(oracle/resample.py) for smooth and Bluestein line lengths, a band-limited
image checked in closed form at a WSClean-sized pair (1764 -> 2460), and
the FinishInversion / StartPredict resampling entry points."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
from oracle.resample import resample as oracle_resample  # noqa: E402

from wsclean_amd import WGriddingGridder_Simple, synth  # noqa: E402

pytestmark = pytest.mark.gpu


def gridder(n=64, dx=10.0 * synth.ARCSEC):
    return WGriddingGridder_Simple(n, n, n, n, dx, dx, 0.0, 0.0, 1, 1e-4)


@pytest.mark.parametrize("wi,hi,wo,ho", [(64, 48, 96, 80), (96, 80, 64, 48), (63, 50, 90, 71),
                                         (90, 71, 63, 50), (300, 260, 410, 350),
                                         (250, 250, 250, 250), (2000, 16, 1000, 16)])
def test_resample_matches_oracle(wi, hi, wo, ho):
    import torch
    rng = np.random.default_rng(wi + 3 * ho)
    img = rng.normal(size=(hi, wi)).astype(np.float32)
    g = gridder()
    src = torch.from_numpy(img).to("cuda:0")
    out = torch.empty((ho, wo), dtype=torch.float32, device="cuda:0")
    g.ResampleDevice(src, wi, hi, out, wo, ho)
    want = oracle_resample(img, wo, ho)
    got = out.cpu().numpy()
    err = np.abs(got - want).max() / np.abs(want).max()
    print(wi, hi, wo, ho, "max rel err", err, flush=True)
    assert err < 2e-5
    g.close()


def test_band_limited_image_at_wsclean_sizes():
    # small inversion 1764^2 (2^2 3^2 7^2, Bluestein) to the padded 2460^2 image
    import torch
    wi = hi = 1764
    wo = ho = 2460
    y, x = np.mgrid[0:hi, 0:wi].astype(np.float64)
    terms = [(1.0, 37, -12, 0.3), (0.5, -400, 250, 1.1), (0.25, 800, 870, 2.0)]
    f = lambda x, y: sum(a * np.cos(2 * np.pi * (kx * x / wi + ky * y / hi) + p)
                         for a, kx, ky, p in terms)
    img = f(x, y).astype(np.float32)
    g = gridder()
    src = torch.from_numpy(img).to("cuda:0")
    out = torch.empty((ho, wo), dtype=torch.float32, device="cuda:0")
    g.ResampleDevice(src, wi, hi, out, wo, ho)
    yo, xo = np.mgrid[0:ho, 0:wo].astype(np.float64)
    want = f(xo * wi / wo, yo * hi / ho)
    err = np.abs(out.cpu().numpy() - want).max() / 1.75
    print("band-limited max err", err, flush=True)
    assert err < 2e-5
    g.close()


def test_finish_inversion_and_start_predict_resampling():
    import torch
    p = synth.small_problem(seed=71, n_ant=14, n_times=20, n_chan=2, npix=96)
    g = WGriddingGridder_Simple(96, 96, 96, 96, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-5)
    g.InitializeInversion()
    g.AddInversionData(p["uvw"].shape[0], 2, p["uvw"], p["freq"], p["vis"])
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(96, 96)
    big = g.RealImageResampled(120, 120).reshape(120, 120)
    want = oracle_resample(img, 120, 120)
    assert np.abs(big - want).max() < 2e-5 * np.abs(want).max()
    # StartPredict: the 120^2 model resampled back to 96^2 gives the same
    # prediction as predicting from the oracle-resampled model
    rng = np.random.default_rng(2)
    model = np.zeros((120, 120), np.float32)
    model[rng.integers(10, 110, 6), rng.integers(10, 110, 6)] = 1.0
    g.InitializePredictionResampled(model, 120, 120)
    out1 = np.zeros((p["uvw"].shape[0], 2), np.complex64)
    g.PredictVisibilities(p["uvw"].shape[0], 2, p["uvw"], p["freq"], out1)
    g.InitializePrediction(oracle_resample(model, 96, 96).astype(np.float32))
    out2 = np.zeros_like(out1)
    g.PredictVisibilities(p["uvw"].shape[0], 2, p["uvw"], p["freq"], out2)
    assert np.linalg.norm(out1 - out2) < 1e-5 * np.linalg.norm(out2)
    g.close()
