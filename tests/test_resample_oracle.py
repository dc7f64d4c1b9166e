"""The resampler oracle (oracle/resample.py) pinned by closed forms:
trigonometric polynomials below both Nyquist limits are reproduced at the
new pixel positions (output pixel m at input coordinate m N_in / N_out), a
Nyquist cosine keeps its amplitude, and upsampling then downsampling is the
identity."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
from oracle.resample import resample  # noqa: E402


def trig(w, h, terms, sx=1.0, sy=1.0):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    x *= sx
    y *= sy
    f = np.zeros((h, w))
    for a, kx, ky, ph, nx, ny in terms:
        f += a * np.cos(2 * np.pi * (kx * x / nx + ky * y / ny) + ph)
    return f


@pytest.mark.parametrize("wi,hi,wo,ho", [(64, 48, 96, 80), (96, 80, 64, 48), (63, 50, 90, 71),
                                         (90, 71, 63, 50), (64, 64, 64, 64)])
def test_band_limited_reproduced(wi, hi, wo, ho):
    rng = np.random.default_rng(wi * 7 + ho)
    lim_x = (min(wi, wo) - 1) // 2
    lim_y = (min(hi, ho) - 1) // 2
    terms = [(rng.uniform(0.2, 1), int(rng.integers(-lim_x + 1, lim_x)),
              int(rng.integers(-lim_y + 1, lim_y)), rng.uniform(0, 6), wi, hi) for _ in range(6)]
    src = trig(wi, hi, terms)
    want = trig(wo, ho, terms, sx=wi / wo, sy=hi / ho)
    np.testing.assert_allclose(resample(src, wo, ho), want, atol=1e-10)


def test_nyquist_cosine_keeps_its_amplitude():
    # cos(pi x) on 32 pixels, upsampled to 80: cos(pi x 32 / 80)
    src = trig(32, 8, [(1.0, 16, 0, 0.0, 32, 8)])
    want = trig(80, 8, [(1.0, 16, 0, 0.0, 32, 8)], sx=32 / 80)
    np.testing.assert_allclose(resample(src, 80, 8), want, atol=1e-12)
    # a cosine at the output's Nyquist frequency survives downsampling
    src = trig(64, 8, [(0.7, 16, 0, 0.0, 64, 8)])
    want = trig(32, 8, [(0.7, 16, 0, 0.0, 64, 8)], sx=2.0)
    np.testing.assert_allclose(resample(src, 32, 8), want, atol=1e-12)


def test_up_then_down_is_identity():
    rng = np.random.default_rng(3)
    img = rng.normal(size=(40, 56))
    np.testing.assert_allclose(resample(resample(img, 112, 80), 56, 40), img, atol=1e-12)
    np.testing.assert_allclose(resample(resample(img, 75, 61), 56, 40), img, atol=1e-12)
