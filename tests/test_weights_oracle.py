"""The imaging-weights / row-preparation oracle (oracle/weights.py) pinned
by closed forms and brute force: uniform weights sum to one per uv cell,
natural weights are one, Briggs interpolates between them, the Gaussian
taper gives the requested image-plane FWHM, the rank filter's window is
the reference's half-open one, GetGrid is point-symmetric about the uv
origin, and the row preparation's PSF, collapse, selection and weight
sums follow from their definitions."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
from oracle import weights as ow  # noqa: E402

C = ow.C_LIGHT


def uv_sample(seed, n_rows=300, n_chan=3):
    rng = np.random.default_rng(seed)
    uvw = rng.uniform(-800, 800, (n_rows, 3))
    freq = np.linspace(120e6, 180e6, n_chan)
    return uvw, freq


def cell_brute(iw, u, v):
    if v < 0:
        u, v = -u, -v
    x = int(np.floor(u * iw.W * iw.psx + iw.W // 2))
    y = int(np.floor(v * iw.H * iw.psy))
    return (x, y) if (0 <= x < iw.W and y < iw.H // 2) else None


def test_cell_mapping_brute_force_and_hermitian():
    iw = ow.ImageWeightsOracle(ow.UNIFORM, 64, 48, 1e-3, 1.2e-3)
    rng = np.random.default_rng(0)
    u = rng.uniform(-600, 600, 500)
    v = rng.uniform(-600, 600, 500)
    x, y = iw.uv_to_xy(u, v)
    ok = iw.inside(x, y)
    for k in range(500):
        b = cell_brute(iw, u[k], v[k])
        assert (b is not None) == bool(ok[k])
        if b:
            assert b == (x[k], y[k])
    x2, y2 = iw.uv_to_xy(-u, -v)
    nz = v != 0
    np.testing.assert_array_equal(x[nz], x2[nz])
    np.testing.assert_array_equal(y[nz], y2[nz])


def test_uniform_weights_sum_to_one_per_cell():
    uvw, freq = uv_sample(1)
    iw = ow.ImageWeightsOracle(ow.UNIFORM, 64, 64, 2e-3, 2e-3)
    wts = np.ones((uvw.shape[0], len(freq), 1), np.float32)
    iw.grid_rows(uvw, freq, wts)
    occupied = int(np.count_nonzero(iw.grid))
    iw.finish()
    lam = C / freq
    w = iw.get_weight(uvw[:, 0:1] / lam, uvw[:, 1:2] / lam)
    assert abs(w.sum() - occupied) < 1e-9
    assert occupied > 50


def test_natural_and_briggs_limits():
    uvw, freq = uv_sample(2)
    wts = np.random.default_rng(3).uniform(0.5, 2.0, (uvw.shape[0], len(freq), 2)).astype(
        np.float32)
    grids = {}
    for mode, r in ((ow.NATURAL, 0.0), (ow.UNIFORM, 0.0), (ow.BRIGGS, 8.0), (ow.BRIGGS, -8.0)):
        iw = ow.ImageWeightsOracle(mode, 64, 64, 2e-3, 2e-3, robustness=r)
        iw.grid_rows(uvw, freq, wts)
        counts = iw.grid.copy()
        iw.finish()
        grids[(mode, r)] = (iw.grid.copy(), counts)
    nat, counts = grids[(ow.NATURAL, 0.0)]
    nz = counts != 0
    assert np.all(nat[nz] == 1.0) and np.all(nat[~nz] == 0.0)
    uni = grids[(ow.UNIFORM, 0.0)][0]
    np.testing.assert_allclose(uni[nz] * counts[nz], 1.0, rtol=1e-15)
    # robust +8: natural; robust -8: proportional to uniform
    np.testing.assert_allclose(grids[(ow.BRIGGS, 8.0)][0][nz], 1.0, rtol=1e-6)
    b = grids[(ow.BRIGGS, -8.0)][0][nz]
    ratio = b / uni[nz]
    np.testing.assert_allclose(ratio, ratio[0], rtol=1e-6)


def test_briggs_uses_sum_of_squares_over_total():
    # two cells with weight sums 1 and 3 (total 4): avg = (1 + 9) / 4
    iw = ow.ImageWeightsOracle(ow.BRIGGS, 8, 8, 1.0, 1.0, robustness=0.5)
    iw.grid[0, 0] = 1.0
    iw.grid[1, 2] = 3.0
    iw.total = 4.0
    iw.finish()
    s2 = (5.0 * 10 ** -0.5) ** 2 / 2.5
    assert iw.grid[0, 0] == pytest.approx(1 / (1 + s2), rel=1e-14)
    assert iw.grid[1, 2] == pytest.approx(1 / (1 + 3 * s2), rel=1e-14)


def test_get_grid_is_point_symmetric_about_the_uv_origin():
    iw = ow.ImageWeightsOracle(ow.NATURAL, 16, 12, 1.0, 1.0)
    iw.grid[:] = np.arange(iw.grid.size).reshape(iw.grid.shape) + 1.0
    img = iw.get_grid()
    H, W = img.shape
    # the uv origin (u = 0, v = 0) is cell (x = W/2, y = 0) -> image (H/2, W/2)
    assert img[H // 2, W // 2] == iw.grid[0, W // 2]
    for y in range(H // 2):
        for x in range(W):
            assert img[H // 2 + y, x] == iw.grid[y, x]
            assert img[H // 2 - 1 - y, W - 1 - x] == iw.grid[y, x]


def test_rank_filter_half_open_window():
    iw = ow.ImageWeightsOracle(ow.NATURAL, 8, 8, 1.0, 1.0)
    iw.grid[:] = 1.0
    iw.grid[1, 3] = 50.0
    # window 4 (d = 2) at (x = 3, y = 1): x in [1, 5), y in [0, 3) = 12 cells
    mean = iw.window_mean(3, 1, 4)
    assert mean == pytest.approx((11 + 50) / 12)
    # a symmetric window [x - 2, x + 2] would hold 15 cells: the mean differs
    assert mean != pytest.approx((14 + 50) / 15)
    iw.rank_filter(2.0, 4)
    assert iw.grid[1, 3] == pytest.approx(2.0 * (11 + 50) / 12)
    assert iw.grid[0, 0] == 1.0


def test_gaussian_taper_gives_the_requested_fwhm():
    beam = 60.0 / 206265.0  # 60 arcsec FWHM in the image
    W = 512
    psx = beam / 8          # image pixels of beam / 8: uv cells of 1 / (W psx)
    iw = ow.ImageWeightsOracle(ow.NATURAL, W, W, psx, psx)
    iw.grid[:] = 1.0
    iw.set_gaussian_taper(beam)
    t = iw.grid[0, :]                     # v = 0 row, u = (x - W/2) / (W psx)
    u = (np.arange(W) - W // 2) / (W * psx)
    ls = np.linspace(0, beam, 401)
    psf = np.array([np.sum(t * np.cos(2 * np.pi * u * l)) for l in ls])
    psf /= psf[0]
    half = ls[np.argmin(np.abs(psf - 0.5))]
    assert 2 * half == pytest.approx(beam, rel=0.01)


def test_tukey_taper_endpoints():
    iw = ow.ImageWeightsOracle(ow.NATURAL, 64, 64, 1.0 / 64, 1.0 / 64)  # uv cell = 1 lambda
    iw.grid[:] = 1.0
    iw.set_tukey_taper(8.0, 20.0)
    u, v = iw._uv_of_cells()
    d = np.sqrt(u * u + v * v)
    assert np.all(iw.grid[d <= 12.0] == 1.0)
    assert np.all(iw.grid[d >= 20.0] == 0.0)   # tukey(0) = 0 at the limit itself
    mid = (d > 12.0) & (d < 20.0)
    assert np.all((iw.grid[mid] > 0) & (iw.grid[mid] < 1))
    # monotone towards the edge
    order = np.argsort(d[mid])
    assert np.all(np.diff(iw.grid[mid][order]) <= 1e-15)


def test_prepare_rows_closed_forms():
    uvw, freq = uv_sample(5, n_rows=40, n_chan=3)
    rng = np.random.default_rng(6)
    vis = (rng.normal(size=(40, 3, 2)) + 1j * rng.normal(size=(40, 3, 2))).astype(np.complex64)
    ones = np.ones((40, 3, 2), np.float32)
    out, st = ow.prepare_rows(uvw, freq, vis, ones)
    np.testing.assert_allclose(out, vis[:, :, 0] + vis[:, :, 1], rtol=1e-6)
    assert st["total_weight"] == 240 and st["gridded_visibility_count"] == 120
    # unselected channel: zero visibilities, no weight
    out, st = ow.prepare_rows(uvw, freq, vis, ones, selected=[True, False, True])
    assert np.all(out[:, 1] == 0) and st["total_weight"] == 160
    # squared / unit visibility weights
    w = np.full((40, 3, 2), 3.0, np.float32)
    _, st = ow.prepare_rows(uvw, freq, vis, w, weighting=ow.VIS_SQUARED)
    assert st["total_weight"] == 240 * 9 and st["visibility_weight_sum"] == 120 * 9
    _, st = ow.prepare_rows(uvw, freq, vis, w, weighting=ow.VIS_UNIT)
    assert st["total_weight"] == 240 and st["max_gridded_weight"] == 1.0
    # PSF at (dl, dm): the visibilities of a unit point source there
    dl, dm = 0.01, -0.02
    out, _ = ow.prepare_rows(uvw, freq, None, ones[:, :, :1], psf=True, psf_dl=dl, psf_dm=dm)
    dn = np.sqrt(1 - dl * dl - dm * dm) - 1
    ph = np.exp(2j * np.pi * (uvw[:, 0:1] * dl + uvw[:, 1:2] * dm + uvw[:, 2:3] * dn) *
                freq[None, :] / C)
    np.testing.assert_allclose(out, ph, atol=2e-6)
    # 4 polarizations: Stokes I = (XX + YY) / 2, model subtracted first
    v4 = (rng.normal(size=(40, 3, 4)) + 1j * rng.normal(size=(40, 3, 4))).astype(np.complex64)
    m4 = (0.1 * v4).astype(np.complex64)
    out, _ = ow.prepare_rows(uvw, freq, v4, np.ones((40, 3, 4), np.float32), model=m4)
    np.testing.assert_allclose(out, 0.45 * (v4[:, :, 0] + v4[:, :, 3]), rtol=1e-5)


def test_prepare_rows_uniform_total_is_occupied_cells():
    uvw, freq = uv_sample(7, n_rows=200, n_chan=2)
    iw = ow.ImageWeightsOracle(ow.UNIFORM, 64, 64, 2e-3, 2e-3)
    ones = np.ones((200, 2, 1), np.float32)
    iw.grid_rows(uvw, freq, ones)
    occupied = int(np.count_nonzero(iw.grid))
    iw.finish()
    _, st = ow.prepare_rows(uvw, freq, np.ones((200, 2, 1), np.complex64), ones,
                            image_weights=iw)
    assert st["total_weight"] == pytest.approx(occupied, rel=1e-6)
