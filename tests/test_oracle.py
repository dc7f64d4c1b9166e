"""CPU tests that pin the oracle (FP64 direct FT restating DirectMSGridder).

Each check is chosen so that a plausible mistake fails it: a transposed or
mirrored pixel mapping fails the literal-wrap comparison and the
point-source peak; a wrong predict sign fails the adjoint test and the
delta-image closed form; a dropped w term fails the w != 0 point source.
"""
import os

import numpy as np
import pytest

from wsclean_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))


def _rand_problem(rng, n_rows=7, n_chan=3, scale=300.0):
    uvw = rng.normal(0, scale, (n_rows, 3))
    freq = 140e6 + 10e6 * np.arange(n_chan)
    vis = (rng.normal(size=(n_rows, n_chan)) + 1j * rng.normal(size=(n_rows, n_chan)))
    return uvw, freq, vis.astype(np.complex64)


@pytest.mark.parametrize("w,h,ls,ms", [(8, 8, 0.0, 0.0), (10, 6, 0.0, 0.0),
                                       (8, 12, 0.003, -0.002)])
def test_dft_matches_literal_directmsgridder(oracle, w, h, ls, ms):
    rng = np.random.default_rng(5)
    uvw, freq, vis = _rand_problem(rng)
    dx, dy = 0.01, 0.012
    a = oracle.dft_invert(w, h, dx, dy, ls, ms, uvw, freq, vis)
    b = oracle.dft_invert_reference_wrap(w, h, dx, dy, ls, ms, uvw, freq, vis)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-10)


def test_point_source_peak_equals_flux_times_count(oracle):
    # A source exactly on pixel (x0, y0) with V = S exp(+2 pi i (ul+vm+w(n-1)))
    # gives I(x0, y0) = S * N_vis (natural weights before 1/ImageWeight).
    p = synth.small_problem(seed=11, n_src=1, on_pixel=True, npix=32, n_chan=2)
    S = p["flux"][0]
    img = oracle.dft_invert(32, 32, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"], p["vis"])
    nvis = p["vis"].size
    x0, y0 = int(p["xs"][0]), int(p["ys"][0])
    assert abs(img[y0, x0] / nvis - S) < 1e-6 * S
    assert np.argmax(img) == y0 * 32 + x0
    # l decreases with x, m increases with y (CDELT1 < 0, CDELT2 > 0)
    l0 = (16 - x0) * p["dx"]
    m0 = (y0 - 16) * p["dx"]
    assert abs(l0 - p["l"][0]) < 1e-15 and abs(m0 - p["m"][0]) < 1e-15


def test_w_term_matters_and_is_matched(oracle):
    # Large w and a far-from-centre source: dropping the w(n-1) term would
    # move the peak; the oracle must still give S * N at the source pixel.
    rng = np.random.default_rng(3)
    n = 64
    dx = np.deg2rad(0.5)
    x0, y0 = 5, 58
    l0, m0 = (n // 2 - x0) * dx, (y0 - n // 2) * dx
    uvw = rng.normal(0, 30.0, (40, 3))
    uvw[:, 2] *= 20.0
    freq = np.array([150e6])
    s = freq[0] / synth.C_LIGHT
    nm1 = np.sqrt(1 - l0 ** 2 - m0 ** 2) - 1
    vis = np.exp(2j * np.pi * s * (uvw[:, 0] * l0 + uvw[:, 1] * m0 + uvw[:, 2] * nm1))
    img = oracle.dft_invert(n, n, dx, dx, 0.0, 0.0, uvw, freq, vis[:, None].astype(np.complex64))
    assert abs(img[y0, x0] - 40.0) < 1e-4
    assert np.argmax(img) == y0 * n + x0


def test_delta_image_predict_closed_form(oracle):
    rng = np.random.default_rng(9)
    uvw, freq, _ = _rand_problem(rng, n_rows=11, n_chan=2)
    n, dx = 16, 0.02
    img = np.zeros((n, n))
    x0, y0, S = 3, 12, 2.5
    img[y0, x0] = S
    V = oracle.dft_predict(n, n, dx, dx, 0.001, 0.002, img, uvw, freq)
    l = (n // 2 - x0) * dx + 0.001
    m = (y0 - n // 2) * dx + 0.002
    nm1 = np.sqrt(1 - l * l - m * m) - 1
    s = freq / synth.C_LIGHT
    ph = 2 * np.pi * (uvw[:, 0:1] * s * l + uvw[:, 1:2] * s * m + uvw[:, 2:3] * s * nm1)
    np.testing.assert_allclose(V, S * np.exp(1j * ph), rtol=1e-12, atol=1e-12)


def test_adjoint_dot_product(oracle):
    rng = np.random.default_rng(21)
    uvw, freq, y = _rand_problem(rng, n_rows=9, n_chan=2)
    n, dx = 12, 0.015
    x = rng.normal(size=(n, n))
    Ax = oracle.dft_predict(n, n, dx, dx, 0.0, 0.0, x, uvw, freq)
    AHy = oracle.dft_invert(n, n, dx, dx, 0.0, 0.0, uvw, freq, y)
    lhs = np.sum(x * AHy)
    rhs = np.real(np.sum(y.astype(np.complex128) * np.conj(Ax)))
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_conjugate_symmetry_and_frequency_reversal(oracle):
    rng = np.random.default_rng(4)
    uvw, freq, vis = _rand_problem(rng)
    n, dx = 10, 0.01
    a = oracle.dft_invert(n, n, dx, dx, 0.0, 0.0, uvw, freq, vis)
    b = oracle.dft_invert(n, n, dx, dx, 0.0, 0.0, -uvw, freq, np.conj(vis))
    c = oracle.dft_invert(n, n, dx, dx, 0.0, 0.0, uvw, freq[::-1].copy(),
                          np.ascontiguousarray(vis[:, ::-1]))
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(a, c, rtol=1e-10, atol=1e-10)


def test_pixel_subset_and_vis_subset(oracle):
    rng = np.random.default_rng(8)
    uvw, freq, vis = _rand_problem(rng)
    n, dx = 10, 0.01
    full = oracle.dft_invert(n, n, dx, dx, 0.0, 0.0, uvw, freq, vis)
    xs, ys = np.array([0, 3, 9]), np.array([1, 5, 9])
    sub = oracle.dft_invert_pixels(n, n, dx, dx, 0.0, 0.0, uvw, freq, vis, xs, ys)
    np.testing.assert_allclose(sub, full[ys, xs], rtol=1e-12)
    img = rng.normal(size=(n, n))
    V = oracle.dft_predict(n, n, dx, dx, 0.0, 0.0, img, uvw, freq)
    sel = np.array([0, 5, 20])
    Vs = oracle.dft_predict(n, n, dx, dx, 0.0, 0.0, img, uvw, freq, sel=sel)
    np.testing.assert_allclose(Vs, V.ravel()[sel], rtol=1e-12)


@pytest.mark.parametrize("uniform", [True, False])
def test_fast_pixel_dft_matches_direct(oracle, uniform):
    # the row-parallel recurrence form against the per-term cos / sin form:
    # 64 evenly spaced channels (the recurrence's longest chain) and, with
    # one channel moved, the per-channel fallback; a wrong step phasor or a
    # dropped channel fails at far more than 1e-11
    p = synth.small_problem(seed=41, n_ant=12, n_times=12, n_chan=64, npix=128,
                            pixel_arcsec=30.0, radius_m=4000.0, df=0.9e6, noise=0.1)
    freq = p["freq"].copy()
    if not uniform:
        freq[17] += 0.3e6
    rng = np.random.default_rng(2)
    xs, ys = rng.integers(0, 128, 40), rng.integers(0, 128, 40)
    vis = p["vis"].copy()
    vis[3, 5] = 0
    ref = oracle.dft_invert_pixels(128, 128, p["dx"], p["dx"], 0.01, -0.02, p["uvw"], freq,
                                   vis, xs, ys)
    got, used = oracle.dft_invert_pixels_fast(128, 128, p["dx"], p["dx"], 0.01, -0.02,
                                              p["uvw"], freq, vis, xs, ys, nthreads=3,
                                              return_mode=True)
    assert used == uniform
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("uniform", [True, False])
def test_per_channel_pixel_dft_matches_direct(oracle, uniform):
    # one sum per channel against the per-term form on that channel alone:
    # a channel's sum landing in another channel's column, or a phasor step
    # applied out of order, fails at far more than 1e-11
    p = synth.small_problem(seed=43, n_ant=12, n_times=12, n_chan=16, npix=128,
                            pixel_arcsec=30.0, radius_m=4000.0, df=2e6, noise=0.1)
    freq = p["freq"].copy()
    if not uniform:
        freq[5] += 0.4e6
    rng = np.random.default_rng(3)
    xs, ys = rng.integers(0, 128, 30), rng.integers(0, 128, 30)
    got, used = oracle.dft_invert_pixels_per_channel(128, 128, p["dx"], p["dx"], 0.01, -0.02,
                                                     p["uvw"], freq, p["vis"], xs, ys,
                                                     nthreads=3, return_mode=True)
    assert used == uniform
    assert got.shape == (30, 16)
    for c in range(16):
        ref = oracle.dft_invert_pixels(128, 128, p["dx"], p["dx"], 0.01, -0.02, p["uvw"],
                                       freq[c:c + 1], np.ascontiguousarray(p["vis"][:, c:c + 1]),
                                       xs, ys)
        np.testing.assert_allclose(got[:, c], ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_golden_fixture(oracle):
    g = np.load(os.path.join(HERE, "golden", "dft_small.npz"))
    img = oracle.dft_invert(int(g["nx"]), int(g["ny"]), float(g["dx"]), float(g["dy"]),
                            float(g["l_shift"]), float(g["m_shift"]), g["uvw"], g["freq"],
                            g["vis"])
    np.testing.assert_allclose(img, g["image"], rtol=1e-10, atol=1e-10)
    V = oracle.dft_predict(int(g["nx"]), int(g["ny"]), float(g["dx"]), float(g["dy"]),
                           float(g["l_shift"]), float(g["m_shift"]), g["model"], g["uvw"],
                           g["freq"])
    np.testing.assert_allclose(V, g["pred"], rtol=1e-10, atol=1e-10)


def test_es_table_phi_hat_against_adaptive_quadrature(oracle):
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "gen_es_table", os.path.join(os.path.dirname(HERE), "wsclean_amd", "csrc",
                                     "gen_es_table.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    for W, beta in [(8, 18.0), (10, 23.0), (6, 13.5)]:
        ts = np.array([0.0, 0.2, 0.41, 1.3])
        got = gen.es_phi_hat(W, np.array([beta]), ts)[0]
        ref = np.array([oracle.es_phi_hat(W, beta, t) for t in ts])
        np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-12)


def test_es_table_error_estimate_brute_force(oracle):
    # The tabulated 1-D aliasing error equals a brute-force evaluation of
    # sqrt(sum_{k != 0} Phi(t+k)^2)/Phi(t) maximised over |t| <= 1/(2 sigma).
    rows = []
    path = os.path.join(os.path.dirname(HERE), "wsclean_amd", "csrc", "es_table.inc")
    for line in open(path):
        line = line.strip()
        if line.startswith("{"):
            W, s, b, e = [float(v) for v in line.strip("{},").split(",")][:4]
            rows.append((int(W), s, b, e))
    # The error oscillates in t (the alias terms Phi(t + k) have zeros), so
    # the brute force samples the same 48-point grid the table used.
    for W, s, b, e in [r for r in rows if (r[0], r[1]) in ((8, 2.0), (10, 1.25))]:
        ts = np.linspace(0, 0.5 / s, 48)[24:]
        errs = []
        for t in ts:
            p0 = oracle.es_phi_hat(W, b, t)
            al = [oracle.es_phi_hat(W, b, t + k) for k in range(-8, 9) if k]
            errs.append(np.sqrt(np.sum(np.square(al))) / p0)
        assert abs(max(errs) - e) < 0.05 * e


# ---- the CPU w-gridder restatement (timed CPU baseline) is pinned to the DFT

@pytest.mark.parametrize("W,sigma,beta,eps", [(8, 2.0, 17.68, 1e-5), (6, 2.0, 13.71, 1e-4),
                                              (10, 1.5, 20.65, 1e-5)])
def test_cpu_wgridder_invert_matches_dft(oracle, W, sigma, beta, eps):
    p = synth.small_problem(seed=41, n_ant=14, n_times=30, n_chan=2, npix=128,
                            pixel_arcsec=40.0, radius_m=3000.0, dec_deg=-70.0, dt_s=600.0,
                            noise=0.05)
    n = 128
    prm = oracle.cpu_wgrid_params(n, n, W, sigma, beta)
    # shift of ~20% of the field keeps the sky inside the image (sources
    # outside any FFT-gridded image alias at the kernel error times their flux)
    img = oracle.cpu_wgrid_invert(n, n, p["dx"], p["dx"], 0.003, -0.002, p["uvw"], p["freq"],
                                  p["vis"], prm["W"], prm["beta"], prm["sigma"], prm["nu"],
                                  prm["nv"])
    ref = oracle.dft_invert(n, n, p["dx"], p["dx"], 0.003, -0.002, p["uvw"], p["freq"], p["vis"])
    err = np.linalg.norm(img - ref) / np.linalg.norm(ref)
    assert err < eps, err


def test_cpu_wgridder_predict_matches_dft(oracle):
    p = synth.small_problem(seed=42, n_ant=14, n_times=30, n_chan=2, npix=128,
                            pixel_arcsec=40.0, radius_m=3000.0, dec_deg=-70.0, dt_s=600.0)
    n = 128
    rng = np.random.default_rng(0)
    model = np.zeros((n, n))
    model[rng.integers(0, n, 20), rng.integers(0, n, 20)] = rng.uniform(-1, 1, 20)
    prm = oracle.cpu_wgrid_params(n, n, 8, 2.0, 17.68)
    V = oracle.cpu_wgrid_predict(n, n, p["dx"], p["dx"], 0.0, 0.0, model, p["uvw"], p["freq"],
                                 **prm)
    ref = oracle.dft_predict(n, n, p["dx"], p["dx"], 0.0, 0.0, model, p["uvw"], p["freq"])
    assert np.linalg.norm(V - ref) / np.linalg.norm(ref) < 1e-5


# ------------------------------------------------- gain corrections (f3)

def _rand_jones(rng, n, diag=False):
    g = (rng.normal(size=(n, 2, 2)) + 1j * rng.normal(size=(n, 2, 2))) * 0.3
    g += np.eye(2)
    if diag:
        g[:, 0, 1] = g[:, 1, 0] = 0
    return g


def test_gain_identity_leaves_visibilities(oracle):
    rng = np.random.default_rng(3)
    for mode, npol in oracle.GAIN_NPOL.items():
        v = rng.normal(size=npol) + 1j * rng.normal(size=npol)
        for conj in (False, True):
            np.testing.assert_allclose(oracle._apply_gain(mode, v, np.eye(2), np.eye(2), conj), v)


def test_gain_modes_against_matrix_algebra(oracle):
    # xx / yy / trace and the 2-vis diagonal mode are the entries of the full
    # 2x2 form G1^H V G2 (conjugate) and G1 V G2^H (forward) that they select
    rng = np.random.default_rng(4)
    for _ in range(20):
        a, b = _rand_jones(rng, 2)
        ad, bd = _rand_jones(rng, 2, diag=True)
        V = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        full_c = a.conj().T @ V @ b
        full_f = a @ V @ b.conj().T
        np.testing.assert_allclose(oracle._apply_gain(oracle.GAIN_FULL, V.ravel(), a, b, True),
                                   full_c.ravel())
        np.testing.assert_allclose(oracle._apply_gain(oracle.GAIN_FULL, V.ravel(), a, b, False),
                                   full_f.ravel())
        s = V[0, 0]
        np.testing.assert_allclose(oracle._apply_gain(oracle.GAIN_XX, [s], a, b, True)[0],
                                   np.conj(a[0, 0]) * s * b[0, 0])
        np.testing.assert_allclose(oracle._apply_gain(oracle.GAIN_YY, [s], a, b, False)[0],
                                   a[1, 1] * s * np.conj(b[1, 1]))
        # trace: 0.5 tr(G1^H G2) (conjugate), 0.5 tr(G1 G2^H) (forward)
        np.testing.assert_allclose(oracle._apply_gain(oracle.GAIN_TRACE, [s], a, b, True)[0],
                                   s * 0.5 * np.trace(a.conj().T @ b))
        np.testing.assert_allclose(oracle._apply_gain(oracle.GAIN_TRACE, [s], a, b, False)[0],
                                   s * 0.5 * np.trace(a @ b.conj().T))
        # diagonal gains: the 2-vis mode is the diagonal of the full form
        d = oracle._apply_gain(oracle.GAIN_2VIS_DIAGONAL, [V[0, 0], V[1, 1]], ad, bd, True)
        full_d = ad.conj().T @ V @ bd
        np.testing.assert_allclose(d, [full_d[0, 0], full_d[1, 1]])


def test_correct_and_collapse_unitary_round_trip(oracle):
    # forward then conjugate gain with unitary Jones matrices is the identity
    rng = np.random.default_rng(5)
    n_rows, n_chan, n_ant, n_times = 6, 3, 4, 2
    q, _ = np.linalg.qr(rng.normal(size=(n_times * n_chan * n_ant, 2, 2))
                        + 1j * rng.normal(size=(n_times * n_chan * n_ant, 2, 2)))
    gains = q.reshape(n_times, n_chan, n_ant, 4)
    vis = rng.normal(size=(n_rows, n_chan, 4)) + 1j * rng.normal(size=(n_rows, n_chan, 4))
    ant = rng.integers(0, n_ant, (n_rows, 2))
    toff = rng.integers(0, n_times, n_rows)
    out = oracle.correct_and_collapse(vis, ant, toff, gains, n_ant, 4, oracle.GAIN_FULL, True)
    np.testing.assert_allclose(out, 0.5 * (vis[..., 0] + vis[..., 3]), atol=1e-12)


@pytest.mark.parametrize("W,beta", [(4, 8.8), (6, 12.48), (8, 16.0), (8, 17.68), (10, 20.65),
                                    (12, 25.08), (14, 29.54), (16, 34.32)])
def test_es_kernel_polynomial_against_closed_form(W, beta):
    # the gridding kernels' per-tap polynomial of the ES kernel, evaluated as
    # the device does (FP32 Horner in g = 2 f - 1), against phi(x) =
    # exp(beta (sqrt(1 - x^2) - 1)) in FP64 on 4e5 fractional coordinates:
    # phi-weighted RMS relative error below 1e-7 (W >= 8; 2e-7 at W 6, where
    # the support edge's e^-beta = 4e-6 limits the fit) and every tap's
    # largest error below e^-beta plus 3e-7 of the peak; taps >= W are zero
    import ctypes
    from wsclean_amd import _lib
    lib = _lib.load()
    coef = np.zeros((16, 12), np.float32)
    deg = ctypes.c_int()
    _lib.check(lib.wgx_es_poly(W, beta, coef.ctypes.data_as(ctypes.c_void_p), ctypes.byref(deg)))
    D = deg.value
    assert np.all(coef[W:] == 0) and np.all(coef[:, D + 1:] == 0)
    f = np.random.default_rng(W).random(400000).astype(np.float32)
    g = (np.float32(2) * f - np.float32(1)).astype(np.float32)
    num = den = 0.0
    for k in range(W):
        acc = np.full_like(g, coef[k, D])
        for d in range(D - 1, -1, -1):
            acc = (acc * g + coef[k, d]).astype(np.float32)
        x = (k - W / 2 + 1 - f.astype(np.float64)) * 2.0 / W
        ref = np.exp(beta * (np.sqrt(np.maximum(0.0, 1.0 - x * x)) - 1.0))
        err = acc.astype(np.float64) - ref
        assert np.abs(err).max() < np.exp(-beta) + 3e-7, (k, np.abs(err).max())
        num += (err ** 2).sum()
        den += (ref ** 2).sum()
    rms = np.sqrt(num / den)
    assert rms < (1e-7 if W >= 8 else 2e-7 if W == 6 else 2e-5), rms


def test_polarization_conversion_closed_forms(oracle):
    # The oracle's conversion from linear correlations (coherency-matrix
    # traces with the Pauli matrices) against the closed forms of the IAU
    # definitions XX = I + Q, YY = I - Q, XY = U + iV, YX = U - iV, the
    # circular identities RR + LL = 2 I, RR - LL = 2 V, RL + LR = 2 Q,
    # RL - LR = 2 i U, and the physics of a feed rotation (I and V invariant,
    # Q + iU turns by 2 theta); a swapped XY / YX or a wrong V sign fails
    rng = np.random.default_rng(4)
    I, Q, U, V = (rng.normal(size=50) + 1j * rng.normal(size=50) for _ in range(4))
    lin = np.stack([I + Q, U + 1j * V, U - 1j * V, I - Q], axis=-1)
    got = [oracle.convert_from_linear(lin, p) for p in (1, 2, 3, 4)]
    for a, b in zip(got, (I, Q, U, V)):
        np.testing.assert_allclose(a, b, atol=1e-12)
    for p, k in ((-5, 0), (-7, 1), (-8, 2), (-6, 3)):
        np.testing.assert_array_equal(oracle.convert_from_linear(lin, p), lin[:, k])
    rr, ll, rl, lr = (oracle.convert_from_linear(lin, p) for p in (-1, -2, -3, -4))
    np.testing.assert_allclose(rr + ll, 2 * I, atol=1e-12)
    np.testing.assert_allclose(rr - ll, 2 * V, atol=1e-12)
    np.testing.assert_allclose(rl + lr, 2 * Q, atol=1e-12)
    np.testing.assert_allclose(rl - lr, 2j * U, atol=1e-12)
    # a real sky (I, Q, U, V real) seen through feeds rotated by theta:
    # C' = R C R^T with R the rotation of the (X, Y) feed pair
    Ir, Qr, Ur, Vr = 1.0, 0.3, -0.2, 0.1
    C = np.array([[Ir + Qr, Ur + 1j * Vr], [Ur - 1j * Vr, Ir - Qr]])
    th = 0.37
    R = np.array([[np.cos(th), np.sin(th)], [-np.sin(th), np.cos(th)]])
    Cr = (R @ C @ R.T).reshape(4)
    I2, Q2, U2, V2 = oracle.stokes_from_linear(Cr)
    assert abs(I2 - Ir) < 1e-12 and abs(V2 - Vr) < 1e-12
    np.testing.assert_allclose(Q2 + 1j * U2, (Qr + 1j * Ur) * np.exp(-2j * th), atol=1e-12)
    # a physical coherency (C = J J^H) is positive semi-definite: I >= |(Q, U, V)|
    for _ in range(20):
        J = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        s = [np.real(x) for x in oracle.stokes_from_linear((J @ J.conj().T).reshape(4))]
        assert s[0] >= np.sqrt(s[1] ** 2 + s[2] ** 2 + s[3] ** 2) - 1e-12
