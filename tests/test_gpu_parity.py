"""GPU parity tests: the HIP w-gridder through the C-ABI against the FP64
direct-FT oracle (DirectMSGridder restatement) and against exact
properties.  Tolerances are the requested epsilon in relative L2 (the
meaning of epsilon in doc/source/wgridding.rst:19-20) unless stated.
"""
import functools

import numpy as np
import pytest

from wsclean_amd import WGriddingGridder_Simple, WgError, fft_lines, synth

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def make(n, dx, eps, ls=0.0, ms=0.0, width=None):
    w = width or n
    return WGriddingGridder_Simple(w, w, n, n, dx, dx, ls, ms, 1, eps)


def invert(p, eps, ls=0.0, ms=0.0, n=None):
    n = n or p["npix"]
    g = make(n, p["dx"], eps, ls, ms)
    g.InitializeInversion()
    g.AddInversionData(p["uvw"].shape[0], len(p["freq"]), p["uvw"], p["freq"], p["vis"])
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(n, n)
    return img, g


# ------------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", [16, 18, 60, 81, 96, 162, 486, 1000, 1458, 2048, 2560, 3000, 4096,
                               4374, 10240, 13122, 13500, 13824, 14580, 16384,
                               # two-pass lines: R = 3, 2, 2, 2, 3, 5, 4, 5 subsequences
                               19683, 20000, 24576, 32768, 39366, 50625, 65536, 81920])
def test_lds_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = (rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))).astype(np.complex64)
    f = fft_lines(x)
    ref = np.fft.fft(x.astype(np.complex128), axis=1)
    assert rel_l2(f, ref) < 2e-6
    b = fft_lines(x, inverse=True)
    refb = np.fft.ifft(x.astype(np.complex128), axis=1) * n
    assert rel_l2(b, refb) < 2e-6


# ------------------------------------------------------------------ sort

@pytest.mark.parametrize("eps", [1e-3, 1e-5])
def test_bucket_sort_bit_exact(oracle, eps):
    p = synth.small_problem(seed=7, n_ant=16, n_times=30, n_chan=3, npix=128)
    vis = p["vis"].copy()
    vis[::7, 1] = 0  # zero visibilities are skipped
    g = make(128, p["dx"], eps)
    keys, perm, plan = g.sort_keys(p["uvw"], p["freq"], vis)
    ref_keys, valid = oracle.bucket_keys(plan, p["uvw"], p["freq"], vis)
    np.testing.assert_array_equal(keys, ref_keys)
    ref_perm = oracle.bucket_perm(plan, p["uvw"], p["freq"], vis)
    assert ref_perm.size == int(valid.sum())
    np.testing.assert_array_equal(perm, ref_perm)


@pytest.mark.parametrize("n_ant,n_times,n_chan", [(15, 31, 3), (21, 67, 5)])
def test_bucket_sort_ragged_counts(oracle, n_ant, n_times, n_chan):
    # element counts that are odd (not a multiple of the 4-key loads of the
    # digit histogram) and not a multiple of a radix tile
    p = synth.small_problem(seed=23, n_ant=n_ant, n_times=n_times, n_chan=n_chan, npix=256,
                            pixel_arcsec=30.0, radius_m=4000.0, dec_deg=-50.0, dt_s=300.0)
    n = p["uvw"].shape[0] * n_chan
    assert n % 4 != 0 and n % 8192 != 0
    g = make(256, p["dx"], 1e-4)
    keys, perm, plan = g.sort_keys(p["uvw"], p["freq"], None)
    ref_keys, _ = oracle.bucket_keys(plan, p["uvw"], p["freq"], None)
    np.testing.assert_array_equal(keys, ref_keys)
    np.testing.assert_array_equal(perm, oracle.bucket_perm(plan, p["uvw"], p["freq"], None))


def test_bucket_sort_bit_exact_large(oracle):
    # enough visibilities for several radix blocks and all three digit passes
    p = synth.small_problem(seed=17, n_ant=40, n_times=60, n_chan=8, npix=512,
                            pixel_arcsec=30.0, radius_m=6000.0, dec_deg=-70.0, dt_s=600.0)
    g = make(512, p["dx"], 1e-5)
    keys, perm, plan = g.sort_keys(p["uvw"], p["freq"], None)
    ref_keys, _ = oracle.bucket_keys(plan, p["uvw"], p["freq"], None)
    assert plan["key_bits"] > 8
    np.testing.assert_array_equal(keys, ref_keys)
    np.testing.assert_array_equal(perm, oracle.bucket_perm(plan, p["uvw"], p["freq"], None))


def _sort_patterns(rng, n, key_bits):
    top = 1 << key_bits
    runs = np.repeat(rng.integers(0, top, max(1, n // 7 + 1), dtype=np.uint64),
                     7)[:n]
    yield "random", rng.integers(0, top, n, dtype=np.uint64)
    yield "equal", np.full(n, top - 1, dtype=np.uint64)
    yield "runs", runs
    yield "sorted", np.sort(rng.integers(0, top, n, dtype=np.uint64))
    # equal digits in separate runs of one wave
    yield "alternating", (np.arange(n, dtype=np.uint64) % 2) * (top - 1)
    yield "few", rng.integers(0, 3, n, dtype=np.uint64) << np.uint64(max(0, key_bits - 2))


@pytest.mark.parametrize("n,key_bits", [(1, 9), (2, 5), (100, 9), (8192, 18), (8193, 27),
                                        (65536 * 3 + 17, 27), (300001, 21)])
def test_radix_sort_patterns(n, key_bits):
    # the device sort against numpy's stable argsort, on key patterns that
    # take the run-based ranking (runs, sorted, equal) and the per-bit
    # ballot ranking (random, alternating) of the scatter
    from wsclean_amd.gridder import radix_sort
    rng = np.random.default_rng(n + key_bits)
    for name, k in _sort_patterns(rng, n, key_bits):
        k = k.astype(np.uint32)
        ko, po = radix_sort(k, key_bits)
        ref = np.argsort(k, kind="stable")
        np.testing.assert_array_equal(po, ref, err_msg=name)
        np.testing.assert_array_equal(ko, k[ref], err_msg=name)


# ---------------------------------------------------------------- invert

@pytest.mark.parametrize("eps", [1e-3, 1e-4, 1e-5])
def test_invert_matches_direct_ft(oracle, eps):
    p = synth.small_problem(seed=1001, n_ant=16, n_times=24, n_chan=2, npix=96, noise=0.01)
    img, g = invert(p, eps)
    ref = oracle.dft_invert(96, 96, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"], p["vis"])
    err = rel_l2(img, ref)
    print("eps", eps, "rel_l2", err, g.plan())
    assert err < eps


def test_invert_with_phase_centre_shift(oracle):
    # a facet-like image far from the phase centre, with its sky inside it
    ls, ms = 0.05, -0.03
    p = synth.small_problem(seed=3, n_ant=14, n_times=20, n_chan=2, npix=80, sky_lm=(ls, ms))
    img, g = invert(p, 1e-5, ls, ms)
    ref = oracle.dft_invert(80, 80, p["dx"], p["dx"], ls, ms, p["uvw"], p["freq"], p["vis"])
    assert rel_l2(img, ref) < 1e-5


def test_invert_wide_field_many_w_planes(oracle):
    # 4 degree field, 12 km baselines, 6.5 h track at dec -70: P >> W
    p = synth.small_problem(seed=5, n_ant=20, n_times=40, n_chan=2, npix=512,
                            pixel_arcsec=30.0, radius_m=6000.0, n_src=8, dec_deg=-70.0,
                            dt_s=600.0)
    img, g = invert(p, 1e-4)
    plan = g.plan()
    assert plan["P"] > plan["W"] + 4, plan
    rng = np.random.default_rng(0)
    xs = rng.integers(0, 512, 600)
    ys = rng.integers(0, 512, 600)
    ref = oracle.dft_invert_pixels(512, 512, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"],
                                   p["vis"], xs, ys)
    assert rel_l2(img[ys, xs], ref) < 1e-4


def dense_core_problem():
    # compact 300 m array, 40 antennas, 240 x 60 s, 16 channels: 3.0e6
    # visibilities in a few hundred uv cells, so work items are full (65536
    # visibilities) and runs of one cell are thousands long
    return synth.small_problem(seed=23, n_ant=40, n_times=240, n_chan=16, npix=256,
                               pixel_arcsec=20.0, radius_m=300.0, n_src=12, dt_s=60.0)


def test_invert_dense_core(oracle):
    # the fixed-point subgrid of the matrix-core gridder at its largest
    # per-item range; tolerance epsilon = 1e-5 on 400 sampled pixels
    p = dense_core_problem()
    img, g = invert(p, 1e-5)
    assert g.plan()["W"] <= 8
    rng = np.random.default_rng(1)
    xs = rng.integers(0, 256, 400)
    ys = rng.integers(0, 256, 400)
    ref = oracle.dft_invert_pixels(256, 256, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"],
                                   p["vis"], xs, ys)
    assert rel_l2(img[ys, xs], ref) < 1e-5


@functools.lru_cache(maxsize=1)
def uniform_weighted_problem():
    # 48 antennas in a log-radial 3 km layout (dense core, sparse outskirts),
    # 240 x 60 s, 8 channels: 2.2e6 visibilities times uniform density
    # weights (1 / visibilities in the image's uv cell), which span about
    # four decades; plus 20 long-baseline visibilities x 1e3 (unflagged
    # interference).  The gridders' fixed-point scale must follow each work
    # item's own visibilities for the faint, densely sampled core to keep
    # epsilon.
    rng = np.random.Generator(np.random.PCG64(31))
    enu = synth.antennas_logradial(rng, 48, 3000.0)
    uvw = synth.earth_rotation_uvw(enu, 240, 60.0, np.deg2rad(-30.0))
    freq = 140e6 + 4e6 * np.arange(8)
    npix, dx = 256, 20.0 * synth.ARCSEC
    l, m, flux, _, _ = synth.point_sky(rng, 12, npix, dx)
    vis = synth.predict_sources(uvw, freq, l, m, flux).astype(np.complex64)
    s = freq[None, :] / 299792458.0 * npix * dx
    u, v = uvw[:, 0:1] * s, uvw[:, 1:2] * s
    flip = v < 0
    u, v = np.where(flip, -u, u), np.where(flip, -v, v)
    cell = (np.floor(u).astype(np.int64) % npix) * npix + np.floor(v).astype(np.int64) % npix
    count = np.bincount(cell.ravel(), minlength=npix * npix)
    wt = 1.0 / count[cell]
    vis = (vis * wt).astype(np.complex64)
    r = np.hypot(uvw[:, 0], uvw[:, 1])
    far = np.argsort(r)[-20:]
    vis[far, 0] *= 1e3
    return dict(uvw=uvw, freq=freq, vis=vis, npix=npix, dx=dx, wt_range=wt.max() / wt.min())


@pytest.mark.parametrize("precision,sub_slab", [(-1, 0), (0, 0), (-1, 2)])
def test_invert_uniform_weighted_dynamic_range(oracle, precision, sub_slab):
    # default split kernels with the fused gather; FP32 matrix kernels; and
    # several plane sub-slabs (pre-gathered visibilities)
    p = uniform_weighted_problem()
    assert p["wt_range"] > 1e3
    n = p["npix"]
    g = make(n, p["dx"], 1e-5)
    g.set_precision(precision)
    if sub_slab:
        g.set_tuning(sub_slab=sub_slab)
    g.InitializeInversion()
    g.AddInversionData(p["uvw"].shape[0], len(p["freq"]), p["uvw"], p["freq"], p["vis"])
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(n, n)
    plan = g.plan()
    assert plan["W"] <= 8
    if sub_slab:
        assert plan["P"] > sub_slab
    rng = np.random.default_rng(2)
    xs = rng.integers(0, n, 300)
    ys = rng.integers(0, n, 300)
    ref = oracle.dft_invert_pixels(n, n, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"],
                                   p["vis"], xs, ys)
    assert rel_l2(img[ys, xs], ref) < 1e-5


@pytest.mark.parametrize("bad", [float("nan"), float("inf")])
@pytest.mark.parametrize("precision,sub_slab,one_call",
                         [(-1, 0, True), (-1, 0, False), (0, 0, False), (-1, 2, False)])
def test_non_finite_visibility_reaches_the_image(bad, precision, sub_slab, one_call):
    # one NaN or infinite visibility makes the dirty image non-finite, as an
    # FP gridder's FFT spreads it, instead of being dropped with its work
    # item: its row maximum and item bounds carry it, and the item's
    # subgrid is written as NaN.  Paths: the one-call row maxima from the
    # keys pass, the prepared row maxima, the FP32 matrix kernel, and the
    # pre-gathered components of several sub-slabs.
    import torch
    p = uniform_weighted_problem()
    vis = p["vis"].copy()
    vis[len(vis) // 3, 5] = complex(bad, 0.0)
    n = p["npix"]
    g = make(n, p["dx"], 1e-5)
    g.set_precision(precision)
    if sub_slab:
        g.set_tuning(sub_slab=sub_slab)
    n_rows, n_chan = vis.shape
    g.InitializeInversion()
    if one_call:
        g.AddInversionData(n_rows, n_chan, p["uvw"], p["freq"], vis)
    else:
        dev = torch.device("cuda:0")
        g.PrepareVisibilities(torch.from_numpy(p["uvw"]).to(dev), p["freq"], n_rows, n_chan)
        g.AddInversionPrepared(
            torch.from_numpy(vis.view(np.float32).reshape(n_rows, n_chan, 2).copy()).to(dev))
    g.FinalizeImage(1.0)
    img = g.RealImage()
    if sub_slab:
        assert g.plan()["P"] > sub_slab
    assert not np.isfinite(img).any()


def test_invert_is_deterministic():
    # integer subgrid accumulation + fixed-order gather: bitwise repeatable
    p = dense_core_problem()
    a, _ = invert(p, 1e-5)
    b, _ = invert(p, 1e-5)
    np.testing.assert_array_equal(a, b)


def test_invert_wrapping_uv(oracle):
    # pixels too coarse for the longest baselines: the uv box wraps around
    p = synth.small_problem(seed=9, n_ant=12, n_times=10, n_chan=2, npix=64,
                            pixel_arcsec=60.0, radius_m=3000.0)
    img, g = invert(p, 1e-4)
    assert g.plan()["wrap_u"] == 1 or g.plan()["wrap_v"] == 1
    ref = oracle.dft_invert(64, 64, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"], p["vis"])
    assert rel_l2(img, ref) < 1e-4


def test_point_source_peak(oracle):
    p = synth.small_problem(seed=12, n_src=1, on_pixel=True, npix=64, n_chan=2)
    img, g = invert(p, 1e-5)
    x0, y0 = int(p["xs"][0]), int(p["ys"][0])
    peak = img[y0, x0] / p["vis"].size
    assert abs(peak - p["flux"][0]) < 1e-5 * p["flux"][0] * 10


def test_chunk_additivity_and_frequency_reversal():
    p = synth.small_problem(seed=2, n_ant=12, n_times=20, n_chan=3, npix=64)
    one, _ = invert(p, 1e-4)
    g = make(64, p["dx"], 1e-4)
    g.InitializeInversion()
    h = p["uvw"].shape[0] // 2
    g.AddInversionData(h, 3, p["uvw"][:h], p["freq"], p["vis"][:h])
    g.AddInversionData(p["uvw"].shape[0] - h, 3, p["uvw"][h:], p["freq"], p["vis"][h:])
    g.FinalizeImage(1.0)
    two = g.RealImage().reshape(64, 64)
    assert rel_l2(two, one) < 1e-4
    q = dict(p)
    q["freq"] = p["freq"][::-1].copy()
    q["vis"] = np.ascontiguousarray(p["vis"][:, ::-1])
    rev, _ = invert(q, 1e-4)
    assert rel_l2(rev, one) < 1e-5


def test_zero_visibilities_give_zero_image():
    p = synth.small_problem(seed=2, npix=32)
    p["vis"][:] = 0
    img, _ = invert(p, 1e-4)
    assert np.all(img == 0)


def test_real_image_trim_and_nan_border():
    p = synth.small_problem(seed=4, npix=48)
    g = WGriddingGridder_Simple(64, 60, 48, 48, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-4)
    g.InitializeInversion()
    g.AddInversionData(p["uvw"].shape[0], 2, p["uvw"], p["freq"], p["vis"])
    g.FinalizeImage(0.5)
    img = g.RealImage().reshape(60, 64)
    inner = img[6:54, 8:56]
    assert np.all(np.isfinite(inner))
    assert np.isnan(img[:6]).all() and np.isnan(img[54:]).all()
    assert np.isnan(img[:, :8]).all() and np.isnan(img[:, 56:]).all()
    ref, _ = invert(p, 1e-4, n=48)
    assert rel_l2(inner, 0.5 * ref) < 1e-5


# --------------------------------------------------------------- predict

def predict(n, dx, eps, model, uvw, freq, ls=0.0, ms=0.0, width=None):
    w = width or n
    g = WGriddingGridder_Simple(w, w, n, n, dx, dx, ls, ms, 1, eps)
    g.InitializePrediction(model.astype(np.float32).ravel())
    out = np.zeros((uvw.shape[0], len(freq)), np.complex64)
    g.PredictVisibilities(uvw.shape[0], len(freq), uvw, freq, out)
    return out, g


@pytest.mark.parametrize("eps", [1e-3, 1e-4, 1e-5])
def test_predict_matches_direct_ft(oracle, eps):
    p = synth.small_problem(seed=31, n_ant=16, n_times=24, n_chan=2, npix=96)
    rng = np.random.default_rng(1)
    model = np.zeros((96, 96))
    ys = rng.integers(5, 90, 12)
    xs = rng.integers(5, 90, 12)
    model[ys, xs] = rng.uniform(0.1, 1.0, 12)
    out, g = predict(96, p["dx"], eps, model, p["uvw"], p["freq"])
    ref = oracle.dft_predict(96, 96, p["dx"], p["dx"], 0.0, 0.0, model, p["uvw"], p["freq"])
    err = rel_l2(out, ref)
    print("eps", eps, "rel_l2", err, g.plan())
    assert err < eps


def test_predict_shift_and_wide_field(oracle):
    p = synth.small_problem(seed=8, n_ant=20, n_times=30, n_chan=2, npix=384,
                            pixel_arcsec=30.0, radius_m=5000.0)
    rng = np.random.default_rng(2)
    model = np.zeros((384, 384))
    model[rng.integers(0, 384, 30), rng.integers(0, 384, 30)] = rng.uniform(-1, 1, 30)
    ls, ms = -0.01, 0.02  # the model lives on the shifted image
    out, g = predict(384, p["dx"], 1e-4, model, p["uvw"], p["freq"], ls, ms)
    sel = rng.choice(out.size, 500, replace=False)
    ref = oracle.dft_predict(384, 384, p["dx"], p["dx"], ls, ms, model, p["uvw"], p["freq"],
                             sel=sel)
    assert rel_l2(out.ravel()[sel], ref) < 1e-4


def test_predict_zero_image_and_untrim():
    p = synth.small_problem(seed=4, npix=48)
    out, _ = predict(48, p["dx"], 1e-4, np.zeros((48, 48)), p["uvw"], p["freq"])
    assert np.all(out == 0)


def test_adjoint_invert_predict():
    # Re<A x, y> = <x, A^H y> holds to epsilon for the gridder pair
    p = synth.small_problem(seed=13, n_ant=14, n_times=20, n_chan=2, npix=64, noise=0.1)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(64, 64))
    Ax, _ = predict(64, p["dx"], 1e-5, x, p["uvw"], p["freq"])
    AHy, _ = invert(p, 1e-5)
    lhs = float(np.sum(x * AHy))
    rhs = float(np.real(np.sum(p["vis"].astype(np.complex128) * np.conj(Ax))))
    assert abs(lhs - rhs) < 1e-5 * np.linalg.norm(x) * np.linalg.norm(AHy)


def test_major_cycle_prepared_matches_stateless():
    import torch
    p = synth.small_problem(seed=21, n_ant=14, n_times=20, n_chan=2, npix=64)
    ref_img, _ = invert(p, 1e-4)
    g = make(64, p["dx"], 1e-4)
    dev = torch.device("cuda:0")
    uvw = torch.from_numpy(p["uvw"]).to(dev)
    vis = torch.from_numpy(p["vis"].view(np.float32)).to(dev)
    g.PrepareVisibilities(uvw, p["freq"], p["uvw"].shape[0], 2)
    g.InitializeInversion()
    g.AddInversionPrepared(vis)
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(64, 64)
    assert rel_l2(img, ref_img) < 1e-5
    out = torch.zeros_like(vis)
    torch.cuda.synchronize()  # the fill runs on torch's stream, the library on its own
    g.InitializePrediction(img.ravel())
    g.PredictPrepared(out)
    ref_out, _ = predict(64, p["dx"], 1e-4, img, p["uvw"], p["freq"])
    got = out.cpu().numpy().view(np.complex64).reshape(ref_out.shape)
    assert rel_l2(got, ref_out) < 1e-5


@pytest.mark.parametrize("n_chan", [3, 16, 64])
def test_one_call_inversion_equals_prepared_bitwise(n_chan):
    """The one-call inversion takes the row maxima of |V| from its keys
    kernel, the prepared path from its own pass over the visibilities (pairs
    of channels per load for an even channel count, in caller row order,
    then gathered into renumbered order): both give the same bits, so the
    work items' gridding scales and the images are identical.  3, 16 and 64
    channels put several rows, rows split between waves, and one row in a
    wave's elements."""
    import torch
    p = synth.small_problem(seed=23, n_ant=16, n_times=24, n_chan=n_chan, npix=96,
                            df=40e6 / n_chan)
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    assert np.all(p["vis"] != 0)   # no zero skipping: the same sorted set
    dev = torch.device("cuda:0")
    uvw = torch.from_numpy(p["uvw"]).to(dev)
    vis = torch.from_numpy(
        np.ascontiguousarray(p["vis"]).view(np.float32).reshape(n_rows, n_chan, 2)).to(dev)
    torch.cuda.synchronize()
    g1 = make(96, p["dx"], 1e-5)
    g1.InitializeInversion()
    g1.AddInversionDataDevice(uvw, p["freq"], vis)
    g1.FinalizeImage(1.0)
    g2 = make(96, p["dx"], 1e-5)
    g2.PrepareVisibilities(uvw, p["freq"], n_rows, n_chan)
    g2.InitializeInversion()
    g2.AddInversionPrepared(vis)
    g2.FinalizeImage(1.0)
    assert np.array_equal(g1.RealImage(), g2.RealImage())
    g1.close()
    g2.close()


# ---------------------------------------------------------------- errors

def test_errors_leave_handle_usable():
    with pytest.raises(WgError):
        WGriddingGridder_Simple(64, 64, 64, 64, 1e-4, 1e-4, 0, 0, 1, 1e-4, verbosity=3)
    p = synth.small_problem(seed=4, npix=32)
    g = make(32, p["dx"], 1e-4)
    g.InitializeInversion()
    with pytest.raises(WgError):
        g.AddInversionData(p["uvw"].shape[0], 2, p["uvw"], -p["freq"], p["vis"])
    g.AddInversionData(p["uvw"].shape[0], 2, p["uvw"], p["freq"], p["vis"])
    g.FinalizeImage(1.0)
    assert np.isfinite(g.RealImage()).all()


@pytest.mark.parametrize("axis", [0, 2])
def test_non_finite_uvw_is_refused(axis):
    # a NaN coordinate reaches the extents (their min / max keep NaN) and
    # the call fails with the non-finite-coordinates error instead of
    # gridding visibilities located from NaN; the handle stays usable
    import torch
    p = synth.small_problem(seed=5, npix=32)
    n_rows = p["uvw"].shape[0]
    bad = p["uvw"].copy()
    bad[n_rows // 2, axis] = np.nan
    g = make(32, p["dx"], 1e-4)
    g.InitializeInversion()
    with pytest.raises(WgError, match="non-finite uvw"):
        g.AddInversionData(n_rows, 2, bad, p["freq"], p["vis"])
    with pytest.raises(WgError, match="non-finite uvw"):
        g.PrepareVisibilities(torch.from_numpy(bad).to("cuda:0"), p["freq"], n_rows, 2)
    out = np.zeros((n_rows, 2), np.complex64)
    g.InitializePrediction(np.zeros((32, 32), np.float32))
    with pytest.raises(WgError, match="non-finite uvw"):
        g.PredictVisibilities(n_rows, 2, bad, p["freq"], out)
    g.InitializeInversion()
    g.AddInversionData(n_rows, 2, p["uvw"], p["freq"], p["vis"])
    g.FinalizeImage(1.0)
    assert np.isfinite(g.RealImage()).all()


# ------------------------------------------------ correction callback (f3)

@pytest.mark.parametrize("mode,nparms,forward,pol", [
    (2, 2, False, 1), (3, 2, False, 1), (4, 4, False, 1), (4, 4, True, 1), (0, 4, False, 1),
    # the 4-polarization collapse to the other polarizations (CollapseData<4>,
    # gridding/msgridderdata.h:45-49; wgridder/callback_implementation.h:92-95)
    (4, 4, False, 2), (4, 4, False, 3), (4, 4, True, 4), (4, 4, False, -1), (4, 4, False, -3),
    (4, 4, False, -4), (4, 4, False, -7), (4, 4, False, -6)])
def test_corrected_inversion_matches_host_correction(oracle, mode, nparms, forward, pol):
    # GPU gain correction + collapse + gridding against the oracle's
    # correction on the host followed by the plain inversion; the two paths
    # differ only by FP32 rounding of the correction (tolerance 1e-5)
    p = synth.small_problem(seed=31, n_ant=12, n_times=16, n_chan=3, npix=96)
    rng = np.random.default_rng(mode * 10 + nparms)
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    npol = oracle.GAIN_NPOL[mode]
    n_ant, n_times = 12, 4
    vis = np.repeat(p["vis"][:, :, None], npol, axis=2)
    vis = vis + 0.1 * (rng.normal(size=vis.shape) + 1j * rng.normal(size=vis.shape))
    vis = vis.astype(np.complex64)
    g = (1.0 + 0.2 * (rng.normal(size=(n_times, n_chan, n_ant, nparms))
                      + 1j * rng.normal(size=(n_times, n_chan, n_ant, nparms))))
    if nparms == 4:
        g[..., 1:3] -= 1.0
    g = g.astype(np.complex64)
    ant = rng.integers(0, n_ant, (n_rows, 2)).astype(np.uint32)
    toff = rng.integers(0, n_times, n_rows).astype(np.uint64)
    host = oracle.correct_and_collapse(vis, ant, toff, g, n_ant, nparms, mode, forward, pol)
    ref_g = make(96, p["dx"], 1e-5)
    ref_g.InitializeInversion()
    ref_g.AddInversionData(n_rows, n_chan, p["uvw"], p["freq"], host.astype(np.complex64))
    ref_g.FinalizeImage(1.0)
    ref = ref_g.RealImage()
    gg = make(96, p["dx"], 1e-5)
    gg.InitializeInversion()
    gg.AddInversionDataWithCorrectionCallback(mode, npol, n_rows, p["uvw"], p["freq"], n_chan,
                                              ant, vis, toff, g, apply_forward=forward,
                                              polarization=pol)
    gg.FinalizeImage(1.0)
    assert rel_l2(gg.RealImage(), ref) < 1e-5
