"""GPU parity of the kernel paths that wide fields and tight epsilons
select: the wave-cooperative VALU gridder / degridder (kernel support
W >= 10), the unfused path with several plane sub-slabs (gather kernel,
per-sub-slab box gather and degridding into an accumulator, scatter
kernel), several FFT slabs, and the default plan of a wide field.

Every result is compared with the FP64 direct FT (the DirectMSGridder
restatement); the tolerance is the requested epsilon in relative L2
(doc/source/wgridding.rst:19-20).  Forced shapes are chosen with an error
estimate below the epsilon of the test.
"""
import numpy as np
import pytest

from wsclean_amd import WGriddingGridder_Simple, synth

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def p96():
    return synth.small_problem(seed=1001, n_ant=16, n_times=24, n_chan=2, npix=96, noise=0.01)


def p512():
    # 4 degree field, 12 km baselines, 6.5 h at dec -70: P >> W
    return synth.small_problem(seed=5, n_ant=20, n_times=40, n_chan=2, npix=512,
                               pixel_arcsec=30.0, radius_m=6000.0, n_src=8, dec_deg=-70.0,
                               dt_s=600.0)


def p512b():
    # 40 antennas, 8 channels: P = 21 planes at eps 1e-5
    return synth.small_problem(seed=17, n_ant=40, n_times=60, n_chan=8, npix=512,
                               pixel_arcsec=30.0, radius_m=6000.0, dec_deg=-70.0, dt_s=600.0)


PROBLEMS = {"96": p96, "512": p512, "512b": p512b}


def run_pair(oracle, p, eps, tuning, n_pix_sample=600, n_vis_sample=2000, seed=0):
    """Invert p's visibilities and predict a sparse model with the given
    tuning; returns (invert error, predict error, plan)."""
    n = p["npix"]
    rng = np.random.default_rng(seed)
    g = WGriddingGridder_Simple(n, n, n, n, p["dx"], p["dx"], 0.0, 0.0, 1, eps)
    if tuning:
        g.set_tuning(**tuning)
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    g.InitializeInversion()
    g.AddInversionData(n_rows, n_chan, p["uvw"], p["freq"], p["vis"])
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(n, n)
    plan_inv = g.plan()
    if n <= 128:
        ref = oracle.dft_invert(n, n, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"], p["vis"])
        e_inv = rel_l2(img, ref)
    else:
        xs, ys = rng.integers(0, n, n_pix_sample), rng.integers(0, n, n_pix_sample)
        ref = oracle.dft_invert_pixels(n, n, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"],
                                       p["vis"], xs, ys)
        e_inv = rel_l2(img[ys, xs], ref)
    model = np.zeros((n, n))
    k = 24
    model[rng.integers(2, n - 2, k), rng.integers(2, n - 2, k)] = rng.uniform(-1.0, 1.0, k)
    g.InitializePrediction(model.astype(np.float32).ravel())
    out = np.zeros((n_rows, n_chan), np.complex64)
    g.PredictVisibilities(n_rows, n_chan, p["uvw"], p["freq"], out)
    plan_pred = g.plan()
    sel = rng.choice(out.size, min(out.size, n_vis_sample), replace=False)
    refv = oracle.dft_predict(n, n, p["dx"], p["dx"], 0.0, 0.0, model, p["uvw"], p["freq"],
                              sel=sel)
    e_pred = rel_l2(out.ravel()[sel], refv)
    g.close()
    assert plan_inv == plan_pred
    return e_inv, e_pred, plan_inv


@pytest.mark.parametrize("W,sigma", [(10, 1.5), (12, 1.6), (14, 1.7), (16, 1.8)])
@pytest.mark.parametrize("prob", ["96", "512"])
def test_valu_kernels_match_direct_ft(oracle, W, sigma, prob):
    p = PROBLEMS[prob]()
    print("valu W=%d sigma=%.2f on %s" % (W, sigma, prob), flush=True)
    e_inv, e_pred, plan = run_pair(oracle, p, 1e-5, dict(W=W, sigma=sigma))
    print("  plan W %d sigma %.2f P %d Sg %d err_est %.2e: invert %.2e predict %.2e"
          % (plan["W"], plan["sigma"], plan["P"], plan["Sg"], plan["err_est"], e_inv, e_pred),
          flush=True)
    assert plan["W"] == W and abs(plan["sigma"] - sigma) < 1e-9
    assert plan["err_est"] < 1e-5
    if prob == "512":
        assert plan["P"] > W + 4
    assert e_inv < 1e-5
    assert e_pred < 1e-5


@pytest.mark.parametrize("prob,eps,tuning", [
    ("512", 1e-4, dict(sub_slab=4)),                            # W 8, P 15: 4 sub-slabs
    ("512b", 1e-5, dict(sub_slab=4)),                           # W 8, P 21: 6 sub-slabs
    ("512b", 1e-5, dict(sub_slab=4, slab_bytes=1 << 20)),       # and one FFT slab per sub-slab
    ("512", 1e-5, dict(W=10, sigma=1.5, sub_slab=4)),           # VALU kernels, 5 sub-slabs
    ("512", 1e-5, dict(W=12, sigma=1.6, sub_slab=3, slab_bytes=1 << 20)),
])
def test_sub_slabs_match_direct_ft(oracle, prob, eps, tuning):
    p = PROBLEMS[prob]()
    print("sub-slabs on %s eps %.0e %s" % (prob, eps, tuning), flush=True)
    e_inv, e_pred, plan = run_pair(oracle, p, eps, tuning)
    nsub = -(-plan["P"] // plan["Sg"])
    print("  plan W %d P %d Sg %d Sf %d (%d sub-slabs): invert %.2e predict %.2e"
          % (plan["W"], plan["P"], plan["Sg"], plan["Sf"], nsub, e_inv, e_pred), flush=True)
    assert nsub >= 3
    if "slab_bytes" in tuning:
        assert plan["Sf"] < plan["P"]
    assert e_inv < eps
    assert e_pred < eps


def test_sub_slabs_equal_single_slab():
    # the same shape with one sub-slab and with four: the images differ only
    # by FP32 summation order and fixed-point quanta, far below epsilon
    # (tolerance epsilon / 10)
    p = p512()
    n = p["npix"]
    n_rows = p["uvw"].shape[0]
    imgs = []
    for sub in (0, 4):
        g = WGriddingGridder_Simple(n, n, n, n, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-4)
        g.set_tuning(W=8, sigma=1.35, sub_slab=sub)
        g.InitializeInversion()
        g.AddInversionData(n_rows, 2, p["uvw"], p["freq"], p["vis"])
        g.FinalizeImage(1.0)
        imgs.append(g.RealImage())
        g.close()
    d = rel_l2(imgs[1], imgs[0])
    print("one sub-slab vs four: %.2e" % d, flush=True)
    assert d < 1e-5


def wide_field_default():
    # 128 antennas uniform in a 2.5 km disc, 60 x 120 s at dec -30, one
    # channel at 150 MHz, 4096^2 at 20 arcsec (a 22.8 degree field)
    rng = np.random.Generator(np.random.PCG64(77))
    enu = synth.antennas_disc(rng, 128, 2500.0)
    uvw = synth.earth_rotation_uvw(enu, 60, 120.0, np.deg2rad(-30.0))
    freq = np.array([150e6])
    dx = 20.0 * synth.ARCSEC
    l, m, flux, _, _ = synth.point_sky(rng, 30, 4096, dx)
    vis = synth.add_noise(rng, synth.predict_sources(uvw, freq, l, m, flux), 0.01)
    return dict(uvw=uvw, freq=freq, vis=vis.astype(np.complex64), npix=4096, dx=dx)


def test_default_plan_wide_field(oracle):
    # the default plan of this field is W 10, sigma 1.5, P 41 in three
    # sub-slabs of 20 planes: the VALU kernels on the unfused path
    p = wide_field_default()
    print("wide field 4096^2 at 20 arcsec, %d rows" % p["uvw"].shape[0], flush=True)
    e_inv, e_pred, plan = run_pair(oracle, p, 1e-5, None, n_pix_sample=1000, n_vis_sample=20000)
    print("  plan %s: invert %.2e predict %.2e" % (plan, e_inv, e_pred), flush=True)
    assert plan["W"] >= 10 and plan["P"] > plan["Sg"]
    assert e_inv < 1e-5
    assert e_pred < 1e-5


@pytest.mark.parametrize("eps", [2e-6, 1e-6, 5e-7, 3e-7])
def test_tight_epsilon_fp32_floor(oracle, eps):
    # epsilons down to the FP32 pipeline's floor (about 2.7e-7): the planner
    # picks W 8 / 10 at large oversampling
    p = synth.small_problem(seed=1001, n_ant=16, n_times=24, n_chan=2, npix=96, noise=0.01)
    e_inv, e_pred, plan = run_pair(oracle, p, eps, None)
    print("eps %.0e plan W %d sigma %.2f est %.2e: invert %.2e predict %.2e"
          % (eps, plan["W"], plan["sigma"], plan["err_est"], e_inv, e_pred), flush=True)
    assert plan["err_est"] <= eps
    assert e_inv < eps
    assert e_pred < eps


def test_epsilon_below_fp32_floor_is_refused():
    # the reference grids in FP64 there (wgridder/wgriddingmsgridder.cpp:43-48);
    # this library refuses with a message naming its floor
    from wsclean_amd import WgError
    p = synth.small_problem(seed=1001, n_ant=16, n_times=24, n_chan=2, npix=96)
    g = WGriddingGridder_Simple(96, 96, 96, 96, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-8)
    g.InitializeInversion()
    with pytest.raises(WgError, match="FP32"):
        g.AddInversionData(p["uvw"].shape[0], 2, p["uvw"], p["freq"], p["vis"])
    g.close()


def coarse_key_problem():
    # the 128-antenna 10 km track, one of every 256 time steps, at the band
    # edges 120 / 180 MHz, imaged at 8192^2 with 6 arcsec pixels: the uv box
    # times the w planes needs more than 2^32 (tile, plane, cell) keys
    uvw, _, _, _ = synth.large_wgridding_inputs(64, 8192)
    sub = np.ascontiguousarray(uvw.reshape(-1, 8128, 3)[::256].reshape(-1, 3))
    freq = np.array([120e6, 180e6])
    rng = np.random.Generator(np.random.PCG64(66))
    dx = 6.0 * synth.ARCSEC
    l, m, flux, _, _ = synth.point_sky(rng, 25, 8192, dx)
    vis = synth.predict_sources(sub, freq, l, m, flux).astype(np.complex64)
    return dict(uvw=sub, freq=freq, vis=vis, npix=8192, dx=dx)


def test_coarse_sort_keys_bit_exact(oracle):
    p = coarse_key_problem()
    g = WGriddingGridder_Simple(8192, 8192, 8192, 8192, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-5)
    keys, perm, plan = g.sort_keys(p["uvw"], p["freq"], p["vis"])
    print("coarse keys: key_shift %d key_T %d nkeys %d" % (plan["key_shift"], plan["key_T"],
                                                           plan["nkeys"]), flush=True)
    assert plan["key_shift"] >= 1
    ref_keys, _ = oracle.bucket_keys(plan, p["uvw"], p["freq"], p["vis"])
    np.testing.assert_array_equal(keys, ref_keys)
    np.testing.assert_array_equal(perm, oracle.bucket_perm(plan, p["uvw"], p["freq"], p["vis"]))
    g.close()


def test_coarse_sort_keys_match_direct_ft(oracle):
    p = coarse_key_problem()
    e_inv, e_pred, plan = run_pair(oracle, p, 1e-5, None, n_pix_sample=300, n_vis_sample=3000)
    print("coarse keys plan W %d P %d Sg %d key_shift %d: invert %.2e predict %.2e"
          % (plan["W"], plan["P"], plan["Sg"], plan["key_shift"], e_inv, e_pred), flush=True)
    assert plan["key_shift"] >= 1
    assert e_inv < 1e-5
    assert e_pred < 1e-5
