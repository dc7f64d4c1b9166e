"""Split-precision matrix-core gridding (each FP32 operand as a pair of f16
halves on v_mfma_f32_16x16x32_f16, W <= 8) against the FP32 matrix-core
kernels on the same plan and against the FP64 direct FT.

The two arithmetics differ only by the rounding of the split products
(about 2^-21 of each product), so their images agree to ~1e-6 relative;
each is within the requested epsilon of the direct FT (relative L2,
doc/source/wgridding.rst:19-20).  Covered: W = 4, 6, 8 (forced), the fused
single-sub-slab gather and the unfused multi-sub-slab path, predict
(degridding) and the adjoint identity <G x, y> = <x, G^H y>.
"""
import numpy as np
import pytest

from test_gpu_wide import p96, p512, p512b, rel_l2
from wsclean_amd import WGriddingGridder_Simple

pytestmark = pytest.mark.gpu

PROBLEMS = {"96": p96, "512": p512, "512b": p512b}


def grid_image(p, eps, split, tuning=None):
    n = p["npix"]
    g = WGriddingGridder_Simple(n, n, n, n, p["dx"], p["dx"], 0.0, 0.0, 1, eps)
    g.set_precision(split)
    if tuning:
        g.set_tuning(**tuning)
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    g.InitializeInversion()
    g.AddInversionData(n_rows, n_chan, p["uvw"], p["freq"], p["vis"])
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(n, n)
    plan = g.plan()
    g.close()
    return img, plan


def predict(p, eps, split, model, tuning=None):
    n = p["npix"]
    g = WGriddingGridder_Simple(n, n, n, n, p["dx"], p["dx"], 0.0, 0.0, 1, eps)
    g.set_precision(split)
    if tuning:
        g.set_tuning(**tuning)
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    g.InitializePrediction(model.astype(np.float32).ravel())
    out = np.zeros((n_rows, n_chan), np.complex64)
    g.PredictVisibilities(n_rows, n_chan, p["uvw"], p["freq"], out)
    plan = g.plan()
    g.close()
    return out, plan


# forced shapes (support and oversampling) with an epsilon each reaches
CASES = [("96", 2e-2, {"W": 4, "sigma": 2.0}), ("96", 1e-3, {"W": 6, "sigma": 2.0}),
         ("96", 1e-5, {"W": 8, "sigma": 1.6}), ("512", 1e-5, {"W": 8, "sigma": 1.6}),
         ("512b", 1e-3, {"W": 6, "sigma": 2.0}),
         ("512b", 1e-5, {"W": 8, "sigma": 1.6, "sub_slab": 4})]


@pytest.mark.parametrize("name,eps,tuning", CASES)
def test_split_gridding_matches_fp32_and_direct_ft(oracle, name, eps, tuning):
    p = PROBLEMS[name]()
    n = p["npix"]
    print("split gridding case", name, eps, tuning, flush=True)
    img32, plan32 = grid_image(p, eps, 0, tuning)
    img16, plan16 = grid_image(p, eps, 1, tuning)
    assert plan32["split"] == 0 and plan16["split"] == 1
    assert plan32["W"] == plan16["W"] == tuning["W"]
    if "sub_slab" in tuning:
        assert plan16["Sg"] < plan16["P"]  # several sub-slabs: the unfused path
    d = rel_l2(img16, img32)
    rng = np.random.default_rng(3)
    if n <= 128:
        ref = oracle.dft_invert(n, n, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"], p["vis"])
        e32, e16 = rel_l2(img32, ref), rel_l2(img16, ref)
    else:
        xs, ys = rng.integers(0, n, 600), rng.integers(0, n, 600)
        ref = oracle.dft_invert_pixels(n, n, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"],
                                       p["vis"], xs, ys)
        e32, e16 = rel_l2(img32[ys, xs], ref), rel_l2(img16[ys, xs], ref)
    print("  split vs fp32 %.3e; vs direct FT: fp32 %.3e split %.3e" % (d, e32, e16),
          flush=True)
    assert d < 2e-6
    assert e16 < eps
    assert abs(e16 - e32) < 2e-6


@pytest.mark.parametrize("name,eps,tuning", CASES)
def test_split_degridding_matches_fp32_and_direct_ft(oracle, name, eps, tuning):
    p = PROBLEMS[name]()
    n = p["npix"]
    print("split degridding case", name, eps, tuning, flush=True)
    rng = np.random.default_rng(4)
    model = np.zeros((n, n))
    k = 24
    model[rng.integers(2, n - 2, k), rng.integers(2, n - 2, k)] = rng.uniform(-1.0, 1.0, k)
    out32, plan32 = predict(p, eps, 0, model, tuning)
    out16, plan16 = predict(p, eps, 1, model, tuning)
    assert plan32["split"] == 0 and plan16["split"] == 1
    d = rel_l2(out16, out32)
    sel = rng.choice(out16.size, min(out16.size, 2000), replace=False)
    refv = oracle.dft_predict(n, n, p["dx"], p["dx"], 0.0, 0.0, model, p["uvw"], p["freq"],
                              sel=sel)
    e32, e16 = rel_l2(out32.ravel()[sel], refv), rel_l2(out16.ravel()[sel], refv)
    print("  split vs fp32 %.3e; vs direct FT: fp32 %.3e split %.3e" % (d, e32, e16),
          flush=True)
    assert d < 2e-6
    assert e16 < eps
    assert abs(e16 - e32) < 2e-6


def test_split_adjoint():
    p = p512b()
    n = p["npix"]
    rng = np.random.default_rng(5)
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    y = (rng.normal(size=(n_rows, n_chan)) + 1j * rng.normal(size=(n_rows, n_chan))).astype(
        np.complex64)
    x = rng.normal(size=(n, n))
    g = WGriddingGridder_Simple(n, n, n, n, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-5)
    g.set_precision(1)
    g.InitializeInversion()
    g.AddInversionData(n_rows, n_chan, p["uvw"], p["freq"], y)
    g.FinalizeImage(1.0)
    ghy = g.RealImage().reshape(n, n).astype(np.float64)
    g.InitializePrediction(x.astype(np.float32).ravel())
    gx = np.zeros((n_rows, n_chan), np.complex64)
    g.PredictVisibilities(n_rows, n_chan, p["uvw"], p["freq"], gx)
    assert g.plan()["split"] == 1
    g.close()
    lhs = float(np.sum(x * ghy))
    rhs = float(np.real(np.vdot(gx.astype(np.complex128), y.astype(np.complex128))))
    scale = np.linalg.norm(x) * np.linalg.norm(ghy)
    print("split adjoint", lhs, rhs, abs(lhs - rhs) / scale, flush=True)
    assert abs(lhs - rhs) < 1e-5 * scale
