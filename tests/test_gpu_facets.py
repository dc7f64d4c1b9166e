"""Facet batch on the GPU: every facet's image and predicted visibilities
equal those of its own single gridder on the same rows (the batch shares
the sort and the gridding launches, not the arithmetic), and match the
FP64 direct FT to epsilon."""
import numpy as np
import pytest

from wsclean_amd import WGriddingGridder_Simple, synth
from wsclean_amd.facets import FacetBatch

pytestmark = pytest.mark.gpu

ARCSEC = synth.ARCSEC
EPS = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def facet_problem(seed=11):
    import torch
    rng = np.random.default_rng(seed)
    enu = synth.antennas_disc(rng, 20, 1500.0)
    uvw = synth.earth_rotation_uvw(enu, 30, 60.0, np.deg2rad(-30.0))
    freq = np.array([130e6, 170e6])
    n_rows = uvw.shape[0]
    dx = 12 * ARCSEC
    facets = [
        dict(width=96, height=96, width_t=96, height_t=96, l_shift=0.0, m_shift=0.0),
        dict(width=64, height=80, width_t=64, height_t=80, l_shift=0.012, m_shift=-0.008),
        dict(width=128, height=64, width_t=128, height_t=64, l_shift=-0.015, m_shift=0.01),
        dict(width=48, height=48, width_t=48, height_t=48, l_shift=0.02, m_shift=0.02),
        dict(width=80, height=80, width_t=64, height_t=64, l_shift=-0.006, m_shift=-0.018),
    ]
    for f in facets:
        f.update(pixel_size_x=dx, pixel_size_y=dx, epsilon=EPS)
    # facets take overlapping row ranges; facet 3 has none
    rows = [(0, n_rows), (0, 2000), (1500, n_rows - 1500), (300, 0), (100, 800)]
    sky_l = rng.uniform(-0.02, 0.02, 12)
    sky_m = rng.uniform(-0.02, 0.02, 12)
    vis = synth.predict_sources(uvw, freq, sky_l, sky_m, rng.uniform(0.1, 1.0, 12))
    vis = synth.add_noise(rng, vis, 0.05).astype(np.complex64)
    dev = torch.device("cuda:0")
    uvw_d = torch.from_numpy(uvw).to(dev)
    vis_d = torch.from_numpy(vis.view(np.float32).reshape(n_rows, 2, 2).copy()).to(dev)
    return dict(uvw=uvw, freq=freq, vis=vis, facets=facets, rows=rows, uvw_d=uvw_d,
                vis_d=vis_d)


def single_gridder(f, plan):
    g = WGriddingGridder_Simple(f["width"], f["height"], f["width_t"], f["height_t"],
                                f["pixel_size_x"], f["pixel_size_y"], f["l_shift"],
                                f["m_shift"], 1, f["epsilon"])
    g.set_tuning(W=plan["W"], sigma=plan["sigma"], tile=plan["T"], sub_slab=plan["Sg"])
    return g


def test_facet_batch_invert_matches_single_gridders_and_direct_ft(oracle):
    p = facet_problem()
    fb = FacetBatch(p["facets"])
    fb.Prepare(p["uvw_d"], p["freq"], [r[0] for r in p["rows"]], [r[1] for r in p["rows"]])
    plan = fb.plan()
    print("atlas", {k: plan["atlas"][k] for k in ("ntu", "ntv", "P", "Sg", "Sf", "W")},
          flush=True)
    fb.InitializeInversion()
    fb.AddInversion(p["vis_d"])
    first = []
    for f, (r0, nr) in enumerate(p["rows"]):
        fac = p["facets"][f]
        nvis = max(1, nr * 2)
        fb.FinalizeImage(f, 1.0 / nvis)
        img = fb.RealImage(f).reshape(fac["height"], fac["width"])
        first.append(img)
        if nr == 0:
            assert np.all(img == 0)
            continue
        g = single_gridder(fac, plan["facets"][f]["plan"])
        g.PrepareVisibilities(p["uvw_d"][r0:r0 + nr].contiguous(), p["freq"], nr, 2)
        g.InitializeInversion()
        g.AddInversionPrepared(p["vis_d"][r0:r0 + nr].contiguous())
        g.FinalizeImage(1.0 / nvis)
        ref = g.RealImage().reshape(fac["height"], fac["width"])
        sp = g.plan()
        for k in ("nu", "nv", "P", "BU", "BV", "box_u0", "box_v0"):
            assert sp[k] == plan["facets"][f]["plan"][k], (f, k)
        np.testing.assert_array_equal(np.isnan(img), np.isnan(ref))
        ok = ~np.isnan(ref)
        err = rel_l2(img[ok], ref[ok])
        print("facet", f, "batch vs single rel_l2", err, flush=True)
        assert err < 2e-6
        # FP64 direct FT of the facet's rows on its own (trimmed) image
        wt, ht = fac["width_t"], fac["height_t"]
        dft = oracle.dft_invert(wt, ht, fac["pixel_size_x"], fac["pixel_size_y"],
                                fac["l_shift"], fac["m_shift"], p["uvw"][r0:r0 + nr],
                                p["freq"], p["vis"][r0:r0 + nr]) / nvis
        y0, x0 = (fac["height"] - ht) // 2, (fac["width"] - wt) // 2
        err = rel_l2(img[y0:y0 + ht, x0:x0 + wt], dft)
        print("facet", f, "vs direct FT rel_l2", err, flush=True)
        assert err < EPS
    # the batch is deterministic
    fb.InitializeInversion()
    fb.AddInversion(p["vis_d"])
    for f, (r0, nr) in enumerate(p["rows"]):
        fb.FinalizeImage(f, 1.0 / max(1, nr * 2))
        np.testing.assert_array_equal(
            fb.RealImage(f).reshape(first[f].shape), first[f])
    fb.close()


def test_facet_batch_predict_matches_single_gridders_and_direct_ft(oracle):
    import torch
    p = facet_problem(seed=12)
    fb = FacetBatch(p["facets"])
    fb.Prepare(p["uvw_d"], p["freq"], [r[0] for r in p["rows"]], [r[1] for r in p["rows"]])
    plan = fb.plan()
    rng = np.random.default_rng(5)
    models = []
    for f, fac in enumerate(p["facets"]):
        m = np.zeros((fac["height"], fac["width"]), np.float32)
        k = 8
        m[rng.integers(4, fac["height"] - 4, k), rng.integers(4, fac["width"] - 4, k)] = \
            rng.uniform(-1, 1, k)
        models.append(m)
        if f % 2:
            fb.InitializePredictionDevice(f, torch.from_numpy(m.ravel()).to("cuda:0"))
        else:
            fb.InitializePrediction(f, m)
    blocks = fb.block_offsets()
    out = torch.full((int(blocks[-1]), 2), float("nan"), dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()  # the fill runs on torch's stream, the library on its own
    fb.Predict(out)
    outh = out.cpu().numpy().view(np.complex64).ravel()
    assert np.isfinite(outh).all()
    for f, (r0, nr) in enumerate(p["rows"]):
        if nr == 0:
            assert blocks[f + 1] == blocks[f]
            continue
        fac = p["facets"][f]
        got = outh[blocks[f]:blocks[f + 1]].reshape(nr, 2)
        g = single_gridder(fac, plan["facets"][f]["plan"])
        g.PrepareVisibilities(p["uvw_d"][r0:r0 + nr].contiguous(), p["freq"], nr, 2)
        g.InitializePrediction(models[f])
        ref = torch.empty((nr, 2, 2), dtype=torch.float32, device="cuda:0")
        g.PredictPrepared(ref)
        refh = ref.cpu().numpy().view(np.complex64).reshape(nr, 2)
        err = rel_l2(got, refh)
        print("facet", f, "predict batch vs single rel_l2", err, flush=True)
        assert err < 1e-6
        wt, ht = fac["width_t"], fac["height_t"]
        y0, x0 = (fac["height"] - ht) // 2, (fac["width"] - wt) // 2
        dft = oracle.dft_predict(wt, ht, fac["pixel_size_x"], fac["pixel_size_y"],
                                 fac["l_shift"], fac["m_shift"],
                                 models[f][y0:y0 + ht, x0:x0 + wt], p["uvw"][r0:r0 + nr],
                                 p["freq"])
        err = rel_l2(got, dft)
        print("facet", f, "predict vs direct FT rel_l2", err, flush=True)
        assert err < EPS
    fb.close()


def test_facet_batch_reprepare_and_trim_border():
    p = facet_problem(seed=13)
    fb = FacetBatch(p["facets"])
    # a second prepare with other rows replaces the first
    fb.Prepare(p["uvw_d"], p["freq"], [0] * 5, [10] * 5)
    fb.Prepare(p["uvw_d"], p["freq"], [r[0] for r in p["rows"]], [r[1] for r in p["rows"]])
    fb.InitializeInversion()
    fb.AddInversion(p["vis_d"])
    fb.FinalizeImage(4, 1.0)
    img = fb.RealImage(4).reshape(80, 80)
    assert np.isnan(img[:8]).all() and np.isnan(img[:, :8]).all()
    assert np.isfinite(img[8:72, 8:72]).all()
    with pytest.raises(Exception):
        fb.RealImage(5)
    fb.close()


def test_facet_batch_invert_uniform_weighted_dynamic_range(oracle):
    # 40 antennas in a log-radial 3 km layout, 120 x 60 s, 4 channels:
    # 3.7e5 visibilities times uniform density weights on a 192-cell uv grid
    # (about four and a half decades), plus 20 long-baseline visibilities
    # x 1e3.  Each facet's work items scale their fixed-point accumulation
    # by bounds on their own visibilities (the facet's row maxima gathered
    # through the atlas rows), so every facet keeps epsilon.
    import torch
    rng = np.random.Generator(np.random.PCG64(41))
    enu = synth.antennas_logradial(rng, 40, 3000.0)
    uvw = synth.earth_rotation_uvw(enu, 120, 60.0, np.deg2rad(-30.0))
    freq = 140e6 + 8e6 * np.arange(4)
    n_rows = uvw.shape[0]
    dx = 12 * ARCSEC
    vis = synth.predict_sources(uvw, freq, rng.uniform(-0.02, 0.02, 12),
                                rng.uniform(-0.02, 0.02, 12), rng.uniform(0.1, 1.0, 12))
    ncell = 192
    s = freq[None, :] / 299792458.0 * ncell * dx
    u, v = uvw[:, 0:1] * s, uvw[:, 1:2] * s
    flip = v < 0
    u, v = np.where(flip, -u, u), np.where(flip, -v, v)
    cell = (np.floor(u).astype(np.int64) % ncell) * ncell + np.floor(v).astype(np.int64) % ncell
    wt = 1.0 / np.bincount(cell.ravel(), minlength=ncell * ncell)[cell]
    assert wt.max() / wt.min() > 1e4
    vis = (vis * wt).astype(np.complex64)
    far = np.argsort(np.hypot(uvw[:, 0], uvw[:, 1]))[-20:]
    vis[far, 0] *= 1e3
    facets = [
        dict(width=96, height=96, width_t=96, height_t=96, l_shift=0.0, m_shift=0.0),
        dict(width=64, height=80, width_t=64, height_t=80, l_shift=0.012, m_shift=-0.008),
        dict(width=128, height=64, width_t=128, height_t=64, l_shift=-0.015, m_shift=0.01),
        dict(width=80, height=80, width_t=64, height_t=64, l_shift=-0.006, m_shift=-0.018),
    ]
    for f in facets:
        f.update(pixel_size_x=dx, pixel_size_y=dx, epsilon=EPS)
    rows = [(0, n_rows), (0, 30000), (20000, n_rows - 20000), (1000, 12000)]
    dev = torch.device("cuda:0")
    uvw_d = torch.from_numpy(uvw).to(dev)
    vis_d = torch.from_numpy(vis.view(np.float32).reshape(n_rows, len(freq), 2).copy()).to(dev)
    fb = FacetBatch(facets)
    fb.Prepare(uvw_d, freq, [r[0] for r in rows], [r[1] for r in rows])
    assert fb.plan()["atlas"]["W"] <= 8
    fb.InitializeInversion()
    fb.AddInversion(vis_d)
    prng = np.random.default_rng(3)
    for f, (r0, nr) in enumerate(rows):
        fac = facets[f]
        fb.FinalizeImage(f, 1.0)
        img = fb.RealImage(f).reshape(fac["height"], fac["width"])
        wt_, ht = fac["width_t"], fac["height_t"]
        y0, x0 = (fac["height"] - ht) // 2, (fac["width"] - wt_) // 2
        xs = prng.integers(0, wt_, 200)
        ys = prng.integers(0, ht, 200)
        ref = oracle.dft_invert_pixels(wt_, ht, dx, dx, fac["l_shift"], fac["m_shift"],
                                       uvw[r0:r0 + nr], freq, vis[r0:r0 + nr], xs, ys)
        err = rel_l2(img[y0 + ys, x0 + xs], ref)
        print("facet", f, "uniform-weighted vs direct FT rel_l2", err, flush=True)
        assert err < EPS
    fb.close()
