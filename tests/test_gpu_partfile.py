"""Gridding a reordered part straight from its files (wg_part_grid: disk
reads overlapped with the device's weighting and gridding) against the same
rows read into memory, weighted with wg_prepare_rows and gridded with
AddInversionData, and against the FP64 direct FT of the weighted
visibilities (tolerance epsilon = 1e-5, doc/source/wgridding.rst:19-20)."""
import numpy as np
import pytest

from test_partfile import write_part
from wsclean_amd import WGriddingGridder_Simple, synth
from wsclean_amd.partfile import ReorderedPart
from wsclean_amd.weighting import prepare_rows

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def test_part_grid_matches_in_memory_path_and_direct_ft(oracle, tmp_path):
    import torch
    n_chan = 16
    rng = np.random.Generator(np.random.PCG64(9))
    enu = synth.antennas_disc(rng, 40, 2000.0)
    uvw = synth.earth_rotation_uvw(enu, 26, 60.0, np.deg2rad(-30.0))[:20000]
    freq = np.linspace(140e6, 160e6, n_chan)
    dx, n = 10.0 * synth.ARCSEC, 512
    l, m, flux, _, _ = synth.point_sky(rng, 12, n, dx)
    vis = synth.add_noise(rng, synth.predict_sources(uvw, freq, l, m, flux), 0.05)
    vis = vis.astype(np.complex64)[:, :, None]
    paths, v = write_part(tmp_path, n_rows=20000, n_chan=n_chan, n_pol=1, seed=9, uvw=uvw,
                          vis=vis)
    wts = v["weights"]
    part = ReorderedPart(paths["meta"], paths["data"], paths["weights"], n_pol=1)
    g = WGriddingGridder_Simple(n, n, n, n, dx, dx, 0.0, 0.0, 1, 1e-5)
    g.InitializeInversion()
    st = part.Grid(g, freq, chunk_rows=3001)
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(n, n)
    print("part grid stats", st, flush=True)
    # the same rows through memory: wg_prepare_rows + AddInversionData
    dev = torch.device("cuda:0")
    rows = part.Read()
    uvw_d = torch.from_numpy(rows["uvw"]).to(dev)
    wts_d = torch.from_numpy(rows["weights"]).to(dev)
    vis_d = torch.from_numpy(rows["vis"].view(np.float32).reshape(20000, n_chan, 1, 2)).to(dev)
    out = torch.empty((20000, n_chan, 2), dtype=torch.float32, device=dev)
    ref_st = prepare_rows(g, uvw_d, freq, wts_d, out, vis_dev=vis_d)
    g2 = WGriddingGridder_Simple(n, n, n, n, dx, dx, 0.0, 0.0, 1, 1e-5)
    g2.InitializeInversion()
    g2.AddInversionDataDevice(uvw_d, freq, out)
    g2.FinalizeImage(1.0)
    img2 = g2.RealImage().reshape(n, n)
    for k in ("total_weight", "visibility_weight_sum", "max_gridded_weight"):
        assert st[k] == pytest.approx(ref_st[k], rel=1e-12), k
    assert st["gridded_visibility_count"] == ref_st["gridded_visibility_count"]
    d = rel_l2(img, img2)
    weighted = (vis[:, :, 0] * wts[:, :, 0]).astype(np.complex64)
    xs, ys = rng.integers(0, n, 400), rng.integers(0, n, 400)
    ref = oracle.dft_invert_pixels(n, n, dx, dx, 0.0, 0.0, uvw, freq, weighted, xs, ys)
    e = rel_l2(img[ys, xs], ref)
    print("part grid vs in-memory %.2e, vs direct FT %.2e" % (d, e), flush=True)
    assert d < 1e-6
    assert e < 1e-5
    g.close()
    g2.close()
    part.close()
