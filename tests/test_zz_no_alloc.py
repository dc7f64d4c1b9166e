"""The allocation-free row preparation (wg_prepare_rows_noalloc) through the
C-ABI against wg_prepare_rows on the same inputs: two calls in a row, the
second with fewer rows and channels, so the shared device scratch keeps the
first call's contents when the second runs.  The two entry points run the
same kernels, so the outputs and weight sums agree to FP32 rounding
(tolerance 1e-6 relative) and are expected to be bitwise equal."""
import numpy as np
import pytest

from wsclean_amd import WGriddingGridder_Simple, WgError, synth
from wsclean_amd.weighting import UNIFORM, ImageWeights, prepare_rows

pytestmark = pytest.mark.gpu


def rows_case(seed, n_rows, n_chan, n_pol):
    import torch
    rng = np.random.default_rng(seed)
    enu = synth.antennas_disc(rng, 24, 1500.0)
    uvw = synth.earth_rotation_uvw(enu, n_rows // 276 + 1, 60.0, np.deg2rad(-30.0))[:n_rows]
    freq = np.linspace(120e6, 180e6, n_chan)
    wts = rng.uniform(0.5, 2.0, (n_rows, n_chan, n_pol)).astype(np.float32)
    wts[rng.random((n_rows, n_chan, n_pol)) < 0.05] = 0.0
    vis = (rng.normal(size=wts.shape) + 1j * rng.normal(size=wts.shape)).astype(np.complex64)
    dev = torch.device("cuda:0")
    as_dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return (freq, as_dev(uvw), as_dev(wts),
            as_dev(vis.view(np.float32).reshape(*vis.shape, 2)))


def test_prepare_rows_noalloc_matches_prepare_rows():
    import torch
    psx = 8.0 / 206265.0
    g = WGriddingGridder_Simple(64, 64, 64, 64, psx, psx, 0.0, 0.0, 1, 1e-4)
    calls = [  # (seed, rows, channels, polarizations, selection, polarization, image weights)
        (5, 3000, 6, 4, [1, 0, 1, 1, 1, 0], 3, True),
        (6, 700, 3, 2, None, 1, False),
    ]
    for seed, n_rows, n_chan, n_pol, sel, pol, use_iw in calls:
        print("noalloc call: %d rows, %d channels, %d pol" % (n_rows, n_chan, n_pol), flush=True)
        freq, uvw_d, wts_d, vis_d = rows_case(seed, n_rows, n_chan, n_pol)
        iw = None
        if use_iw:
            iw = ImageWeights(UNIFORM, 200, 200, psx, psx)
            iw.Grid(uvw_d, freq, wts_d)
            iw.FinishGridding()
        outs, stats = [], []
        for no_alloc in (False, True):
            out = torch.full((n_rows, n_chan, 2), float("nan"), dtype=torch.float32,
                             device=uvw_d.device)
            st = prepare_rows(g, uvw_d, freq, wts_d, out, vis_dev=vis_d, selected=sel,
                              polarization=pol, image_weights=iw, no_alloc=no_alloc)
            outs.append(out.cpu().numpy())
            stats.append(st)
        ref, got = outs
        assert np.isfinite(got).all()
        print("  bitwise equal:", bool(np.array_equal(ref, got)), flush=True)
        np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())
        for k in ("total_weight", "visibility_weight_sum", "max_gridded_weight"):
            assert stats[1][k] == pytest.approx(stats[0][k], rel=1e-12), k
        assert stats[1]["gridded_visibility_count"] == stats[0]["gridded_visibility_count"]
        if iw is not None:
            iw.close()
    g.close()


def test_prepare_rows_noalloc_rejects_too_many_channels():
    # refused on the host before any device work
    import torch
    psx = 8.0 / 206265.0
    g = WGriddingGridder_Simple(64, 64, 64, 64, psx, psx, 0.0, 0.0, 1, 1e-4)
    n_chan = 65537
    dev = torch.device("cuda:0")
    uvw = torch.zeros((1, 3), dtype=torch.float64, device=dev)
    wts = torch.ones((1, n_chan, 1), dtype=torch.float32, device=dev)
    vis = torch.zeros((1, n_chan, 1, 2), dtype=torch.float32, device=dev)
    out = torch.zeros((1, n_chan, 2), dtype=torch.float32, device=dev)
    freq = np.linspace(100e6, 200e6, n_chan)
    with pytest.raises(WgError, match="65536 channels"):
        prepare_rows(g, uvw, freq, wts, out, vis_dev=vis, no_alloc=True)
    g.close()
