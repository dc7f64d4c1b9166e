"""A major cycle captured as a HIP graph: every replay gives the eager
result bit for bit, also with other eager device work and
synchronisations between replays, and replays read the bound buffers'
current contents."""
import numpy as np
import pytest

from wsclean_amd import WGriddingGridder_Simple, WgError, synth

pytestmark = pytest.mark.gpu


def test_major_cycle_graph_replays_match_eager():
    import torch
    dev = torch.device("cuda:0")
    p = synth.small_problem(seed=41, n_ant=16, n_times=30, n_chan=3, npix=128,
                            pixel_arcsec=15.0, noise=0.05)
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    uvw = torch.from_numpy(p["uvw"]).to(dev)
    vis = torch.from_numpy(p["vis"].view(np.float32).reshape(n_rows, n_chan, 2).copy()).to(dev)
    fac = 1.0 / (n_rows * n_chan)
    g = WGriddingGridder_Simple(128, 128, 128, 128, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-5)
    g.PrepareVisibilities(uvw, p["freq"], n_rows, n_chan)
    g.InitializeInversion()
    g.AddInversionPrepared(vis)
    g.FinalizeImage(fac)
    img_e = g.RealImage()
    out_e = torch.empty_like(vis)
    g.PredictPrepared(out_e)
    out_g = torch.full_like(vis, float("nan"))
    torch.cuda.synchronize()
    g.CaptureMajorCycle(vis, out_g, fac)
    other = WGriddingGridder_Simple(64, 64, 64, 64, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-4)
    for rep in range(4):
        out_g.fill_(float("nan"))
        g.InitializePrediction(np.zeros(128 * 128, np.float32))  # clobber the image
        # eager work of another handle and of torch between the replays
        other.InitializeInversion()
        other.AddInversionData(n_rows, n_chan, p["uvw"], p["freq"], p["vis"])
        torch.cuda.synchronize()
        g.ReplayMajorCycle(1)
        np.testing.assert_array_equal(g.RealImage(), img_e)
        assert torch.equal(out_g, out_e), "replay %d" % rep
    # a replay reads the current contents of the bound buffers
    vis.mul_(2.0)
    torch.cuda.synchronize()
    g.ReplayMajorCycle(1)
    np.testing.assert_allclose(g.RealImage(), 2 * img_e, rtol=1e-5,
                               atol=1e-6 * np.abs(img_e).max())
    # an invert-only graph replayed twice leaves the same image (it clears)
    g.CaptureMajorCycle(vis_dev=vis, factor=fac)
    g.ReplayMajorCycle(2)
    img2 = g.RealImage()
    g.ReplayMajorCycle(1)
    np.testing.assert_array_equal(g.RealImage(), img2)
    # a new prepare drops the graph
    g.PrepareVisibilities(uvw, p["freq"], n_rows, n_chan)
    with pytest.raises(WgError):
        g.ReplayMajorCycle(1)
    g.close()
    other.close()


def test_facet_batch_graph_replays_match_eager():
    import torch
    from test_gpu_facets import facet_problem

    from wsclean_amd.facets import FacetBatch
    p = facet_problem(seed=21)
    fb = FacetBatch(p["facets"])
    ro, nr = [r[0] for r in p["rows"]], [r[1] for r in p["rows"]]
    fb.Prepare(p["uvw_d"], p["freq"], ro, nr)
    factors = np.array([1.0 / max(1, n * 2) for n in nr])
    fb.InitializeInversion()
    fb.AddInversion(p["vis_d"])
    for f in range(len(nr)):
        fb.FinalizeImage(f, factors[f])
    imgs = [fb.RealImage(f) for f in range(len(nr))]
    n_virtual = int(fb.block_offsets()[-1])
    out_e = torch.empty((n_virtual, 2), dtype=torch.float32, device="cuda:0")
    fb.Predict(out_e)
    out_g = torch.full_like(out_e, float("nan"))
    torch.cuda.synchronize()
    fb.CaptureMajorCycle(p["vis_d"], out_g, factors)
    for rep in range(3):
        out_g.fill_(float("nan"))
        torch.cuda.synchronize()
        fb.ReplayMajorCycle(1)
        for f in range(len(nr)):
            np.testing.assert_array_equal(fb.RealImage(f), imgs[f])
        assert torch.equal(out_g, out_e), "replay %d" % rep
    fb.close()


def test_graph_refuses_replay_after_buffers_move():
    # an eager call that needs larger shared buffers (box, line buffers,
    # subgrid scratch) than the captured cycle reallocates them: the replay
    # then refuses instead of touching freed memory, and a new capture works
    import torch
    dev = torch.device("cuda:0")
    p = synth.small_problem(seed=41, n_ant=16, n_times=30, n_chan=3, npix=128,
                            pixel_arcsec=15.0, noise=0.05)
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    uvw = torch.from_numpy(p["uvw"]).to(dev)
    vis = torch.from_numpy(p["vis"].view(np.float32).reshape(n_rows, n_chan, 2).copy()).to(dev)
    g = WGriddingGridder_Simple(128, 128, 128, 128, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-5)
    g.PrepareVisibilities(uvw, p["freq"], n_rows, n_chan)
    out = torch.empty_like(vis)
    g.CaptureMajorCycle(vis, out, 1.0)
    img_e = g.RealImage()
    out_e = out.clone()
    # same-size eager calls on the handle keep the graph
    g.InitializeInversion()
    g.AddInversionData(n_rows, n_chan, p["uvw"], p["freq"], p["vis"])
    g.ReplayMajorCycle(1)
    np.testing.assert_array_equal(g.RealImage(), img_e)
    # a wider uv box (longer baselines, finer tiling) grows the shared buffers
    big = synth.small_problem(seed=42, n_ant=24, n_times=40, n_chan=3, npix=128,
                              pixel_arcsec=15.0, radius_m=6000.0)
    g.set_tuning(sub_slab=2)
    g.InitializeInversion()
    g.AddInversionData(big["uvw"].shape[0], 3, big["uvw"], big["freq"], big["vis"])
    g.set_tuning()
    with pytest.raises(WgError):
        g.ReplayMajorCycle(1)
    with pytest.raises(WgError):   # the graph was dropped
        g.ReplayMajorCycle(1)
    g.CaptureMajorCycle(vis, out, 1.0)
    g.ReplayMajorCycle(2)
    np.testing.assert_array_equal(g.RealImage(), img_e)
    assert torch.equal(out, out_e)
    g.close()
