"""Autotuner: the trials are measured and ordered, the handle keeps the
fastest, and the tuned plan still meets epsilon against the FP64 direct
FT; a handle created with tuning on tunes itself."""
import numpy as np
import pytest

from wsclean_amd import WGriddingGridder_Simple, synth

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def test_autotune_orders_trials_and_keeps_the_fastest(oracle):
    import torch
    p = synth.small_problem(seed=51, n_ant=16, n_times=30, n_chan=3, npix=128,
                            pixel_arcsec=15.0, noise=0.02)
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    uvw = torch.from_numpy(p["uvw"]).to("cuda:0")
    vis = torch.from_numpy(p["vis"].view(np.float32).reshape(n_rows, n_chan, 2).copy()).to(
        "cuda:0")
    g = WGriddingGridder_Simple(128, 128, 128, 128, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-5)
    trials = g.Autotune(uvw, p["freq"], n_rows, n_chan, max_sample_vis=4000)
    print(trials, flush=True)
    ok = [t for t in trials if t["ok"]]
    assert len(ok) >= 3
    est = [t["est_ms"] for t in ok]
    assert est == sorted(est) and est[0] > 0
    assert any(t["T"] == 16 for t in ok)
    best = ok[0]
    g.PrepareVisibilities(uvw, p["freq"], n_rows, n_chan)
    plan = g.plan()
    assert (plan["W"], plan["T"]) == (best["W"], best["T"])
    assert abs(plan["sigma"] - best["sigma"]) < 1e-3
    g.InitializeInversion()
    g.AddInversionPrepared(vis)
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(128, 128)
    ref = oracle.dft_invert(128, 128, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"], p["vis"])
    assert rel_l2(img, ref) < 1e-5
    g.close()


def test_tuning_flag_tunes_the_handle(oracle):
    p = synth.small_problem(seed=52, n_ant=14, n_times=20, n_chan=2, npix=96, noise=0.02)
    g = WGriddingGridder_Simple(96, 96, 96, 96, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-5,
                                tuning=True)
    g.InitializeInversion()
    g.AddInversionData(p["uvw"].shape[0], len(p["freq"]), p["uvw"], p["freq"], p["vis"])
    g.FinalizeImage(1.0)
    img = g.RealImage().reshape(96, 96)
    ref = oracle.dft_invert(96, 96, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"], p["vis"])
    assert rel_l2(img, ref) < 1e-5
    # a second handle of the same geometry reuses the choice
    g2 = WGriddingGridder_Simple(96, 96, 96, 96, p["dx"], p["dx"], 0.0, 0.0, 1, 1e-5,
                                 tuning=True)
    g2.InitializeInversion()
    g2.AddInversionData(p["uvw"].shape[0], len(p["freq"]), p["uvw"], p["freq"], p["vis"])
    assert g2.plan()["W"] == g.plan()["W"] and g2.plan()["sigma"] == g.plan()["sigma"]
    g.close()
    g2.close()
