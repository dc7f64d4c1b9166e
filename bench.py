"""Benchmark: grid + degrid throughput of the MI355X w-gridder.

Workload (default): 8192 x 8192 image, 2 arcsec pixels, 128 antennas in a
seeded 10 km log-radial layout (8128 baselines), 2880 x 10 s at dec -30,
64 channels 120-180 MHz (1.498e9 visibilities), 500 power-law point
sources + 5 Gaussian blobs + noise 0.05, epsilon 1e-5.  Inputs are
generated on the device and resident in HBM before the timed region.

One step = one major cycle on the GPU: plan + bucket sort of the
visibilities (wg_prepare_visibilities), invert (grid) and predict (degrid)
of all visibilities.  value = N_vis / step time (vis/s, grid + degrid).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--channels C]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--channels", type=int, default=64)
    ap.add_argument("--npix", type=int, default=8192)
    ap.add_argument("--eps", type=float, default=1e-5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-per-channel", action="store_true",
                    help="skip the one-call-per-channel GPU comparison leg")
    ap.add_argument("--cpu-stride", type=int, default=200,
                    help="CPU baseline sample: every n-th small-w row of the track")
    ap.add_argument("--serve-requests", type=int, default=40,
                    help="requests in the served-stream leg (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=16.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the served-stream leg and print its JSON")
    ap.add_argument("--no-autotune", action="store_true", help="skip the autotuner leg")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph leg")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the direct-FT error of the timed results")
    ap.add_argument("--no-host-leg", action="store_true",
                    help="skip the host-buffer (PCIe-inclusive) leg")
    ap.add_argument("--facet-steps", type=int, default=2,
                    help="timed steps of the facet-batch leg (0 skips it)")
    ap.add_argument("--facets-only", action="store_true",
                    help="run only the facet-batch leg and print its JSON")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--sg", type=int, default=0)
    ap.add_argument("--W", type=int, default=0)
    ap.add_argument("--sigma", type=float, default=0.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the calibration "
                         "kernel for the FETCH_SIZE / WRITE_SIZE counters) and exit")
    return ap.parse_args()


def copy_probe(mib):
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "copies": 5}))


def make_inputs(n_chan, npix, seed=1002):
    from wsclean_amd import synth
    return synth.large_wgridding_inputs(n_chan, npix, seed)


def roofline_bytes(plan, n_vis, n_rows, nx, ny):
    """Compulsory traffic of one grid + degrid (CTM): vis + index + uvw per
    direction, and per w plane the flush, a 2-pass FFT and the image
    accumulate / model read, evaluated on the plan used (nu, nv, P)."""
    G = 8.0 * plan["nu"] * plan["nv"]
    P = plan["P"]
    b_vis = 12.0 * n_vis + 24.0 * n_rows
    inv = 5 * G + 16.0 * nx * ny
    pred = 6 * G + 4.0 * nx * ny
    return 2 * b_vis + P * (inv + pred)


MFMA_F32_PEAK = 157.3e12   # MI355X dense FP32 matrix FLOP/s


def _kernel_entry(name, ms, launches, flops, nbytes, W):
    """Roofline entry of one gridding kernel from its HIP-event time."""
    t = ms / launches / 1e3
    gbs = nbytes / launches / t / 1e9
    out = {"kernel": name, "launch_ms": t * 1e3, "launches": int(launches),
           "flops_per_launch": flops / launches, "bytes_per_launch": nbytes / launches,
           "hbm_view": {"achieved": gbs, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": gbs / (HBM_PEAK / 1e9)}, "traffic": None}
    if W <= 8:
        ach = flops / launches / t / 1e12
        out.update({"bound": "mfma", "achieved": ach, "peak": MFMA_F32_PEAK / 1e12,
                    "unit": "TFLOP/s", "frac": ach / (MFMA_F32_PEAK / 1e12)})
    else:
        out.update({"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                    "frac": gbs / (HBM_PEAK / 1e9)})
    return out


def kernel_roofline(st, plan):
    """Rooflines of the two gridding kernels from their own HIP events on
    the engine stream; the slower one is the dominant kernel.

    For W <= 8 both run on the matrix cores: GridKernelSplit<W, fused> /
    DegridKernelSplit<W, fused> (FP32 operands as f16 pairs on the f16
    instructions) where the plan says "split", else GridKernelMfma /
    DegridKernelMfma (FP32 instructions).  Bound "mfma", algorithmic FLOPs =
    4 W^3 per visibility pass (each of the W^3 footprint cells takes a
    complex x real multiply-add, counted once, in FP32 terms), against the
    157.3 TFLOP/s dense FP32 MFMA peak: the FP32-equivalent work, whatever
    instructions carry it.  With one
    plane sub-slab (P <= Sg) the kernels gather / scatter the visibilities
    themselves (fused).  The HBM view is reported beside it: algorithmic
    bytes per launch = per visibility pass 16 B coordinates + 8 B visibility
    (fused: + 4 B permutation word + 8 B phasor), read once per sub-slab the
    visibility touches, + 8 B x sub_n x E^2 per work item (its subgrid
    written or read once), against 8 TB/s.  traffic = measured HBM bytes per
    launch (rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel
    calibration) from the committed counter summary of that kernel and plan,
    else None."""
    if not st["grid_launches"] or not st["degrid_launches"]:
        return None
    W = plan["W"]
    mfma = W <= 8
    fused = mfma and plan["P"] <= plan["Sg"]
    fl = "true" if fused else "false"
    if mfma and plan.get("split"):
        names = ("GridKernelSplit<%d, %s>" % (W, fl), "DegridKernelSplit<%d, %s>" % (W, fl))
    elif mfma:
        names = ("GridKernelMfma<%d, %s>" % (W, fl), "DegridKernelMfma<%d, %s>" % (W, fl))
    else:
        names = ("GridKernel<%d>" % W, "DegridKernel<%d>" % W)
    grid = _kernel_entry(names[0], st["grid_kernel_ms"], st["grid_launches"],
                         4.0 * W ** 3 * st["grid_passes"], st["grid_bytes"], W)
    degrid = _kernel_entry(names[1], st["degrid_kernel_ms"], st["degrid_launches"],
                           4.0 * W ** 3 * st["degrid_passes"], st["degrid_bytes"], W)
    for e, stem in ((grid, "gridkernel"), (degrid, "degridkernel")):
        path = os.path.join(ROOT, "profiles", "%s_8192_traffic.json" % stem)
        if os.path.exists(path):
            with open(path) as f:
                tr = json.load(f)
            if tr.get("kernel") == e["kernel"] and \
                    tr.get("plan") == {k: plan[k] for k in tr.get("plan", {})}:
                e["traffic"] = tr["bytes_per_dispatch"]
    main, other = (degrid, grid) if degrid["launch_ms"] >= grid["launch_ms"] else (grid, degrid)
    out = dict(main)
    out["other_gridder"] = other
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(uvw_h, freq, dx, src, npix, plan, stride=200, w_max_m=100.0):
    """FP64 CPU w-gridder (oracle/wgridder_cpu.cpp, 'CPU restatement, not
    ducc0') on a like-for-like sample of the same workload, all channels,
    the full image and the same kernel and grid (W, beta, sigma, nu, nv).
    The CPU times its visibility work (set-up, gridding, degridding) and its
    plane work (FFTs, screens) separately.  The sample is every
    `stride`-th of the track's rows with |w| < w_max_m: those rows span
    nearly the full uv box (so the plane work per plane is the full job's)
    but only the minimum W planes (so the sweep stays short).  The full
    job's CPU time is the visibility work scaled by the visibility ratio
    plus the plane work per plane times the job's plane count.  Returns
    invert + predict visibilities per second of that full-job time."""
    from oracle import oracle
    from wsclean_amd import synth
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    sel = np.flatnonzero(np.abs(uvw_h[:, 2]) < w_max_m)[::stride]
    sub = np.ascontiguousarray(uvw_h[sel])
    l, m, S = src[:64, 0], src[:64, 1], src[:64, 3]
    vis = synth.predict_sources(sub[:20000], freq, l, m, S).astype(np.complex64)
    vis = np.resize(vis, (sub.shape[0], len(freq))).astype(np.complex64)
    kw = dict(W=plan["W"], beta=plan["beta"], sigma=plan["sigma"], nu=plan["nu"],
              nv=plan["nv"], nthreads=cores)
    t = time.perf_counter()
    img = oracle.cpu_wgrid_invert(npix, npix, dx, dx, 0.0, 0.0, sub, freq, vis, **kw)
    tv_i, tp_i, planes_i = oracle.cpu_wgrid_last_timing()
    oracle.cpu_wgrid_predict(npix, npix, dx, dx, 0.0, 0.0, img, sub, freq, **kw)
    tv_p, tp_p, planes_p = oracle.cpu_wgrid_last_timing()
    wall = time.perf_counter() - t
    n_sample = sub.shape[0] * len(freq)
    n_full = uvw_h.shape[0] * len(freq)
    full_s = (tv_i + tv_p) * (n_full / n_sample) + (tp_i + tp_p) * plan["P"] / max(1, planes_i)
    return {"value": n_full / full_s, "unit": "vis/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": "CPU restatement (FP64, not ducc0): invert+predict of every %dth row with "
                      "|w| < %g m (%d rows) x all %d channels = %d visibilities, %d^2 image, same "
                      "kernel and grid W=%d sigma=%.2f nu=%d, %d w planes; visibility work "
                      "scaled by %.0fx, plane work scaled to the job's %d planes; %.1f s wall" %
                      (stride, w_max_m, sub.shape[0], len(freq), n_sample, npix, plan["W"],
                       plan["sigma"], plan["nu"], planes_i, n_full / n_sample, plan["P"], wall),
            "split_s": {"vis_invert": tv_i, "vis_predict": tv_p, "planes_invert": tp_i,
                        "planes_predict": tp_p},
            "full_job_s": full_s,
            "sample_wall_s": wall}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one MI355X and this build "
              "has no multi-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    import torch

    from wsclean_amd import WGriddingGridder_Simple, _lib

    torch.cuda.set_device(0)
    if args.serve_only:
        print(json.dumps(serving_leg(args)))
        return
    if args.facets_only:
        print(json.dumps(facets_leg(args)))
        return
    dev = torch.device("cuda:0")
    t_gen = time.time()
    uvw_h, freq, dx, src = make_inputs(args.channels, args.npix)
    n_rows, n_chan = uvw_h.shape[0], len(freq)
    n_vis = n_rows * n_chan
    uvw = torch.from_numpy(uvw_h).to(dev)
    vis = torch.empty((n_rows, n_chan, 2), dtype=torch.float32, device=dev)
    out = torch.empty_like(vis)
    lib = _lib.load()
    srcc = np.ascontiguousarray(src)
    _lib.check(lib.wgx_synth_visibilities(
        ctypes_ptr(uvw), n_rows, n_chan, freq.ctypes.data_as(_c_void_p()), srcc.shape[0],
        srcc.ctypes.data_as(_c_void_p()), 0.05, 1002, ctypes_ptr(vis)))
    torch.cuda.synchronize()
    t_gen = time.time() - t_gen

    n = args.npix
    g = WGriddingGridder_Simple(n, n, n, n, dx, dx, 0.0, 0.0, 1, args.eps)
    g.set_timing(True)
    if args.tile or args.sg or args.W or args.sigma:
        g.set_tuning(W=args.W, sigma=args.sigma, tile=args.tile, sub_slab=args.sg)

    def step():
        g.PrepareVisibilities(uvw, freq, n_rows, n_chan)
        g.InitializeInversion()
        g.AddInversionPrepared(vis)
        g.FinalizeImage(1.0 / n_vis)
        g.PredictPrepared(out)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    g.stage_times(reset=True)
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = g.stage_times(reset=True)
    plan = g.plan()
    ms = (t1 - t0) * 1e3 / args.steps
    value = n_vis / (ms / 1e3)
    ctm = roofline_bytes(plan, n_vis, n_rows, n, n)
    stages = {k: v / args.steps for k, v in st.items()}
    line = {
        "metric": "vis/sec grid+degrid at eps=1e-5, 8192^2 w-gridding; p50/p99 request latency",
        "value": value,
        "unit": "vis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 coordinates and phases)",
        "data": "synthetic (seeded earth-rotation uvw, 505-source sky, device-generated)",
        "config": {"workload": "large w-gridding: %d^2 image, 2 arcsec, 128 antennas, 2880 x "
                               "10 s, %d channels 120-180 MHz, eps %.0e" % (n, n_chan, args.eps),
                   "n_vis": n_vis, "n_rows": n_rows, "n_chan": n_chan,
                   "plan": {k: plan[k] for k in ("W", "sigma", "nu", "nv", "P", "T", "Sg", "Sf",
                                                 "BU", "BV")}},
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3)},
        "stages_ms": stages,
        "ctm_roofline": {"bytes": ctm, "frac": ctm / (ms / 1e3) / HBM_PEAK},
        "roofline": kernel_roofline(st, plan),
        "cpu_baseline": None,
        "gen_s": t_gen,
    }
    if not args.no_accuracy:
        line["accuracy"] = accuracy_leg(g, uvw_h, vis, out, freq, dx, n, n_vis)
    if not args.no_graph:
        line["graph"] = {"8192": graph_cycle(g, vis, out, n_vis, max(2, args.steps))}
    # the other legs make handles of their own: free this one's HBM first
    g.close()
    torch.cuda.empty_cache()
    line["correction"] = fresh_handle_step(uvw, vis, out, freq, n, dx, args.eps, n_rows, n_chan,
                                           ms)
    if not args.no_autotune:
        line["autotune"] = autotune_leg(uvw, freq, n, dx, args.eps, n_rows, n_chan, plan)
    if not args.no_per_channel:
        line["per_channel_gpu"] = per_channel_leg(uvw, vis, freq, n, dx, args.eps, n_rows, value)
    if not args.no_host_leg:
        line["host_buffers"] = host_buffer_leg(uvw, vis, freq, n, dx, args.eps, n_rows, n_chan)
    del uvw, vis, out
    torch.cuda.empty_cache()
    if not args.no_graph:
        line["graph"]["512"] = small_graph_cycle(args.eps)
    if args.serve_requests > 0:
        line["serving"] = serving_leg(args)
    if args.facet_steps > 0:
        line["facets"] = facets_leg(args)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(uvw_h, freq, dx, src, n, plan, args.cpu_stride)
    print(json.dumps(line))


def accuracy_leg(g, uvw_h, vis, out, freq, dx, n, n_vis, n_pix=16, n_sel=10000, seed=8192):
    """Error of the timed pipeline's own results against the FP64 direct FT
    (oracle, the DirectMSGridder restatement), relative L2 as epsilon is
    defined (doc/source/wgridding.rst:19-20): the last step's dirty image
    (x n_vis) on n_pix random pixels over all visibilities, and a predict of
    a 500-source pixel model on n_sel random visibilities."""
    import torch
    from oracle import oracle
    from wsclean_amd import synth
    t = time.time()
    rng = np.random.Generator(np.random.PCG64(seed))
    img = g.RealImage().reshape(n, n).astype(np.float64) * n_vis
    xs, ys = rng.integers(0, n, n_pix), rng.integers(0, n, n_pix)
    vis_h = vis.cpu().numpy().view(np.complex64).reshape(uvw_h.shape[0], len(freq))
    ref = oracle.dft_invert_pixels_fast(n, n, dx, dx, 0.0, 0.0, uvw_h, freq, vis_h, xs, ys)
    del vis_h
    e_inv = float(np.linalg.norm(img[ys, xs] - ref) / np.linalg.norm(ref))
    model = np.zeros((n, n))
    _, _, flux, mx, my = synth.point_sky(rng, 500, n, dx, on_pixel=True)
    model[my, mx] = flux
    g.InitializePrediction(model.astype(np.float32).ravel())
    g.PredictPrepared(out)
    sel = rng.choice(out.shape[0] * out.shape[1], n_sel, replace=False)
    got = out.view(-1, 2)[torch.from_numpy(sel).to(out.device)].cpu().numpy()
    refv = oracle.dft_predict(n, n, dx, dx, 0.0, 0.0, model, uvw_h, freq, sel=sel)
    e_pred = float(np.linalg.norm(got[:, 0] + 1j * got[:, 1] - refv) / np.linalg.norm(refv))
    return {"invert_rel_l2": e_inv, "invert_pixels": n_pix, "predict_rel_l2": e_pred,
            "predict_visibilities": n_sel, "reference": "FP64 direct FT", "s": time.time() - t}


def fresh_handle_step(uvw, vis, out, freq, n, dx, eps, n_rows, n_chan, ms_cached):
    """One step on a new handle, as WSClean creates a new gridder for every
    inversion and prediction (wgridder/wgriddingmsgridder.cpp:267, :316):
    it pays the grid-correction table (cached per handle in the timed
    steps), the FFT twiddle tables and the buffer allocations."""
    import torch
    from wsclean_amd import WGriddingGridder_Simple
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = WGriddingGridder_Simple(n, n, n, n, dx, dx, 0.0, 0.0, 1, eps)
    g.PrepareVisibilities(uvw, freq, n_rows, n_chan)
    g.InitializeInversion()
    g.AddInversionPrepared(vis)
    g.FinalizeImage(1.0 / (n_rows * n_chan))
    g.PredictPrepared(out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    g.close()
    torch.cuda.empty_cache()
    return {"cached_per_handle": True, "timed_step_ms": ms_cached, "fresh_handle_step_ms": ms,
            "note": "the timed steps reuse one handle, which keeps the grid-correction table; "
                    "a new handle per step also builds it and allocates its buffers"}


def host_buffer_leg(uvw, vis, freq, n, dx, eps, n_rows, n_chan):
    """The reference-shaped call sequence on host buffers: one
    AddInversionData and one PredictVisibilities with all channels of the
    workload in one chunk (wgridder/wgriddingmsgridder.cpp:176-178), uvw and
    visibilities in host memory, so the time includes the PCIe copies both
    ways and the per-call plan and sort.  One pass, wall clock."""
    import torch

    from wsclean_amd import WGriddingGridder_Simple
    uvw_h = uvw.cpu().numpy()
    vis_h = vis.cpu().numpy().view(np.complex64).reshape(n_rows, n_chan)
    out_h = np.empty_like(vis_h)
    g = WGriddingGridder_Simple(n, n, n, n, dx, dx, 0.0, 0.0, 1, eps)

    def once():
        g.InitializeInversion()
        g.AddInversionData(n_rows, n_chan, uvw_h, freq, vis_h)
        g.FinalizeImage(1.0 / (n_rows * n_chan))
        g.PredictVisibilities(n_rows, n_chan, uvw_h, freq, out_h)

    once()   # buffers, twiddles, kernels
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    once()
    dt = time.perf_counter() - t0
    g.close()
    del vis_h, out_h, uvw_h
    return {"value": n_rows * n_chan / dt, "unit": "vis/s", "ms": dt * 1e3,
            "bytes_over_pcie": 2 * (vis.numel() * 4 + uvw.numel() * 8),
            "calls": "AddInversionData + PredictVisibilities, all channels, host buffers"}


def per_channel_leg(uvw, vis, freq, n, dx, eps, n_rows, batched_value):
    """The same workload as a plain one-call-per-channel GPU gridder: for
    every channel a separate plan + sort, invert and predict (the channel's
    visibilities copied to contiguous device buffers before timing).  One
    pass, timed with the device synchronised on both sides."""
    import torch

    from wsclean_amd import WGriddingGridder_Simple
    n_chan = len(freq)
    vis_t = vis.permute(1, 0, 2).contiguous()     # [chan][rows][2]
    out_t = torch.empty_like(vis_t)
    g = WGriddingGridder_Simple(n, n, n, n, dx, dx, 0.0, 0.0, 1, eps)

    def one(c):
        f = freq[c:c + 1]
        g.PrepareVisibilities(uvw, f, n_rows, 1)
        g.AddInversionPrepared(vis_t[c])
        g.PredictPrepared(out_t[c])

    g.InitializeInversion()
    one(0)   # warm the buffers
    torch.cuda.synchronize()
    g.InitializeInversion()
    t0 = time.perf_counter()
    for c in range(n_chan):
        one(c)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    v = n_rows * n_chan / dt
    del vis_t, out_t
    torch.cuda.empty_cache()
    return {"value": v, "unit": "vis/s", "ms": dt * 1e3, "calls": n_chan,
            "batched_over_per_channel": batched_value / v}


def autotune_leg(uvw, freq, n, dx, eps, n_rows, n_chan, plan):
    """The autotuner on the headline workload (a 1.3e8-visibility row
    sample): every trial's extrapolated step time, the choice, and the
    planner's default (W, sigma, T) beside it."""
    from wsclean_amd import WGriddingGridder_Simple
    g = WGriddingGridder_Simple(n, n, n, n, dx, dx, 0.0, 0.0, 1, eps)
    t = time.time()
    trials = g.Autotune(uvw, freq, n_rows, n_chan, max_sample_vis=1 << 27)
    tune_s = time.time() - t
    g.close()
    return {"tune_s": tune_s, "sample_vis": 1 << 27,
            "chosen": {k: trials[0][k] for k in ("W", "sigma", "T", "est_ms")},
            "default": {"W": plan["W"], "sigma": plan["sigma"], "T": plan["T"]},
            "trials": [{k: t[k] for k in ("W", "sigma", "T", "P", "nu", "est_ms", "ok")}
                       for t in trials]}


def graph_cycle(g, vis, out, n_vis, steps):
    """One major cycle on the prepared set (clear, invert, finalize,
    predict) issued eagerly versus replayed from a captured HIP graph."""
    import torch

    def eager():
        g.InitializeInversion()
        g.AddInversionPrepared(vis)
        g.FinalizeImage(1.0 / n_vis)
        g.PredictPrepared(out)

    eager()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eager()
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / steps
    g.CaptureMajorCycle(vis, out, 1.0 / n_vis)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.ReplayMajorCycle(steps)
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t0) * 1e3 / steps
    return {"eager_ms": eager_ms, "graph_ms": graph_ms, "cycles": steps}


def small_graph_cycle(eps):
    """The graph leg on a small request: 512^2 image, ~2e6 visibilities."""
    from wsclean_amd import WGriddingGridder_Simple
    from wsclean_amd.serve import make_requests
    r = make_requests(1, 1.0, seed=7, npix_choices=[512], vis_range=(2e6, 2.0001e6), cycles=1,
                      eps=eps)[0]
    g = WGriddingGridder_Simple(512, 512, 512, 512, r.pixel_rad, r.pixel_rad, 0.0, 0.0, 1, eps)
    g.PrepareVisibilities(r.uvw, r.freq, r.n_rows, r.n_chan)
    res = graph_cycle(g, r.vis, r.out, r.n_vis, 20)
    res["n_vis"] = r.n_vis
    g.close()
    return res


def serving_leg(args):
    """Served request stream: seeded Poisson arrivals of requests with
    image sizes 512^2 ... 8192^2 and 1e5 ... 5e7 visibilities from random
    32-256 antenna arrays over 1-64 channels, each 5 major cycles (invert,
    image to host, model, predict, residual) on device-resident sorted
    visibilities, admitted by free slot and HBM.  Latency is arrival to
    completion; throughput counts one grid + degrid per visibility and
    cycle over the stream's span."""
    import torch

    from wsclean_amd.serve import (HandlePool, NativeServer, Server, latency_stats,
                                   make_requests, run_request)
    # one untimed request loads the kernels' code objects
    warm = make_requests(1, 1.0, seed=1, npix_choices=[512], vis_range=(1e5, 1.1e5), cycles=1,
                         eps=args.eps)
    run_request(warm[0])
    del warm
    t = time.time()
    reqs = make_requests(args.serve_requests, args.serve_rate, seed=1005, cycles=5, eps=args.eps)
    gen_s = time.time() - t
    # the library's request loop (C++ workers, per-geometry handle pool that
    # starts empty, admission by measured free HBM)
    nsrv = NativeServer(slots=args.serve_slots)
    out = nsrv.serve(reqs)
    s = latency_stats(out)
    per_request = [{"npix": r.npix, "n_vis": r.n_vis, "n_chan": r.n_chan,
                    "arrival_s": round(r.arrival, 4),
                    "wait_ms": round((r.start - r.arrival) * 1e3, 2),
                    "latency_ms": round((r.done - r.arrival) * 1e3, 2)} for r in out]
    native_stats = nsrv.stats
    for r in reqs:
        r.start = r.done = 0.0
    # the same stream through the Python loop (threads over the C-ABI calls)
    pool = HandlePool()
    srv = Server(slots=args.serve_slots, pool=pool)
    out = srv.serve(reqs)
    pool.close()
    sp = latency_stats(out)
    sp.update({"loop": "Python threads (wsclean_amd.serve.Server)", "slots": srv.slots,
               "peak_running": srv.peak_running,
               "min_free_gb_at_admission": min(f for _, f, _ in srv.admissions) / 1e9,
               "handle_pool": {"hits": pool.hits, "misses": pool.misses,
                               "evictions": pool.evictions}})
    s.update({"workload": "served stream: %d requests, Poisson %.1f/s, 512^2-8192^2, "
                          "1e5-5e7 vis, 5 major cycles, eps %.0e" % (len(reqs), args.serve_rate,
                                                                     args.eps),
              "loop": "library (wg_server_*: C++ worker threads)", "slots": nsrv.slots,
              "peak_running": native_stats["peak_running"], "gen_s": gen_s,
              "admission": "measured free HBM at each admission, less the running "
                           "requests' outstanding bytes, 1 GiB reserve",
              "min_free_gb_at_admission": native_stats["min_free_bytes_at_admission"] / 1e9,
              "handle_pool": {"hits": native_stats["pool_hits"],
                              "misses": native_stats["pool_misses"],
                              "closed": native_stats["handles_closed"]},
              "python_loop": sp,
              "per_request": per_request})
    del reqs, out
    torch.cuda.empty_cache()
    return s


def make_facet_inputs(share=8, seed=1004):
    from wsclean_amd import synth
    return synth.batched_facets_workload(share, seed)


def facets_leg(args):
    """Facet batch (all 36 facets in one sort and one gridding / degridding
    launch per sub-slab) against one gridder per facet on the same rows.
    A step is prepare (plan + sort) + invert + finalize + predict of every
    facet; value counts every facet's visibilities once."""
    import torch

    from wsclean_amd import WGriddingGridder_Simple, _lib
    from wsclean_amd.facets import FacetBatch
    dev = torch.device("cuda:0")
    t = time.time()
    uvw_h, freq, facets, row_off, n_rows, src = make_facet_inputs()
    n_chan = len(freq)
    uvw = torch.from_numpy(uvw_h).to(dev)
    R = uvw_h.shape[0]
    vis = torch.empty((R, n_chan, 2), dtype=torch.float32, device=dev)
    lib = _lib.load()
    srcc = np.ascontiguousarray(src)
    _lib.check(lib.wgx_synth_visibilities(
        ctypes_ptr(uvw), R, n_chan, freq.ctypes.data_as(_c_void_p()), srcc.shape[0],
        srcc.ctypes.data_as(_c_void_p()), 0.05, 1004, ctypes_ptr(vis)))
    torch.cuda.synchronize()
    gen_s = time.time() - t
    n_virtual = int(sum(n_rows)) * n_chan
    out = torch.empty((n_virtual, 2), dtype=torch.float32, device=dev)
    fb = FacetBatch(facets)

    def step():
        fb.Prepare(uvw, freq, row_off, n_rows)
        fb.InitializeInversion()
        fb.AddInversion(vis)
        for f in range(len(facets)):
            fb.FinalizeImage(f, 1.0 / max(1, n_rows[f] * n_chan))
        fb.Predict(out)

    fb.set_timing(True)
    step()
    torch.cuda.synchronize()
    fb.stage_times(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.facet_steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.facet_steps
    st = fb.stage_times(reset=True)
    plan = fb.plan()
    fb.close()
    del out
    torch.cuda.empty_cache()
    # one gridder per facet (WSClean's one-gridder-per-facet structure)
    gs = []
    for f, fac in enumerate(facets):
        g = WGriddingGridder_Simple(fac["width"], fac["height"], fac["width"], fac["height"],
                                    fac["pixel_size_x"], fac["pixel_size_y"], fac["l_shift"],
                                    fac["m_shift"], 1, fac["epsilon"])
        gs.append(g)
    sub = [(uvw[row_off[f]:row_off[f] + n_rows[f]], vis[row_off[f]:row_off[f] + n_rows[f]])
           for f in range(len(facets))]
    outs = [torch.empty_like(v) for _, v in sub]

    def per_facet():
        for f, g in enumerate(gs):
            u, v = sub[f]
            g.PrepareVisibilities(u, freq, n_rows[f], n_chan)
            g.InitializeInversion()
            g.AddInversionPrepared(v)
            g.FinalizeImage(1.0 / max(1, n_rows[f] * n_chan))
            g.PredictPrepared(outs[f])

    per_facet()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per_facet()
    torch.cuda.synchronize()
    ms_pf = (time.perf_counter() - t0) * 1e3
    for g in gs:
        g.close()
    del sub, outs, uvw, vis
    torch.cuda.empty_cache()
    a = plan["atlas"]
    return {"workload": "batched facets: 12000^2 field at 2 arcsec, 36 facets 1024-3072 px, "
                        "log-uniform 1e6-5e7 rows each from the 128-antenna 10 km track at "
                        "16 channels, one GPU's 1/8 row share, eps 1e-5",
            "value": n_virtual / (ms / 1e3), "unit": "vis/s", "ms_per_step": ms,
            "steps": args.facet_steps, "n_facets": len(facets), "n_vis": n_virtual,
            "per_facet_gridders_ms": ms_pf,
            "batch_over_per_facet": ms_pf / ms,
            "atlas": {k: a[k] for k in ("W", "sigma", "T", "Sg", "Sf", "ntu", "ntv", "P", "P0")},
            "stages_ms": {k: v / args.facet_steps for k, v in st.items()},
            "n_items": plan["n_items"], "gen_s": gen_s}


def ctypes_ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def _c_void_p():
    import ctypes
    return ctypes.c_void_p


if __name__ == "__main__":
    main()
