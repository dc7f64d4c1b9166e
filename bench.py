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
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--serve-requests", type=int, default=40,
                    help="requests in the served-stream leg (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=16.0, help="Poisson arrivals per second")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the served-stream leg and print its JSON")
    ap.add_argument("--no-autotune", action="store_true", help="skip the autotuner leg")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph leg")
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


def kernel_roofline(st, plan):
    """Roofline of the dominant kernel, the ES-kernel gridding kernel, from
    its own HIP events on the engine stream.

    For W <= 8 the gridding runs on the FP32 matrix cores
    (GridKernelMfma<W, fused>): bound "mfma", algorithmic FLOPs = 4 W^3 per
    visibility pass (each of the W^3 footprint cells takes a complex x real
    multiply-add), against the 157.3 TFLOP/s dense FP32 MFMA peak.  With one
    plane sub-slab (P <= Sg) the kernel gathers the visibilities itself
    (fused).  The HBM view is reported beside it: algorithmic bytes per
    launch = per visibility pass 16 B coordinates + 8 B visibility (fused:
    + 4 B permutation word + 8 B phasor), read once per sub-slab the
    visibility touches, + 8 B x sub_n x E^2 per work item (its subgrid
    written once), against 8 TB/s.  traffic = measured HBM
    bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by the
    copy-kernel calibration) from the committed counter summary of this
    kernel and plan, else None."""
    n = st["grid_launches"]
    if not n or st["grid_kernel_ms"] <= 0:
        return None
    W = plan["W"]
    t = st["grid_kernel_ms"] / n / 1e3
    per_launch_bytes = st["grid_bytes"] / n
    per_launch_flops = 4.0 * W ** 3 * st["grid_passes"] / n
    gbs = per_launch_bytes / t / 1e9
    mfma = W <= 8
    fused = mfma and plan["P"] <= plan["Sg"]
    kernel = ("GridKernelMfma<%d, %s>" % (W, "true" if fused else "false")) if mfma else \
        "GridKernel<%d>" % W
    if mfma:
        achieved = per_launch_flops / t / 1e12
        out = {"bound": "mfma", "kernel": kernel, "achieved": achieved,
               "peak": MFMA_F32_PEAK / 1e12, "unit": "TFLOP/s",
               "frac": achieved / (MFMA_F32_PEAK / 1e12)}
    else:
        out = {"bound": "hbm", "kernel": kernel, "achieved": gbs, "peak": HBM_PEAK / 1e9,
               "unit": "GB/s", "frac": gbs / (HBM_PEAK / 1e9)}
    dt = st["degrid_kernel_ms"] / max(st["degrid_launches"], 1) / 1e3
    out.update({
        "traffic": None, "launch_ms": t * 1e3, "launches": int(n),
        "flops_per_launch": per_launch_flops, "bytes_per_launch": per_launch_bytes,
        "hbm_view": {"achieved": gbs, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": gbs / (HBM_PEAK / 1e9)},
        "degrid": {"kernel": ("DegridKernelMfma<%d, %s>" % (W, "true" if fused else "false"))
                   if mfma else "DegridKernel<%d>" % W,
                   "launch_ms": dt * 1e3,
                   "achieved_tflops": 4.0 * W ** 3 * st["degrid_passes"]
                   / max(st["degrid_launches"], 1) / max(dt, 1e-12) / 1e12}})
    path = os.path.join(ROOT, "profiles", "gridkernel_8192_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            tr = json.load(f)
        if tr.get("kernel") == kernel and \
                tr.get("plan") == {k: plan[k] for k in tr.get("plan", {})}:
            out["traffic"] = tr["bytes_per_dispatch"]
    return out


def cpu_baseline(uvw_h, freq, dx, src, npix, plan, max_seconds):
    """FP64 CPU w-gridder (oracle/wgridder_cpu.cpp, 'CPU restatement, not
    ducc0') on a bounded sample of the same workload: the full image and the
    same kernel (W, beta, nu, nv), one channel (the middle one) and the
    central part of the track, sized for ~max_seconds on the host cores.
    Returns invert + predict visibilities per second of the sample."""
    from oracle import oracle
    from wsclean_amd import synth
    cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "64") or 64))
    n_bl = 8128
    n_t = uvw_h.shape[0] // n_bl
    # central 1/4 of the track of one channel
    t0, t1 = n_t * 3 // 8, n_t * 5 // 8
    sub = uvw_h[t0 * n_bl:t1 * n_bl]
    f = np.array([freq[len(freq) // 2]])
    l, m, S = src[:64, 0], src[:64, 1], src[:64, 3]
    vis = synth.predict_sources(sub[:200000], f, l, m, S).astype(np.complex64)
    vis = np.resize(vis, (sub.shape[0], 1)).astype(np.complex64)
    kw = dict(W=plan["W"], beta=plan["beta"], sigma=plan["sigma"], nu=plan["nu"],
              nv=plan["nv"], nthreads=cores)
    t = time.perf_counter()
    img = oracle.cpu_wgrid_invert(npix, npix, dx, dx, 0.0, 0.0, sub, f, vis, **kw)
    oracle.cpu_wgrid_predict(npix, npix, dx, dx, 0.0, 0.0, img, sub, f, **kw)
    dt = time.perf_counter() - t
    n = sub.shape[0]
    return {"value": n / dt, "unit": "vis/s", "cores": cores, "kind": "port",
            "sample": "CPU restatement (FP64, not ducc0): invert+predict of %d visibilities "
                      "(1 channel at %.1f MHz, central quarter of the track), %d^2 image, "
                      "same kernel W=%d sigma=%.2f; %.1f s" % (n, f[0] / 1e6, npix, plan["W"],
                                                              plan["sigma"], dt)}


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
    if not args.no_graph:
        line["graph"] = {"8192": graph_cycle(g, vis, out, n_vis, max(2, args.steps))}
    # the other legs make handles of their own: free this one's HBM first
    g.close()
    torch.cuda.empty_cache()
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
        line["cpu_baseline"] = cpu_baseline(uvw_h, freq, dx, src, n, plan, args.cpu_seconds)
    print(json.dumps(line))


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

    from wsclean_amd.serve import Server, latency_stats, make_requests, run_request
    # one untimed request loads the kernels' code objects
    warm = make_requests(1, 1.0, seed=1, npix_choices=[512], vis_range=(1e5, 1.1e5), cycles=1,
                         eps=args.eps)
    run_request(warm[0])
    del warm
    t = time.time()
    reqs = make_requests(args.serve_requests, args.serve_rate, seed=1005, cycles=5, eps=args.eps)
    gen_s = time.time() - t
    srv = Server(slots=args.serve_slots)
    out = srv.serve(reqs)
    s = latency_stats(out)
    s.update({"workload": "served stream: %d requests, Poisson %.1f/s, 512^2-8192^2, "
                          "1e5-5e7 vis, 5 major cycles, eps %.0e" % (len(reqs), args.serve_rate,
                                                                     args.eps),
              "slots": srv.slots, "peak_running": srv.peak_running,
              "peak_admitted_gb": srv.peak_bytes / 1e9, "gen_s": gen_s,
              "per_request": [{"npix": r.npix, "n_vis": r.n_vis, "n_chan": r.n_chan,
                               "arrival_s": round(r.arrival, 4),
                               "wait_ms": round((r.start - r.arrival) * 1e3, 2),
                               "latency_ms": round((r.done - r.arrival) * 1e3, 2)}
                              for r in out]})
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
