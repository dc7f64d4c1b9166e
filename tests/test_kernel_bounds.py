"""CPU check of the index ranges of every HIP kernel of the pipeline.

The plan comes from the library itself (wgx_make_plan, host-only); the
bucket keys, the stable sort, the bucket offsets and the work items are
rebuilt in numpy exactly as the device builds them (the key restatement is
pinned bit-for-bit against the device by the GPU sort tests).  Every array
index the kernels form is then checked against the size of the buffer it
addresses, for the geometries the GPU tests and the benchmark use.  No GPU.
"""
import ctypes
import json

import numpy as np
import pytest

from oracle import oracle
from wsclean_amd import _lib, synth

CHUNK = 65536      # wsclean_amd/csrc/grid.hip kChunk
MAX_LDS_FFT = 16384   # largest LDS FFT line
MAX_FFT = 5 * MAX_LDS_FFT  # longest line: R <= 5 subsequences of an LDS line


def split_radix(n):
    """Subsequences of a line of n points (fft.hip LongFFTRadix): 1 up to
    the LDS FFT, else the smallest R in 2..5 with n / R an LDS line."""
    if n <= MAX_LDS_FFT:
        return 1
    for r in range(2, 6):
        if n % r == 0 and n // r <= MAX_LDS_FFT:
            return r
    return 0
C = synth.C_LIGHT


def extents(uvw, freq):
    s = freq / C
    smin, smax = s.min(), s.max()
    u, v, w = uvw[:, 0].copy(), uvw[:, 1].copy(), uvw[:, 2].copy()
    f = w < 0
    u[f], v[f], w[f] = -u[f], -v[f], -w[f]
    out = []
    for a in (u, v, w):
        x1, x2 = a * smin, a * smax
        out += [float(np.minimum(x1, x2).min()), float(np.maximum(x1, x2).max())]
    return np.array(out)


def make_plan(n, dx, eps, uvw, freq, ls=0.0, ms=0.0, **tune):
    lib = _lib.load()
    prm = _lib.WgParams(n, n, n, n, dx, dx, ls, ms, 1, eps, 0, 0)
    e = extents(uvw, freq)
    buf = ctypes.create_string_buffer(4096)
    _lib.check(lib.wgx_make_plan(ctypes.byref(prm), e.ctypes.data_as(ctypes.c_void_p),
                                 uvw.shape[0] * len(freq), tune.get("W", 0),
                                 tune.get("sigma", 0.0), tune.get("tile", 0),
                                 tune.get("sub_slab", 0), buf, 4096))
    p = json.loads(buf.value.decode())
    p["nx"] = p["ny"] = n
    return p


def check_plan_bounds(plan, uvw, freq, vis=None):
    W, T, P0, P = plan["W"], plan["T"], plan["P0"], plan["P"]
    ntu, ntv, BU, BV = plan["ntu"], plan["ntv"], plan["BU"], plan["BV"]
    Sg, Sf, nx, ny = plan["Sg"], plan["Sf"], plan["nx"], plan["ny"]
    E = T + W - 1
    ntiles = ntu * ntv
    nbuckets = ntiles * P0
    kt = plan["key_T"]
    assert kt == ((T - 1) >> plan["key_shift"]) + 1
    assert plan["nbuckets"] == nbuckets and plan["nkeys"] == nbuckets * kt * kt < 2 ** 32
    assert plan["nu"] <= MAX_FFT and plan["nv"] <= MAX_FFT
    for side in (plan["nu"], plan["nv"]):
        r = split_radix(side)
        assert r >= 1 and side % r == 0 and side // r <= MAX_LDS_FFT
    assert Sf % Sg == 0 and 1 <= Sg <= 32
    assert Sg * E * E * 8 <= 160 * 1024                      # LDS subgrid
    if W <= 8:
        # matrix-core gridding: 16-byte aligned subgrid + 8 waves x 2 views
        # x 2 x 64 staged float4 entries
        assert (Sg * E * E * 8 + 15) // 16 * 16 + 8 * 2 * 128 * 16 <= 160 * 1024
        # matrix-core degridding: box block + 8 waves x 128 x (float4 + float2)
        assert (Sg * E * E * 8 + 15) // 16 * 16 + 8 * 128 * 24 <= 160 * 1024
    assert BU == ntu * T + W - 1 and BV == ntv * T + W - 1

    keys, valid = oracle.bucket_keys(plan, uvw, freq, vis)
    kv = keys[valid].astype(np.int64)
    assert kv.size == 0 or (kv.min() >= 0 and kv.max() < plan["nkeys"])
    sk = np.sort(keys, kind="stable").astype(np.int64)
    n_valid = int(valid.sum())
    off = np.searchsorted(sk, np.arange(nbuckets + 1, dtype=np.int64) * kt * kt, side="left")
    assert off[-1] == n_valid

    # decode the sorted keys exactly as the kernels unpack them
    s = sk[:n_valid]
    # the visibilities' locations in sorted order (the keys hold the cell
    # only when key_shift is 0)
    order = np.argsort(keys, kind="stable")[:n_valid]
    loc_tile, loc_p0, loc_du, loc_dv = (a.ravel()[order] for a in
                                        oracle.bucket_locations(plan, uvw, freq))
    du, dv = loc_du, loc_dv
    bucket = s // (kt * kt)
    assert (bucket == loc_tile * P0 + loc_p0).all()
    if plan["key_shift"] == 0:
        assert (du == s % T).all() and (dv == (s // T) % T).all()
    tile, p0 = bucket // P0, bucket % P0
    assert (p0 < P0).all() and (tile < ntiles).all() and (p0 < 32768).all()

    nsub = (P + Sg - 1) // Sg
    counts = np.zeros((nsub, ntiles), np.int64)
    ranges = {}
    for g in range(nsub):
        lo, hi = max(0, g * Sg - W + 1), min(P0 - 1, g * Sg + Sg - 1)
        if lo > hi:
            continue
        t = np.arange(ntiles)
        b = off[t * P0 + lo]
        e = off[t * P0 + hi + 1]
        counts[g] = (e - b + CHUNK - 1) // CHUNK
        ranges[g] = (b, e)
    item_pos = np.concatenate([[0], np.cumsum(counts.ravel())])
    per_sub = np.array([item_pos[(g + 1) * ntiles] - item_pos[g * ntiles] for g in range(nsub)])
    max_items = int(per_sub.max()) if nsub else 0
    scratch_cells = max_items * Sg * E * E

    for g, (b, e) in ranges.items():
        sub_lo, sub_n = g * Sg, min(Sg, P - g * Sg)
        assert 1 <= sub_n <= Sg
        slab_lo = (sub_lo // Sf) * Sf
        assert sub_lo + sub_n <= slab_lo + Sf                 # sub-slab inside the slab
        nz = np.nonzero(e > b)[0]
        for t in nz:
            first = item_pos[g * ntiles + t] - item_pos[g * ntiles]
            nitem = item_pos[g * ntiles + t + 1] - item_pos[g * ntiles + t]
            assert first + nitem <= max_items
            seg = slice(b[t], e[t])
            # GridKernel: vis of the item belong to its tile and sub-slab range
            assert (tile[seg] == t).all()
            assert (p0[seg] <= sub_lo + sub_n - 1).all() and (p0[seg] + W > sub_lo).all()
            q_hi = np.minimum(p0[seg] + W, sub_lo + sub_n) - sub_lo
            lds_max = ((q_hi - 1) * E + dv[seg] + W - 1) * E + du[seg] + W - 1
            assert lds_max.max() < sub_n * E * E
            assert (first + nitem) * sub_n * E * E <= scratch_cells
            # DegridKernel / GatherBoxKernel: the tile footprint inside the box
            tu, tv = t % ntu, t // ntu
            assert tv * T + E <= BV and tu * T + E <= BU
            box_max = ((sub_lo + sub_n - 1 - slab_lo) * BV + tv * T + E - 1) * BU + tu * T + E - 1
            assert box_max < Sf * BU * BV

    # GatherBoxKernel: every cell's covering tiles and slot offsets
    r = np.arange(BV)
    c = np.arange(BU)
    tv0 = np.maximum(0, np.trunc((r - E + T) / T).astype(int))
    tv1 = np.minimum(ntv - 1, r // T)
    tu0 = np.maximum(0, np.trunc((c - E + T) / T).astype(int))
    tu1 = np.minimum(ntu - 1, c // T)
    assert (tv1 >= tv0).all() and (tu1 >= tu0).all()
    for lo_, hi_, x in ((tv0, tv1, r), (tu0, tu1, c)):
        for k in range(3):
            t = lo_ + k
            m = t <= hi_
            d = x[m] - t[m] * T
            assert (d >= 0).all() and (d < E).all()

    # FFT line kernels: folded loads and cropped stores stay inside the line
    for n, origin, L, crop, half in ((plan["nu"], plan["box_u0"], BU, nx, nx // 2),
                                     (plan["nv"], plan["box_v0"], BV, ny, ny // 2)):
        i = np.arange(L)
        pos = (origin + i) % n
        assert pos.min() >= 0 and pos.max() < n
        assert origin + L <= 3 * n
        assert crop <= n
    return dict(n_valid=n_valid, max_items=max_items, nsub=nsub)


PROBLEMS = [
    # (problem kwargs, npix, eps, shift)
    (dict(seed=1001, n_ant=16, n_times=24, n_chan=2, npix=96, noise=0.01), 96, 1e-5, (0, 0)),
    (dict(seed=3, n_ant=14, n_times=20, n_chan=2, npix=80, sky_lm=(0.05, -0.03)), 80, 1e-5,
     (0.05, -0.03)),
    (dict(seed=5, n_ant=20, n_times=40, n_chan=2, npix=512, pixel_arcsec=30.0, radius_m=6000.0,
          n_src=8, dec_deg=-70.0, dt_s=600.0), 512, 1e-4, (0, 0)),
    (dict(seed=9, n_ant=12, n_times=10, n_chan=2, npix=64, pixel_arcsec=60.0,
          radius_m=3000.0), 64, 1e-4, (0, 0)),
    (dict(seed=17, n_ant=40, n_times=60, n_chan=8, npix=512, pixel_arcsec=30.0,
          radius_m=6000.0, dec_deg=-70.0, dt_s=600.0), 512, 1e-5, (0, 0)),
    (dict(seed=8, n_ant=20, n_times=30, n_chan=2, npix=384, pixel_arcsec=30.0,
          radius_m=5000.0), 384, 1e-4, (-0.01, 0.02)),
    # dense core: full work items, runs thousands long
    (dict(seed=23, n_ant=40, n_times=240, n_chan=16, npix=256, pixel_arcsec=20.0,
          radius_m=300.0, n_src=12, dt_s=60.0), 256, 1e-5, (0, 0)),
]


@pytest.mark.parametrize("kw,n,eps,shift", PROBLEMS)
@pytest.mark.parametrize("tune", [{}, dict(tile=32, sub_slab=8), dict(sub_slab=8),
                                  dict(tile=8, sub_slab=32), dict(tile=12, sub_slab=16)])
def test_kernel_index_ranges(kw, n, eps, shift, tune):
    p = synth.small_problem(**kw)
    plan = make_plan(n, p["dx"], eps, p["uvw"], p["freq"], shift[0], shift[1], **tune)
    check_plan_bounds(plan, p["uvw"], p["freq"], p["vis"])
    check_plan_bounds(plan, p["uvw"], p["freq"], None)


@pytest.mark.parametrize("tune", [{}, dict(tile=32, sub_slab=8), dict(tile=8, sub_slab=32),
                                  dict(tile=12, sub_slab=16)])
def test_kernel_index_ranges_benchmark_geometry(tune):
    # the 8192^2, 2 arcsec, 128-antenna track of the benchmark; one channel
    # of every 8th time step (the index ranges depend on the coordinate
    # distribution, not on the visibility count)
    import bench
    uvw, freq, dx, _ = bench.make_inputs(64, 8192)
    nbl = 8128
    uvw = uvw.reshape(-1, nbl, 3)[::8].reshape(-1, 3)
    plan = make_plan(8192, dx, 1e-5, uvw, freq, **tune)
    f1 = freq[[0, 31, 63]]
    # the plan above covers the full band; check the extreme channels
    check_plan_bounds(plan, uvw, f1, None)


@pytest.mark.parametrize("seed,n_chan", [(3, 4), (11, 5)])
def test_bucket_perm_order(seed, n_chan):
    # the device's sorted order: a permutation of the valid visibilities that
    # sorts the keys; with rows already in the order of their middle-channel
    # key the row renumbering is the identity and the order reduces to a
    # stable sort of the keys over the caller's flat index
    p = synth.small_problem(seed=seed, n_ant=14, n_times=20, n_chan=n_chan, npix=128)
    uvw, freq = p["uvw"], p["freq"]
    vis = p["vis"].copy()
    vis[::5, 0] = 0
    plan = make_plan(128, p["dx"], 1e-4, uvw, freq)
    keys, valid = oracle.bucket_keys(plan, uvw, freq, vis)
    perm = oracle.bucket_perm(plan, uvw, freq, vis)
    assert np.array_equal(np.sort(perm), np.flatnonzero(valid).astype(np.uint32))
    assert (np.diff(keys[perm].astype(np.int64)) >= 0).all()

    mid = n_chan // 2
    rkeys, _ = oracle.bucket_keys(plan, uvw, freq[mid:mid + 1], None)
    order = np.argsort(rkeys, kind="stable")
    uvw_s, vis_s = uvw[order], vis[order]
    keys_s, valid_s = oracle.bucket_keys(plan, uvw_s, freq, vis_s)
    ref = np.argsort(keys_s, kind="stable")[: int(valid_s.sum())].astype(np.uint32)
    np.testing.assert_array_equal(oracle.bucket_perm(plan, uvw_s, freq, vis_s), ref)
    # the renumbering is not trivial for the unsorted rows
    assert not np.array_equal(order, np.arange(order.size))


_P96 = dict(seed=1001, n_ant=16, n_times=24, n_chan=2, npix=96, noise=0.01)
_P512 = dict(seed=5, n_ant=20, n_times=40, n_chan=2, npix=512, pixel_arcsec=30.0,
             radius_m=6000.0, n_src=8, dec_deg=-70.0, dt_s=600.0)
_P512B = dict(seed=17, n_ant=40, n_times=60, n_chan=8, npix=512, pixel_arcsec=30.0,
              radius_m=6000.0, dec_deg=-70.0, dt_s=600.0)


@pytest.mark.parametrize("kw,eps,tune,one_fft_slab", [
    # forced VALU shapes (W >= 10), several sub-slabs and FFT slabs
    (_P96, 1e-5, dict(W=10, sigma=1.5), False),
    (_P96, 1e-5, dict(W=16, sigma=1.8), False),
    (_P512, 1e-5, dict(W=10, sigma=1.5), False),
    (_P512, 1e-5, dict(W=12, sigma=1.6), False),
    (_P512, 1e-5, dict(W=14, sigma=1.7), False),
    (_P512, 1e-5, dict(W=16, sigma=1.8), False),
    (_P512, 1e-4, dict(sub_slab=4), False),
    (_P512B, 1e-5, dict(sub_slab=4), False),
    (_P512B, 1e-5, dict(sub_slab=4), True),
    (_P512, 1e-5, dict(W=10, sigma=1.5, sub_slab=4), False),
    (_P512, 1e-5, dict(W=12, sigma=1.6, sub_slab=3), True),
    (_P96, 2e-6, {}, False),
    (_P96, 1e-6, {}, False),
    (_P96, 5e-7, {}, False),
    (_P96, 3e-7, {}, False),
])
def test_kernel_index_ranges_wide_paths(kw, eps, tune, one_fft_slab):
    p = synth.small_problem(**kw)
    n = kw["npix"]
    plan = make_plan(n, p["dx"], eps, p["uvw"], p["freq"], **tune)
    if one_fft_slab:
        plan["Sf"] = plan["Sg"]   # the smallest FFT slab a byte budget can give
    check_plan_bounds(plan, p["uvw"], p["freq"], p["vis"])


def test_kernel_index_ranges_wide_field_default():
    # 4096^2 at 20 arcsec, 128 antennas in a 2.5 km disc (W 10, 2 sub-slabs)
    rng = np.random.Generator(np.random.PCG64(77))
    enu = synth.antennas_disc(rng, 128, 2500.0)
    uvw = synth.earth_rotation_uvw(enu, 60, 120.0, np.deg2rad(-30.0))
    freq = np.array([150e6])
    plan = make_plan(4096, 20.0 * synth.ARCSEC, 1e-5, uvw, freq)
    assert plan["W"] >= 10 and plan["P"] > plan["Sg"]
    check_plan_bounds(plan, uvw, freq, None)


def test_kernel_index_ranges_coarse_sort_keys():
    # 8192^2 at 6 arcsec on the benchmark's 128-antenna track: (tile, plane,
    # cell) keys would exceed 32 bits, so the in-bucket cell order is
    # coarsened (key_shift 1); one channel of every 32nd time step at both
    # band edges
    uvw, freq, _, _ = synth.large_wgridding_inputs(64, 8192)
    sub = uvw.reshape(-1, 8128, 3)[::32].reshape(-1, 3)
    plan = make_plan(8192, 6.0 * synth.ARCSEC, 1e-5, sub, freq)
    assert plan["key_shift"] >= 1 and plan["nkeys"] < 2 ** 32
    assert plan["nbuckets"] * plan["key_T"] ** 2 == plan["nkeys"]
    check_plan_bounds(plan, sub[::4], freq[[0, 63]], None)


@pytest.mark.parametrize("npix", [12288, 20480, 30720])
def test_plans_with_grid_sides_above_the_lds_fft(npix):
    # images whose padded grid at eps 1e-5 exceeds one LDS FFT line (16384)
    # plan with lines split into 2..5 subsequences; bench track, 2 arcsec,
    # every 64th time step of one channel at both band edges
    uvw, freq, _, _ = synth.large_wgridding_inputs(64, 8192)
    sub = uvw.reshape(-1, 8128, 3)[::64].reshape(-1, 3)
    plan = make_plan(npix, 2.0 * synth.ARCSEC, 1e-5, sub, freq[[0, 63]])
    assert plan["nu"] > MAX_LDS_FFT and plan["nv"] > MAX_LDS_FFT
    assert plan["nu"] >= plan["sigma"] * npix and plan["nv"] >= plan["sigma"] * npix
    check_plan_bounds(plan, sub[::8], freq[[0, 63]], None)
