"""Facet-batch layout on the host (no GPU): every facet keeps the plan a
single gridder would make with the batch's shared kernel shape, and the
facets' tile regions are disjoint inside the atlas, with each box's W - 1
overhang inside its own region."""
import numpy as np
import pytest

from wsclean_amd import _lib, synth
from wsclean_amd.facets import facet_atlas

ARCSEC = synth.ARCSEC


def facet_case(seed, n_facets):
    rng = np.random.default_rng(seed)
    enu = synth.antennas_disc(rng, 24, 1500.0)
    uvw = synth.earth_rotation_uvw(enu, 40, 60.0, np.deg2rad(-30.0))
    freq = np.array([120e6, 160e6])
    facets, exts, nvis = [], [], []
    for f in range(n_facets):
        n = int(rng.choice([48, 64, 96, 128, 160]))
        m = int(rng.choice([48, 64, 96, 128]))
        facets.append(dict(width=n, height=m, width_t=n, height_t=m,
                           pixel_size_x=10 * ARCSEC, pixel_size_y=10 * ARCSEC,
                           l_shift=float(rng.uniform(-0.02, 0.02)),
                           m_shift=float(rng.uniform(-0.02, 0.02)), epsilon=1e-5))
        r0 = int(rng.integers(0, uvw.shape[0] // 2))
        r1 = int(rng.integers(r0 + 1, uvw.shape[0]))
        u = uvw[r0:r1].copy()
        u[u[:, 2] < 0] *= -1
        s = freq / 299792458.0
        ext = []
        for k in range(3):
            c = np.concatenate([u[:, k] * s.min(), u[:, k] * s.max()])
            ext += [c.min(), c.max()]
        exts.append(ext)
        nvis.append((r1 - r0) * len(freq))
    return facets, np.array(exts), np.array(nvis)


def single_plan(f, ext, nvis, W, sigma, T, Sg):
    import ctypes
    import json
    lib = _lib.load()
    p = _lib.WgParams(f["width"], f["height"], f["width_t"], f["height_t"], f["pixel_size_x"],
                      f["pixel_size_y"], f["l_shift"], f["m_shift"], 1, f["epsilon"], 0, 0)
    e = np.ascontiguousarray(ext, np.float64)
    buf = ctypes.create_string_buffer(4096)
    _lib.check(lib.wgx_make_plan(ctypes.byref(p), e.ctypes.data_as(ctypes.c_void_p), int(nvis),
                                 W, sigma, T, Sg, buf, 4096))
    return json.loads(buf.value.decode())


@pytest.mark.parametrize("seed,n_facets", [(1, 1), (2, 5), (3, 12), (4, 36)])
def test_atlas_regions_disjoint_and_plans_shared(seed, n_facets):
    facets, exts, nvis = facet_case(seed, n_facets)
    lay = facet_atlas(facets, exts, nvis)
    a = lay["atlas"]
    T, W = a["T"], a["W"]
    gap = (W - 1 + T - 1) // T
    occupied = np.zeros((a["ntv"], a["ntu"]), np.int32)
    for f, fl in enumerate(lay["facets"]):
        p = fl["plan"]
        assert (p["W"], p["T"], p["Sg"], p["beta"], p["sigma"]) == \
            (W, T, a["Sg"], a["beta"], a["sigma"])
        assert p["P"] <= a["P"] and p["P0"] <= a["P0"]
        # the box and its overhang fit the facet's region
        assert p["BU"] <= (p["ntu"] + gap) * T and p["BV"] <= (p["ntv"] + gap) * T
        tu0, tv0 = fl["tu0"], fl["tv0"]
        assert tu0 + p["ntu"] + gap <= a["ntu"] and tv0 + p["ntv"] + gap <= a["ntv"]
        occupied[tv0:tv0 + p["ntv"] + gap, tu0:tu0 + p["ntu"] + gap] += 1
        # the facet's plan is the single gridder's plan with the shared shape
        ref = single_plan(facets[f], exts[f], nvis[f], W, p["sigma"], T, a["Sg"])
        for k in ("nu", "nv", "P", "P0", "ntu", "ntv", "BU", "BV", "box_u0", "box_v0",
                  "w_lo", "dw", "zc"):
            assert p[k] == ref[k], k
    assert occupied.max() == 1
    assert a["BU"] == a["ntu"] * T + W - 1 and a["BV"] == a["ntv"] * T + W - 1
    assert a["nbuckets"] == a["ntu"] * a["ntv"] * a["P0"]
    assert a["nkeys"] == a["nbuckets"] * a["key_T"] ** 2 < 2 ** 32
    # shelf packing wastes at most a modest share of the atlas
    used = sum((fl["plan"]["ntu"] + gap) * (fl["plan"]["ntv"] + gap) for fl in lay["facets"])
    assert used >= 0.5 * a["ntu"] * a["ntv"]
