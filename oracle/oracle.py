"""Oracle for the w-gridder: TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, as the checker.  Nothing in the product imports it.

Contents
--------
* ``dft_invert`` / ``dft_invert_pixels`` / ``dft_predict``: the FP64 direct
  Fourier transform of oracle/dft.c, a restatement of WSClean's
  DirectMSGridder (gridding/directmsgridder.cpp:56-132, :205-229) plus its
  adjoint (sign from gridding/msgridderdata.h:820-843, :929-941).
* ``dft_invert_reference_wrap``: a literal numpy restatement of
  DirectMSGridder::gridSample + FinishInversion (the scratch layout and the
  wrap, directmsgridder.cpp:56-89, :112-131) for tiny inputs.  It pins the
  closed-form pixel -> (l, m) mapping used by dft.c.
* ``bucket_keys``: numpy restatement of the device's FP64 visibility ->
  (tile, first plane) mapping (wsclean_amd/csrc/device.hpp LocateVis), the
  bucketing step that ducc0's ms2dirty/dirty2ms perform internally
  (wgridder/gridder_simple.h:217-225; ducc0 itself is absent, so its own
  bucketing is parity unpinned -- the bucket sort is pinned against a stable
  CPU sort of these keys).
* ``bucket_perm``: the sorted permutation as the device orders ties (rows
  renumbered in uv order first, then a stable sort of the keys).
* ``es_phi_hat``: Fourier transform of the ES kernel by adaptive scipy
  quadrature, to pin the planner's table and correction factors.
* ``cpu_wgridder_*``: the multithreaded C++ CPU w-gridder restatement
  (oracle/wgridder_cpu.cpp), used as the timed CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
C_LIGHT = 299792458.0

_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _nthreads(n):
    return int(n) if n else min(16, os.cpu_count() or 1)


def dft_invert(nx, ny, dx, dy, l_shift, m_shift, uvw, freq, vis, nthreads=0):
    """Dirty image [ny, nx] (index [y, x]), not divided by the image weight."""
    lib = load()
    uvw = np.ascontiguousarray(uvw, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    vis = np.ascontiguousarray(vis, np.complex64)
    out = np.zeros(nx * ny, np.float64)
    f = lib.oracle_dft_invert
    f.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + \
        [ctypes.c_size_t, ctypes.c_size_t] + [ctypes.c_void_p] * 4 + [ctypes.c_int]
    f(nx, ny, dx, dy, l_shift, m_shift, uvw.size // 3, freq.size, _p(uvw), _p(freq), _p(vis),
      _p(out), _nthreads(nthreads))
    return out.reshape(ny, nx)


def dft_invert_pixels(nx, ny, dx, dy, l_shift, m_shift, uvw, freq, vis, xs, ys, nthreads=0):
    lib = load()
    uvw = np.ascontiguousarray(uvw, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    vis = np.ascontiguousarray(vis, np.complex64)
    xs = np.ascontiguousarray(xs, np.int32)
    ys = np.ascontiguousarray(ys, np.int32)
    out = np.zeros(xs.size, np.float64)
    f = lib.oracle_dft_invert_pixels
    f.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + \
        [ctypes.c_size_t, ctypes.c_size_t] + [ctypes.c_void_p] * 3 + [ctypes.c_size_t] + \
        [ctypes.c_void_p] * 3 + [ctypes.c_int]
    f(nx, ny, dx, dy, l_shift, m_shift, uvw.size // 3, freq.size, _p(uvw), _p(freq), _p(vis),
      xs.size, _p(xs), _p(ys), _p(out), _nthreads(nthreads))
    return out


def dft_invert_pixels_fast(nx, ny, dx, dy, l_shift, m_shift, uvw, freq, vis, xs, ys,
                           nthreads=0, return_mode=False):
    """dft_invert_pixels for large visibility sets (threads over rows, the
    per-channel phasors of evenly spaced channels by recurrence).  Pinned
    against dft_invert_pixels by tests/test_oracle.py.  With return_mode,
    also returns whether the recurrence was used."""
    lib = load()
    uvw = np.ascontiguousarray(uvw, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    vis = np.ascontiguousarray(vis, np.complex64)
    xs = np.ascontiguousarray(xs, np.int32)
    ys = np.ascontiguousarray(ys, np.int32)
    out = np.zeros(xs.size, np.float64)
    used = ctypes.c_int(0)
    f = lib.oracle_dft_invert_pixels_fast
    f.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + \
        [ctypes.c_size_t, ctypes.c_size_t] + [ctypes.c_void_p] * 3 + [ctypes.c_size_t] + \
        [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p]
    f(nx, ny, dx, dy, l_shift, m_shift, uvw.size // 3, freq.size, _p(uvw), _p(freq), _p(vis),
      xs.size, _p(xs), _p(ys), _p(out), _nthreads(nthreads), ctypes.byref(used))
    return (out, bool(used.value)) if return_mode else out


def dft_invert_pixels_per_channel(nx, ny, dx, dy, l_shift, m_shift, uvw, freq, vis, xs, ys,
                                  nthreads=0, return_mode=False):
    """dft_invert_pixels_fast with one sum per channel: [len(xs), n_chan],
    column c = dft_invert_pixels of channel c alone.  Pinned against
    dft_invert_pixels per channel by tests/test_oracle.py."""
    lib = load()
    uvw = np.ascontiguousarray(uvw, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    vis = np.ascontiguousarray(vis, np.complex64)
    xs = np.ascontiguousarray(xs, np.int32)
    ys = np.ascontiguousarray(ys, np.int32)
    out = np.zeros(xs.size * freq.size, np.float64)
    used = ctypes.c_int(0)
    f = lib.oracle_dft_invert_pixels_per_channel
    f.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + \
        [ctypes.c_size_t, ctypes.c_size_t] + [ctypes.c_void_p] * 3 + [ctypes.c_size_t] + \
        [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p]
    f(nx, ny, dx, dy, l_shift, m_shift, uvw.size // 3, freq.size, _p(uvw), _p(freq), _p(vis),
      xs.size, _p(xs), _p(ys), _p(out), _nthreads(nthreads), ctypes.byref(used))
    out = out.reshape(xs.size, freq.size)
    return (out, bool(used.value)) if return_mode else out


def dft_predict(nx, ny, dx, dy, l_shift, m_shift, image, uvw, freq, sel=None, nthreads=0):
    """Visibilities from image [ny, nx]; complex128 [n_rows, n_chan] (or [len(sel)])."""
    lib = load()
    img = np.ascontiguousarray(image, np.float64).ravel()
    uvw = np.ascontiguousarray(uvw, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    n_rows, n_chan = uvw.size // 3, freq.size
    if sel is not None:
        sel = np.ascontiguousarray(sel, np.uint64)
        nsel = sel.size
        out = np.zeros(2 * nsel, np.float64)
        selp = _p(sel)
    else:
        nsel = 0
        out = np.zeros(2 * n_rows * n_chan, np.float64)
        selp = None
    f = lib.oracle_dft_predict
    f.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.c_void_p] + \
        [ctypes.c_size_t, ctypes.c_size_t] + [ctypes.c_void_p] * 2 + [ctypes.c_size_t] + \
        [ctypes.c_void_p] * 2 + [ctypes.c_int]
    f(nx, ny, dx, dy, l_shift, m_shift, _p(img), n_rows, n_chan, _p(uvw), _p(freq), nsel, selp,
      _p(out), _nthreads(nthreads))
    v = out[0::2] + 1j * out[1::2]
    return v if sel is not None else v.reshape(n_rows, n_chan)


def dft_invert_reference_wrap(width, height, dx, dy, l_shift, m_shift, uvw, freq, vis):
    """Literal restatement of DirectMSGridder for tiny inputs (pure numpy).

    gridSample (directmsgridder.cpp:112-131) accumulates into a scratch layer
    whose pixel (x, y) uses xSrc = (x + W/2) mod W, ySrc = (H - y + H/2) mod
    H, l = (W/2 - xSrc) dx + l_shift, m = (ySrc - H/2) dy + m_shift; then
    FinishInversion (:73-86) writes image[x + y W] = scratch[xSrc + ySrc W]
    with the same xSrc / ySrc.  Returns image [height, width].
    """
    uvw = np.asarray(uvw, np.float64).reshape(-1, 3)
    freq = np.asarray(freq, np.float64)
    vis = np.asarray(vis, np.complex64).reshape(uvw.shape[0], freq.size)
    scratch = np.zeros(width * height)
    for y in range(height):
        ysrc = (height - y) + height // 2
        if ysrc >= height:
            ysrc -= height
        m = (ysrc - height // 2) * dy + m_shift
        for x in range(width):
            xsrc = x + width // 2
            if xsrc >= width:
                xsrc -= width
            l = (width // 2 - xsrc) * dx + l_shift
            nm1 = np.sqrt(1.0 - l * l - m * m) - 1.0 if l * l + m * m < 1.0 else 0.0
            acc = 0.0
            for r in range(uvw.shape[0]):
                for c in range(freq.size):
                    s = freq[c] / C_LIGHT
                    u, v, w = uvw[r] * s
                    ang = -2.0 * np.pi * (u * l + v * m + w * nm1)
                    val = vis[r, c]
                    acc += float(val.real) * np.cos(ang) - float(val.imag) * np.sin(ang)
            scratch[y * width + x] = acc
    image = np.zeros(width * height)
    for y in range(height):
        ysrc = (height - y) + height // 2
        if ysrc >= height:
            ysrc -= height
        for x in range(width):
            xsrc = x + width // 2
            if xsrc >= width:
                xsrc -= width
            image[x + y * width] = scratch[xsrc + ysrc * width]
    return image.reshape(height, width)


def bucket_locations(plan, uvw, freq):
    """(tile, p0, du, dv) per flat visibility index [rows, chan], FP64 like
    LocateVis: the visibility's uv tile, first w plane and first kernel cell
    relative to the tile origin."""
    uvw = np.asarray(uvw, np.float64).reshape(-1, 3)
    freq = np.asarray(freq, np.float64)
    W, T = plan["W"], plan["T"]
    nu, nv = plan["nu"], plan["nv"]
    dx, dy = plan["dx"], plan["dy"]
    ufac = float(nu) * dx
    vfac = float(nv) * dy
    scale = freq / C_LIGHT
    flip = uvw[:, 2] < 0.0
    um = np.where(flip, -uvw[:, 0], uvw[:, 0])
    vm = np.where(flip, -uvw[:, 1], uvw[:, 1])
    wm = np.where(flip, -uvw[:, 2], uvw[:, 2])
    u = um[:, None] * scale[None, :]
    v = vm[:, None] * scale[None, :]
    w = wm[:, None] * scale[None, :]
    tu = u * ufac + 0.5 * nu
    tv = v * vfac + 0.5 * nv
    inv_dw = 1.0 / plan["dw"]
    tw = (w - plan["w_lo"]) * inv_dw + float(W // 2 - 1)
    hw = W // 2 - 1
    a0u = np.floor(tu).astype(np.int64) - hw
    a0v = np.floor(tv).astype(np.int64) - hw
    p0 = np.floor(tw).astype(np.int64) - hw
    if plan["wrap_u"]:
        lu = np.mod(a0u, nu)
    else:
        lu = a0u - plan["box_u0"]
    if plan["wrap_v"]:
        lv = np.mod(a0v, nv)
    else:
        lv = a0v - plan["box_v0"]
    tile = (lv // T) * plan["ntu"] + (lu // T)
    return tile, p0, lu % T, lv % T


def bucket_keys(plan, uvw, freq, vis=None):
    """(keys, valid) per flat visibility index, FP64 like LocateVis."""
    T = plan["T"]
    tile, p0, du, dv = bucket_locations(plan, uvw, freq)
    bucket = tile * plan["P0"] + p0
    # the in-bucket cell order, coarsened by key_shift when (bucket, cell)
    # keys would exceed 32 bits
    ks = plan.get("key_shift", 0)
    kt = ((T - 1) >> ks) + 1
    keys = ((bucket * kt + (dv >> ks)) * kt + (du >> ks)).astype(np.int64)
    valid = np.ones(keys.shape, bool)
    if vis is not None:
        vis = np.asarray(vis).reshape(keys.shape)
        valid = vis != 0
    nkeys = plan["ntu"] * plan["ntv"] * plan["P0"] * kt * kt
    keys = np.where(valid, keys, nkeys)
    return keys.ravel().astype(np.uint32), valid.ravel()


def bucket_perm(plan, uvw, freq, vis=None):
    """Sorted permutation of the valid visibilities (the caller's flat
    indices row * n_chan + chan) in the device's order: rows renumbered by
    a stable sort of their key at the middle channel, then a stable sort of
    the visibility keys in renumbered (row, channel) order."""
    uvw = np.asarray(uvw, np.float64).reshape(-1, 3)
    freq = np.asarray(freq, np.float64).ravel()
    n_rows, n_chan = uvw.shape[0], freq.size
    mid = n_chan // 2
    rkeys, _ = bucket_keys(plan, uvw, freq[mid:mid + 1], None)
    rperm = np.argsort(rkeys, kind="stable")
    keys, valid = bucket_keys(plan, uvw, freq, vis)
    keys = keys.reshape(n_rows, n_chan)[rperm].ravel()
    order = np.argsort(keys, kind="stable")[: int(valid.sum())]
    r, ch = order // n_chan, order % n_chan
    return (rperm[r] * n_chan + ch).astype(np.uint32)


def _good_size(n):
    m = n + (n & 1)
    while True:
        r = m
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        if r == 1:
            return m
        m += 2


def cpu_wgrid_params(nx, ny, W, sigma, beta):
    return dict(W=W, beta=beta, sigma=sigma, nu=_good_size(int(np.ceil(sigma * nx))),
                nv=_good_size(int(np.ceil(sigma * ny))))


def _wgrid_fn(name, out_type):
    lib = load()
    f = getattr(lib, name)
    f.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.c_int] + \
        [ctypes.c_double] * 2 + [ctypes.c_int] * 2 + [ctypes.c_size_t] * 2 + \
        [ctypes.c_void_p] * 4 + [ctypes.c_int]
    f.restype = ctypes.c_int
    return f


def cpu_wgrid_invert(nx, ny, dx, dy, l_shift, m_shift, uvw, freq, vis, W, beta, sigma, nu, nv,
                     nthreads=0):
    """FP64 CPU w-gridder dirty image [ny, nx] (not divided by the weight)."""
    uvw = np.ascontiguousarray(uvw, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    vis = np.ascontiguousarray(vis, np.complex64)
    out = np.zeros(nx * ny, np.float64)
    f = _wgrid_fn("oracle_wgrid_invert", None)
    f(nx, ny, dx, dy, l_shift, m_shift, W, beta, sigma, nu, nv, uvw.size // 3, freq.size,
      _p(uvw), _p(freq), _p(vis), _p(out), _nthreads(nthreads))
    return out.reshape(ny, nx)


def cpu_wgrid_predict(nx, ny, dx, dy, l_shift, m_shift, image, uvw, freq, W, beta, sigma, nu,
                      nv, nthreads=0):
    """FP64 CPU w-gridder prediction, complex128 [n_rows, n_chan]."""
    uvw = np.ascontiguousarray(uvw, np.float64)
    freq = np.ascontiguousarray(freq, np.float64)
    img = np.ascontiguousarray(image, np.float64)
    n_rows, n_chan = uvw.size // 3, freq.size
    out = np.zeros(2 * n_rows * n_chan, np.float64)
    f = _wgrid_fn("oracle_wgrid_predict", None)
    f(nx, ny, dx, dy, l_shift, m_shift, W, beta, sigma, nu, nv, n_rows, n_chan, _p(uvw),
      _p(freq), _p(img), _p(out), _nthreads(nthreads))
    return (out[0::2] + 1j * out[1::2]).reshape(n_rows, n_chan)


def cpu_wgrid_last_timing():
    """(visibility-work seconds, plane-work seconds, planes) of the last
    cpu_wgrid_invert / cpu_wgrid_predict call."""
    lib = load()
    out = np.zeros(3)
    f = lib.oracle_wgrid_last_timing
    f.argtypes = [ctypes.c_void_p]
    f.restype = None
    f(_p(out))
    return float(out[0]), float(out[1]), int(out[2])


def es_phi_hat(W, beta, t):
    """Phi(t) = int_{-W/2}^{W/2} exp(beta (sqrt(1 - (2s/W)^2) - 1)) cos(2 pi s t) ds."""
    from scipy.integrate import quad

    def f(s):
        x = 2.0 * s / W
        return np.exp(beta * (np.sqrt(max(0.0, 1.0 - x * x)) - 1.0)) * np.cos(2 * np.pi * s * t)

    val, _ = quad(f, -W / 2, W / 2, limit=400, epsabs=1e-14, epsrel=1e-13)
    return val


# ---------------------------------------------------------------------------
# Polarization codes: the values of the FITS STOKES axis, which
# aocommon::PolarizationEnum uses (aocommon is absent from the reference
# mount: the numbering and the conventions below are parity unpinned).
POL_I, POL_Q, POL_U, POL_V = 1, 2, 3, 4
POL_RR, POL_LL, POL_RL, POL_LR = -1, -2, -3, -4
POL_XX, POL_YY, POL_XY, POL_YX = -5, -6, -7, -8
ALL_POLS = (POL_I, POL_Q, POL_U, POL_V, POL_RR, POL_LL, POL_RL, POL_LR, POL_XX, POL_YY,
            POL_XY, POL_YX)
_SIGMA = [np.array(m, np.complex128) for m in
          ([[1, 0], [0, 1]], [[1, 0], [0, -1]], [[0, 1], [1, 0]], [[0, -1j], [1j, 0]])]


def stokes_from_linear(v):
    """(I, Q, U, V) of linear correlations v[..., 4] = XX, XY, YX, YY, from
    the coherency matrix C = [[XX, XY], [YX, YY]] = I s0 + Q s3 + U s1 - V s2
    (Pauli matrices s; the IAU convention XY = U + iV): S = tr(C s) / 2, with
    the sign of s2 for V."""
    v = np.asarray(v, np.complex128)
    C = v.reshape(v.shape[:-1] + (2, 2))
    tr = [np.einsum("...ij,ji->...", C, s) / 2 for s in _SIGMA]
    return tr[0], tr[1], tr[2], -tr[3]


def convert_from_linear(v, pol):
    """One visibility of polarization pol from v[..., 4] = XX, XY, YX, YY:
    the linear products themselves, Stokes parameters, or circular products
    RR = I + V, LL = I - V, RL = Q + iU, LR = Q - iU
    (gridding/msgridderdata.h:45-49, aocommon ConvertFromLinear)."""
    v = np.asarray(v, np.complex128)
    lin = {POL_XX: 0, POL_XY: 1, POL_YX: 2, POL_YY: 3}
    if pol in lin:
        return v[..., lin[pol]]
    I, Q, U, V = stokes_from_linear(v)
    return {POL_I: I, POL_Q: Q, POL_U: U, POL_V: V, POL_RR: I + V, POL_LL: I - V,
            POL_RL: Q + 1j * U, POL_LR: Q - 1j * U}[pol]


# Per-antenna gain corrections and polarization collapse (the correction
# callback of WGriddingGridder_Simple, wgridder/callback_implementation.h:
# 78-97).  Gain modes in the order of GainMode (gridding/gainmode.h:16-31).
GAIN_XX, GAIN_YY, GAIN_TRACE, GAIN_2VIS_DIAGONAL, GAIN_FULL = range(5)
GAIN_NPOL = {GAIN_XX: 1, GAIN_YY: 1, GAIN_TRACE: 1, GAIN_2VIS_DIAGONAL: 2, GAIN_FULL: 4}


def _jones(resp, nparms):
    """2x2 gain matrices [..., 2, 2] from a cached solution row: nparms 2 is
    diag(p0, p1), nparms 4 is [[p0, p1], [p2, p3]]
    (gridding/visibilitymodifier.h:554-557, :578-579)."""
    resp = np.asarray(resp)
    g = np.zeros(resp.shape[:-1] + (2, 2), np.complex128)
    if nparms == 2:
        g[..., 0, 0] = resp[..., 0]
        g[..., 1, 1] = resp[..., 1]
    else:
        g[..., 0, 0], g[..., 0, 1] = resp[..., 0], resp[..., 1]
        g[..., 1, 0], g[..., 1, 1] = resp[..., 2], resp[..., 3]
    return g


def _apply_gain(mode, v, g1, g2, conjugated):
    """One visibility (n_pol values) corrected by the gains of its two
    antennas: ApplyGain (forward) and ApplyConjugatedGain
    (gridding/visibilitymodifier.h:346-401).  Element order of a 2x2 matrix:
    [0] xx, [1] xy, [2] yx, [3] yy."""
    a = g1.reshape(4)
    b = g2.reshape(4)
    v = np.array(v, np.complex128)
    if mode == GAIN_XX:
        return np.array([np.conj(a[0]) * v[0] * b[0] if conjugated else a[0] * v[0] * np.conj(b[0])])
    if mode == GAIN_YY:
        return np.array([np.conj(a[3]) * v[0] * b[3] if conjugated else a[3] * v[0] * np.conj(b[3])])
    if mode == GAIN_TRACE:
        dd = np.sum(b * np.conj(a)) if conjugated else np.sum(a * np.conj(b))
        return np.array([v[0] * 0.5 * dd])
    if mode == GAIN_2VIS_DIAGONAL:
        # the second line uses the updated first value, as the reference does
        if conjugated:
            v0 = np.conj(a[0]) * v[0] * b[0] + np.conj(a[2]) * v[1] * b[2]
            v1 = np.conj(a[3]) * v[1] * b[3] + np.conj(a[1]) * v0 * b[1]
        else:
            v0 = a[0] * v[0] * np.conj(b[0]) + a[1] * v[1] * np.conj(b[1])
            v1 = a[3] * v[1] * np.conj(b[3]) + a[2] * v0 * np.conj(b[2])
        return np.array([v0, v1])
    V = v.reshape(2, 2)
    if conjugated:
        R = g1.conj().T @ V @ g2
    else:
        R = g1 @ V @ g2.conj().T
    return R.reshape(4)


def correct_and_collapse(vis, antennas, time_offsets, gains, n_ant, nparms, mode, forward,
                         pol=POL_I):
    """vis [rows, chan, n_pol] complex -> corrected, collapsed [rows, chan].

    gains: cached solutions [n_times, n_chan, n_ant, nparms] complex, indexed
    (time_offset * n_chan + ch) * n_ant * nparms
    (gridding/visibilitymodifier.h:550-553).  forward applies the forward
    gain first (direction-dependent PSF mode, :559-562).  Collapse:
    2 polarizations -> v0 + v1 (gridding/msgridderdata.h:40-44), 4 ->
    convert_from_linear(v, pol) (:45-49)."""
    vis = np.asarray(vis)
    n_rows, n_chan, npol = vis.shape
    assert npol == GAIN_NPOL[mode]
    out = np.zeros((n_rows, n_chan), np.complex128)
    for r in range(n_rows):
        a1, a2 = antennas[r]
        t = time_offsets[r]
        for c in range(n_chan):
            g1 = _jones(gains[t, c, a1], nparms)
            g2 = _jones(gains[t, c, a2], nparms)
            v = vis[r, c]
            if forward:
                v = _apply_gain(mode, v, g1, g2, False)
            v = _apply_gain(mode, v, g1, g2, True)
            if npol == 1:
                out[r, c] = v[0]
            elif npol == 2:
                out[r, c] = v[0] + v[1]
            else:
                out[r, c] = convert_from_linear(v, pol)
    return out
