"""Oracle for imaging weights and the caller-side row preparation: TEST
INFRASTRUCTURE ONLY (used by tests/ as the checker; nothing in the product
imports it).

* ``ImageWeightsOracle``: numpy restatement of WSClean's ImageWeights
  (structures/imageweights.h:22-132, structures/imageweights.cpp:33-390):
  the half-plane weight grid, uv -> cell mapping (imageweights.h:74-95),
  gridding of row weights (imageweights.cpp:76-114), FinishGridding for
  natural / uniform / Briggs weighting (:116-150), the uv-range limits and
  tapers (:152-334), the rank filter and its window mean (:302-317,
  :358-388), GetGrid (:270-287) and GetWeight (imageweights.h:97-105).
* ``prepare_rows``: restatement of MsGridderData::CalculateWeights +
  ApplyWeights (gridding/msgridderdata.h:761-789, :815-886) with
  RotateVisibilities (:929-941) and the polarization collapse CollapseData
  (:37-52): PSF visibilities, model subtraction, channel selection,
  visibility weighting mode, imaging weights, the weight sums, and the
  Stokes-I collapse.

aocommon (BandData, Polarization) is absent from the reference mount.  The
channel wavelength is taken as c / f in FP64 (BandData::ChannelWavelength),
and the 4-polarization collapse as 0.5 (XX + YY) for Stokes I: both are
parity unpinned beyond the closed forms the tests check.
"""
import numpy as np

C_LIGHT = 299792458.0

NATURAL, UNIFORM, BRIGGS = 0, 1, 2
VIS_NORMAL, VIS_SQUARED, VIS_UNIT = 0, 1, 2


def _tukey(x, n):
    # imageweights.h:112-114
    return 0.5 * (1.0 + np.cos((np.pi / n) * (x - n)))


class ImageWeightsOracle:
    def __init__(self, mode, width, height, psx, psy, weights_as_taper=False, super_weight=1.0,
                 robustness=0.0):
        # imageweights.cpp:33-48 (round, then make even)
        self.mode = mode
        self.robust = robustness
        w = int(round(width / super_weight))
        h = int(round(height / super_weight))
        if w % 2:
            w += 1
        if h % 2:
            h += 1
        self.W, self.H = w, h
        self.psx, self.psy = psx, psy
        self.taper = weights_as_taper
        self.grid = np.zeros((h // 2, w), np.float64)
        self.total = 0.0

    def uv_to_xy(self, u, v):
        # imageweights.h:74-81
        u = np.asarray(u, np.float64)
        v = np.asarray(v, np.float64)
        neg = v < 0.0
        u = np.where(neg, -u, u)
        v = np.where(neg, -v, v)
        x = np.floor(u * self.W * self.psx + self.W // 2).astype(np.int64)
        y = np.floor(v * self.H * self.psy).astype(np.int64)
        return x, y

    def inside(self, x, y):
        return (x >= 0) & (x < self.W) & (y < self.H // 2)

    def grid_rows(self, uvw, freq, weights):
        """weights [rows][chan][pol] float32, in row, channel, polarization
        order (imageweights.cpp:76-114)."""
        uvw = np.asarray(uvw, np.float64)
        wts = np.asarray(weights, np.float32).copy()
        if self.taper:
            wts[wts != 0] = 1.0
        n_rows, n_chan, n_pol = wts.shape
        lam = C_LIGHT / np.asarray(freq, np.float64)
        um, vm = uvw[:, 0].copy(), uvw[:, 1].copy()
        neg = vm < 0.0
        um[neg] = -um[neg]
        vm[neg] = -vm[neg]
        u = um[:, None] / lam[None, :]
        v = vm[:, None] / lam[None, :]
        x, y = self.uv_to_xy(u, v)
        ok = self.inside(x, y)
        idx = np.broadcast_to((x + y * self.W)[:, :, None], wts.shape)
        okp = np.broadcast_to(ok[:, :, None], wts.shape)
        flat = self.grid.reshape(-1)
        np.add.at(flat, idx[okp], wts[okp].astype(np.float64))
        for val in wts[okp].astype(np.float64):
            self.total += val

    def finish(self):
        # imageweights.cpp:116-150
        g = self.grid
        nz = g != 0.0
        if self.mode == BRIGGS:
            avg = 0.0
            for val in g.reshape(-1):
                avg += val * val
            avg /= self.total
            num = 5.0 * np.exp(-np.log(10.0) * self.robust)
            ssq = num * num / avg
            g[nz] = 1.0 / (1.0 + g[nz] * ssq)
        elif self.mode == UNIFORM:
            g[nz] = 1.0 / g[nz]
        else:
            g[nz] = 1.0

    def _uv_of_cells(self):
        # xyToUV for y >= 0 (imageweights.h:83-90)
        y, x = np.mgrid[0:self.H // 2, 0:self.W]
        u = (x - self.W // 2) / (self.W * self.psx)
        v = y / (self.H * self.psy)
        return u, v

    def set_min_uv_range(self, min_uv):
        u, v = self._uv_of_cells()
        self.grid[u * u + v * v < min_uv * min_uv] = 0.0

    def set_max_uv_range(self, max_uv):
        u, v = self._uv_of_cells()
        self.grid[u * u + v * v > max_uv * max_uv] = 0.0

    def set_tukey_taper(self, transition, max_uv):
        u, v = self._uv_of_cells()
        d2 = u * u + v * v
        out = d2 > max_uv * max_uv
        mid = ~out & (d2 > (max_uv - transition) ** 2)
        self.grid[out] = 0.0
        self.grid[mid] *= _tukey(max_uv - np.sqrt(d2[mid]), transition)

    def set_tukey_inner_taper(self, transition, min_uv):
        u, v = self._uv_of_cells()
        d2 = u * u + v * v
        inner = d2 < min_uv * min_uv
        mid = ~inner & (d2 < (min_uv + transition) ** 2)
        self.grid[inner] = 0.0
        self.grid[mid] *= _tukey(np.sqrt(d2[mid]) - min_uv, transition)

    def _max_uv(self):
        # xyToUV(width, height / 2)
        return (self.W - self.W // 2) / (self.W * self.psx), (self.H // 2) / (self.H * self.psy)

    def set_edge_taper(self, size):
        u, v = self._uv_of_cells()
        mu, mv = self._max_uv()
        self.grid[(mu - np.abs(u) < size) | (mv - np.abs(v) < size)] = 0.0

    def set_edge_tukey_taper(self, transition, edge):
        u, v = self._uv_of_cells()
        mu, mv = self._max_uv()
        ud = mu - np.abs(u)
        vd = mv - np.abs(v)
        total = transition + edge
        zero = (ud < edge) | (vd < edge)
        mid = ~zero & ((ud < total) | (vd < total))
        ru = np.minimum(ud - edge, transition)
        rv = np.minimum(vd - edge, transition)
        self.grid[zero] = 0.0
        self.grid[mid] *= _tukey(ru[mid], transition) * _tukey(rv[mid], transition)

    def set_gaussian_taper(self, beam_size):
        # imageweights.cpp:333-356
        half_power_uv = 1.0 / (beam_size * 2.0 * np.pi)
        sigma_to_hp = float(2.0 * np.sqrt(2.0 * np.log(np.longdouble(2.0))))
        m2s = half_power_uv * sigma_to_hp
        m2s *= -2.0 * m2s
        u, v = self._uv_of_cells()
        nz = self.grid != 0.0
        self.grid[nz] = self.grid[nz] * np.exp((u[nz] ** 2 + v[nz] ** 2) / m2s)

    def window_mean(self, x, y, win):
        # imageweights.cpp:358-388: the window [x - d, x + d) x [y - d, y + d)
        d = win // 2
        x1 = 0 if x <= d else x - d
        y1 = 0 if y <= d else y - d
        x2 = self.W if x + d >= self.W else x + d
        y2 = self.H // 2 if y + d >= self.H // 2 else y + d
        cnt = 0
        s = 0.0
        for yi in range(y1, y2):
            for xi in range(x1, x2):
                w = self.grid[yi, xi]
                if w != 0.0:
                    cnt += 1
                    s += w
        return s / cnt

    def rank_filter(self, rank_limit, win):
        # imageweights.cpp:302-317
        new = self.grid.copy()
        for y in range(self.H // 2):
            for x in range(self.W):
                w = self.grid[y, x]
                if w != 0.0:
                    mean = self.window_mean(x, y, win)
                    if w > mean * rank_limit:
                        new[y, x] = mean * rank_limit
        self.grid = new

    def get_grid(self):
        # imageweights.cpp:270-287
        h2 = self.H // 2
        img = np.zeros((self.H, self.W), np.float64)
        for y in range(h2):
            img[h2 - 1 - y, ::-1] = self.grid[y]
            img[h2 + y, :] = self.grid[y]
        return img

    def get_weight(self, u, v):
        x, y = self.uv_to_xy(u, v)
        ok = self.inside(x, y)
        out = np.zeros(np.shape(x), np.float64)
        out[ok] = self.grid.reshape(-1)[(x + y * self.W)[ok]]
        return out


def prepare_rows(uvw, freq, vis, weights, model=None, selected=None, psf=False, psf_dl=0.0,
                 psf_dm=0.0, weighting=VIS_NORMAL, image_weights=None, polarization=1):
    """Returns (out [rows][chan] complex64, stats dict).  vis, weights,
    model: [rows][chan][pol] (complex64 / float32).  Sequential in row,
    channel, polarization order, float32 products as the reference."""
    uvw = np.asarray(uvw, np.float64)
    wts = np.asarray(weights, np.float32).copy()
    n_rows, n_chan, n_pol = wts.shape
    lam = C_LIGHT / np.asarray(freq, np.float64)
    if psf:
        v = np.ones((n_rows, n_chan, n_pol), np.complex64)
        if psf_dl != 0.0 or psf_dm != 0.0:
            dn = np.sqrt(1.0 - psf_dl * psf_dl - psf_dm * psf_dm) - 1.0
            shift = 2.0 * np.pi * (uvw[:, 0] * psf_dl + uvw[:, 1] * psf_dm + uvw[:, 2] * dn)
            rad = shift[:, None] / lam[None, :]
            ph = (np.cos(rad).astype(np.float32) + 1j * np.sin(rad).astype(np.float32)).astype(
                np.complex64)
            v = v * ph[:, :, None]
    else:
        v = np.asarray(vis, np.complex64).copy()
    if model is not None:
        v = (v - np.asarray(model, np.complex64)).astype(np.complex64)
    if selected is not None:
        wts[:, ~np.asarray(selected, bool), :] = 0.0
    if weighting == VIS_SQUARED:
        wts = wts * wts
    elif weighting == VIS_UNIT:
        wts[wts != 0] = 1.0
    if image_weights is not None:
        iw = image_weights.get_weight(uvw[:, 0:1] / lam[None, :],
                                      uvw[:, 1:2] / lam[None, :]).astype(np.float32)
    else:
        iw = np.ones((n_rows, n_chan), np.float32)
    cum = (wts * iw[:, :, None]).astype(np.float32)
    has = cum != 0.0
    stats = dict(total_weight=0.0, visibility_weight_sum=0.0, max_gridded_weight=0.0,
                 gridded_visibility_count=0)
    # sums in the reference's order (row, channel, polarization)
    for r in range(n_rows):
        for c in range(n_chan):
            for p in range(n_pol):
                if p == 0:
                    stats["visibility_weight_sum"] += float(np.float32(wts[r, c, 0] *
                                                                       np.float32(has[r, c, 0])))
                    stats["max_gridded_weight"] = max(float(cum[r, c, 0]),
                                                      stats["max_gridded_weight"])
                    stats["gridded_visibility_count"] += int(has[r, c, 0])
                stats["total_weight"] += float(cum[r, c, p])
    v = (v * cum).astype(np.complex64)
    if n_pol == 1:
        out = v[:, :, 0]
    elif n_pol == 2:
        out = v[:, :, 0] + v[:, :, 1]
    else:
        from .oracle import convert_from_linear
        out = convert_from_linear(v, polarization).astype(np.complex64)
    return out.astype(np.complex64), stats
