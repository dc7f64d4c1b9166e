"""Imaging weights and caller-side row preparation on the MI355X library.

``ImageWeights`` mirrors WSClean's class of the same name
(structures/imageweights.h:22-132): construct with the weighting mode and
image geometry, ``Grid`` every chunk of rows, ``FinishGridding``, optional
uv limits / tapers / rank filter, then ``GetWeight`` / ``GetGrid``.  The
grid lives in HBM.

``prepare_rows`` is MsGridderData's per-row preparation
(gridding/msgridderdata.h:761-789, :815-886, :929-941, :37-52) on the
device: PSF visibilities or data minus model, channel selection,
visibility-weight mode, imaging weights, weight sums and the polarization
collapse, writing visibilities ready for the gridder's device entry points.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import WgError

NATURAL, UNIFORM, BRIGGS = 0, 1, 2
VIS_NORMAL, VIS_SQUARED, VIS_UNIT = 0, 1, 2


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dev(t):
    return ctypes.c_void_p(t.data_ptr() if t is not None else 0)


class ImageWeights:
    def __init__(self, mode, width, height, pixel_scale_x, pixel_scale_y,
                 weights_as_taper=False, super_weight=1.0, briggs_robustness=0.0):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.wg_image_weights_create(
            int(mode), float(briggs_robustness), int(width), int(height), float(pixel_scale_x),
            float(pixel_scale_y), 1 if weights_as_taper else 0, float(super_weight),
            ctypes.byref(h)))
        self._h = h
        w, hh = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.check(self._lib.wg_image_weights_size(h, ctypes.byref(w), ctypes.byref(hh)))
        self.width, self.height = w.value, hh.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.wg_image_weights_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def Width(self):
        return self.width

    def Height(self):
        return self.height

    def Grid(self, uvw_dev, freq, weights_dev):
        """weights_dev: float32 device tensor [rows, chan, pol]."""
        freq = np.ascontiguousarray(freq, np.float64)
        n_rows, n_chan, n_pol = (int(x) for x in weights_dev.shape)
        if n_chan != freq.size or uvw_dev.shape[0] != n_rows:
            raise WgError("shape mismatch")
        _lib.check(self._lib.wg_image_weights_grid(self._h, n_rows, n_chan, n_pol,
                                                   _dev(uvw_dev), _ptr(freq), _dev(weights_dev)))

    def FinishGridding(self):
        _lib.check(self._lib.wg_image_weights_finish(self._h))

    def SetMinUVRange(self, min_uv):
        _lib.check(self._lib.wg_image_weights_set_min_uv_range(self._h, float(min_uv)))

    def SetMaxUVRange(self, max_uv):
        _lib.check(self._lib.wg_image_weights_set_max_uv_range(self._h, float(max_uv)))

    def SetTukeyTaper(self, transition, max_uv):
        _lib.check(self._lib.wg_image_weights_set_tukey_taper(self._h, float(transition),
                                                              float(max_uv)))

    def SetTukeyInnerTaper(self, transition, min_uv):
        _lib.check(self._lib.wg_image_weights_set_tukey_inner_taper(self._h, float(transition),
                                                                    float(min_uv)))

    def SetEdgeTaper(self, size):
        _lib.check(self._lib.wg_image_weights_set_edge_taper(self._h, float(size)))

    def SetEdgeTukeyTaper(self, transition, edge):
        _lib.check(self._lib.wg_image_weights_set_edge_tukey_taper(self._h, float(transition),
                                                                   float(edge)))

    def SetGaussianTaper(self, beam_size):
        _lib.check(self._lib.wg_image_weights_set_gaussian_taper(self._h, float(beam_size)))

    def RankFilter(self, rank_limit, window):
        _lib.check(self._lib.wg_image_weights_rank_filter(self._h, float(rank_limit),
                                                          int(window)))

    def GetGrid(self):
        img = np.empty((self.height, self.width), np.float64)
        _lib.check(self._lib.wg_image_weights_get_grid(self._h, _ptr(img)))
        return img

    def GetWeight(self, u, v):
        u = np.ascontiguousarray(u, np.float64).ravel()
        v = np.ascontiguousarray(v, np.float64).ravel()
        out = np.empty(u.size, np.float64)
        _lib.check(self._lib.wg_image_weights_sample(self._h, u.size, _ptr(u), _ptr(v),
                                                     _ptr(out)))
        return out


def prepare_rows(gridder, uvw_dev, freq, weights_dev, out_dev, vis_dev=None, model_dev=None,
                 selected=None, psf=False, psf_dl=0.0, psf_dm=0.0, weighting=VIS_NORMAL,
                 image_weights=None, polarization=1, no_alloc=False):
    """Weighted, collapsed visibilities [rows, chan] (complex64 as float32
    [..., 2]) into out_dev on the gridder handle's stream; returns the
    weight sums (total_weight, visibility_weight_sum, max_gridded_weight,
    gridded_visibility_count).  weights_dev: float32 [rows, chan, pol];
    vis_dev / model_dev: float32 [rows, chan, pol, 2].  polarization: the
    gridder's polarization (gridder.POL_* codes) 4 linear correlations are
    converted to.  no_alloc: use wg_prepare_rows_noalloc (no device
    allocation; at most 65536 channels)."""
    freq = np.ascontiguousarray(freq, np.float64)
    n_rows, n_chan, n_pol = (int(x) for x in weights_dev.shape)
    if n_chan != freq.size or out_dev.numel() != 2 * n_rows * n_chan:
        raise WgError("shape mismatch")
    for t in (vis_dev, model_dev):
        if t is not None and t.numel() != 2 * n_rows * n_chan * n_pol:
            raise WgError("vis / model must be [rows, chan, pol] complex")
    sel = None
    if selected is not None:
        sel = np.ascontiguousarray(selected, np.uint8)
        if sel.size != n_chan:
            raise WgError("one selection flag per channel")
    stats = np.zeros(4, np.float64)
    fn = gridder._lib.wg_prepare_rows_noalloc if no_alloc else gridder._lib.wg_prepare_rows
    _lib.check(fn(
        gridder._h, n_rows, n_chan, n_pol, _dev(uvw_dev), _ptr(freq), _dev(vis_dev),
        _dev(weights_dev), _dev(model_dev), _ptr(sel) if sel is not None else None,
        1 if psf else 0, float(psf_dl), float(psf_dm), int(weighting), int(polarization),
        image_weights.handle if image_weights is not None else None, _dev(out_dev),
        _ptr(stats)))
    return dict(total_weight=float(stats[0]), visibility_weight_sum=float(stats[1]),
                max_gridded_weight=float(stats[2]), gridded_visibility_count=int(stats[3]))
