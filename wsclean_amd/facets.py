"""Facet batches on the MI355X library (``wg_facet_batch_*`` in
include/wgrid_hip.h).

WSClean images a faceted field with one ``WGriddingGridder_Simple`` per
facet (gridding/msgriddermanager.cpp:41-77) and lets ``BatchInvert``
(gridding/msgriddermanager.cpp:227-396) feed one chunk read to all of them.
``FacetBatch`` is that set of gridders in one object: every facet has the
constructor arguments of its own gridder (image size, pixel size, l/m
shift), and all facets share one sort and one gridding / degridding launch
per w-plane sub-slab on the device.

Visibilities live on the device, in buffers shared by the facets: facet f
reads rows ``[row_offsets[f], row_offsets[f] + n_rows[f])`` of ``uvw``
(float64 [rows, 3]) and ``vis`` (complex64 [rows, n_chan], as float32
[rows, n_chan, 2]).  ``Predict`` writes the facets' visibilities one block
after another: facet f's [n_rows[f], n_chan] block starts at
``block_offsets()[f]``.
"""
import ctypes
import json

import numpy as np

from . import _lib
from ._lib import WgError


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dev(t):
    return ctypes.c_void_p(t.data_ptr())


class FacetBatch:
    """``facets``: a list of dicts with the WGriddingGridder_Simple
    constructor arguments (gridder_simple.h:84-88) of each facet:
    width, height, width_t, height_t, pixel_size_x, pixel_size_y, l_shift,
    m_shift and epsilon (the smallest epsilon applies to all facets)."""

    def __init__(self, facets):
        self._lib = _lib.load()
        self.facets = [dict(f) for f in facets]
        arr = (_lib.WgParams * len(facets))()
        for i, f in enumerate(self.facets):
            arr[i] = _lib.WgParams(int(f["width"]), int(f["height"]),
                                   int(f.get("width_t", f["width"])),
                                   int(f.get("height_t", f["height"])),
                                   float(f["pixel_size_x"]), float(f["pixel_size_y"]),
                                   float(f.get("l_shift", 0.0)), float(f.get("m_shift", 0.0)),
                                   1, float(f.get("epsilon", 1e-4)), 0, 0)
        h = ctypes.c_void_p()
        _lib.check(self._lib.wg_facet_batch_create(arr, len(facets), ctypes.byref(h)))
        self._h = h
        self._blocks = None

    def close(self):
        if getattr(self, "_h", None):
            self._lib.wg_facet_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.facets)

    def Prepare(self, uvw_dev, freq, row_offsets, n_rows):
        freq = np.ascontiguousarray(freq, np.float64)
        ro = np.ascontiguousarray(row_offsets, np.uint64)
        nr = np.ascontiguousarray(n_rows, np.uint64)
        if ro.size != len(self) or nr.size != len(self):
            raise WgError("one row offset and one row count per facet")
        if np.any(ro + nr > uvw_dev.shape[0]):
            raise WgError("facet rows beyond the uvw buffer")
        _lib.check(self._lib.wg_facet_batch_prepare(self._h, freq.size, _ptr(freq),
                                                    _dev(uvw_dev), _ptr(ro), _ptr(nr)))
        self._blocks = np.concatenate([[0], np.cumsum(nr * freq.size)]).astype(np.int64)

    def block_offsets(self):
        """Start of every facet's block in the Predict output (visibilities),
        and the total as the last entry."""
        return self._blocks

    def InitializeInversion(self):
        _lib.check(self._lib.wg_facet_batch_initialize_inversion(self._h))

    def AddInversion(self, vis_dev):
        _lib.check(self._lib.wg_facet_batch_add_inversion(self._h, _dev(vis_dev)))

    def FinalizeImage(self, facet, factor):
        _lib.check(self._lib.wg_facet_batch_finalize_image(self._h, int(facet), float(factor)))

    def _size(self, facet):
        f = self.facets[facet]
        return int(f["width"]) * int(f["height"])

    def RealImage(self, facet):
        out = np.empty(self._size(facet), np.float32)
        _lib.check(self._lib.wg_facet_batch_real_image(self._h, int(facet), _ptr(out)))
        return out

    def RealImageDevice(self, facet, image_dev):
        if image_dev.numel() != self._size(facet) or not image_dev.is_contiguous():
            raise WgError("image must be a contiguous float32 device tensor of width*height")
        _lib.check(self._lib.wg_facet_batch_real_image_device(self._h, int(facet),
                                                              _dev(image_dev)))

    def InitializePrediction(self, facet, image):
        img = np.ascontiguousarray(image, np.float32).ravel()
        if img.size != self._size(facet):
            raise WgError("image has %d pixels, expected %d" % (img.size, self._size(facet)))
        _lib.check(self._lib.wg_facet_batch_initialize_prediction(self._h, int(facet), _ptr(img)))

    def InitializePredictionDevice(self, facet, image_dev):
        if image_dev.numel() != self._size(facet) or not image_dev.is_contiguous():
            raise WgError("image must be a contiguous float32 device tensor of width*height")
        _lib.check(self._lib.wg_facet_batch_initialize_prediction_device(self._h, int(facet),
                                                                         _dev(image_dev)))

    def Predict(self, vis_out_dev):
        n = ctypes.c_size_t()
        _lib.check(self._lib.wg_facet_batch_virtual_count(self._h, ctypes.byref(n)))
        if vis_out_dev.numel() < 2 * n.value:
            raise WgError("output holds %d visibilities, the batch has %d"
                          % (vis_out_dev.numel() // 2, n.value))
        _lib.check(self._lib.wg_facet_batch_predict(self._h, _dev(vis_out_dev)))

    def CaptureMajorCycle(self, vis_dev=None, vis_out_dev=None, factors=None):
        """Capture clear + invert(vis_dev) + FinalizeImage(factors[f]) and / or
        predict(vis_out_dev) of every facet as one HIP graph."""
        what = (1 if vis_dev is not None else 0) | (2 if vis_out_dev is not None else 0)
        fac = np.ascontiguousarray(factors if factors is not None else np.ones(len(self)),
                                   np.float64)
        _lib.check(self._lib.wg_facet_batch_capture_major_cycle(
            self._h, ctypes.c_void_p(vis_dev.data_ptr() if vis_dev is not None else 0),
            ctypes.c_void_p(vis_out_dev.data_ptr() if vis_out_dev is not None else 0),
            _ptr(fac), what))

    def ReplayMajorCycle(self, count=1):
        _lib.check(self._lib.wg_facet_batch_replay_major_cycle(self._h, int(count)))

    def set_timing(self, on=True):
        _lib.check(self._lib.wg_facet_batch_set_timing(self._h, 1 if on else 0))

    def stage_times(self, reset=True):
        out = (ctypes.c_double * 14)()
        _lib.check(self._lib.wg_facet_batch_stage_times(self._h, out, 1 if reset else 0))
        return dict(zip(["plan", "sort", "grid", "fft", "degrid", "total",
                         "grid_kernel_ms", "grid_launches", "grid_bytes",
                         "degrid_kernel_ms", "degrid_launches", "degrid_bytes",
                         "grid_passes", "degrid_passes"], list(out)))

    def plan(self):
        buf = ctypes.create_string_buffer(1 << 20)
        _lib.check(self._lib.wg_facet_batch_plan_json(self._h, buf, 1 << 20))
        s = buf.value.decode()
        return json.loads(s) if s else {}


def facet_atlas(facets, extents, nvis):
    """Host-only layout of a facet batch (no GPU): the plans and atlas
    placement the library makes for facets with visibility extents
    ``extents`` (per facet umin, umax, vmin, vmax, wmin, wmax in
    wavelengths after the w >= 0 flip) and visibility counts ``nvis``."""
    lib = _lib.load()
    arr = (_lib.WgParams * len(facets))()
    for i, f in enumerate(facets):
        arr[i] = _lib.WgParams(int(f["width"]), int(f["height"]),
                               int(f.get("width_t", f["width"])),
                               int(f.get("height_t", f["height"])),
                               float(f["pixel_size_x"]), float(f["pixel_size_y"]),
                               float(f.get("l_shift", 0.0)), float(f.get("m_shift", 0.0)),
                               1, float(f.get("epsilon", 1e-4)), 0, 0)
    ext = np.ascontiguousarray(extents, np.float64).reshape(len(facets), 6)
    nv = np.ascontiguousarray(nvis, np.int64)
    buf = ctypes.create_string_buffer(1 << 20)
    _lib.check(lib.wgx_facet_atlas(arr, len(facets), _ptr(ext), _ptr(nv), buf, 1 << 20))
    return json.loads(buf.value.decode())
