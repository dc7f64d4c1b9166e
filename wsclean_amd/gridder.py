"""Python mirror of WSClean's w-gridder boundary on the MI355X library.

``WGriddingGridder_Simple`` has the constructor arguments and the method set
of ``WGriddingGridderBase`` / ``WGriddingGridder_Simple<NumT>``
(wgridder/gridder_simple.h:20-42, :84-197) and calls the C-ABI of
libwgrid_hip.so (include/wgrid_hip.h) for each of them.  Host numpy buffers
follow the reference layouts:

* ``uvw``: float64 [n_rows, 3] in metres;
* ``freq``: float64 [n_chan] in Hz (increasing or decreasing);
* ``vis``: complex64 [n_rows, n_chan], visibility(row, chan);
* images: float32 [height * width], index x + y * width
  (gridder_simple.h:166-179).

Errors raise ``WgError`` (the reference throws; the C-ABI returns a status
and a message, and the handle stays usable).
"""
import ctypes
import json

import numpy as np

from . import _lib
from ._lib import WgError


# Polarization codes (aocommon::PolarizationEnum values, the FITS STOKES axis)
POL_I, POL_Q, POL_U, POL_V = 1, 2, 3, 4
POL_RR, POL_LL, POL_RL, POL_LR = -1, -2, -3, -4
POL_XX, POL_YY, POL_XY, POL_YX = -5, -6, -7, -8


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.size != int(np.prod(shape)):
        raise WgError("array of size %d, expected %s" % (a.size, shape))
    return a


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.complex64)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class WGriddingGridder_Simple:
    """GPU drop-in for ``WGriddingGridder_Simple<NumT>``.

    Same constructor arguments as gridder_simple.h:84-88.  ``nthreads`` is
    accepted for interface parity; the GPU ignores it.
    """

    def __init__(self, width, height, width_t, height_t, pixel_size_x, pixel_size_y,
                 l_shift, m_shift, nthreads=1, epsilon=1e-4, verbosity=0, tuning=False):
        self._lib = _lib.load()
        self.width, self.height = int(width), int(height)
        self.width_t, self.height_t = int(width_t), int(height_t)
        p = _lib.WgParams(self.width, self.height, self.width_t, self.height_t,
                          float(pixel_size_x), float(pixel_size_y), float(l_shift),
                          float(m_shift), int(nthreads), float(epsilon), int(verbosity),
                          1 if tuning else 0)
        h = ctypes.c_void_p()
        _lib.check(self._lib.wg_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.wg_destroy(self._h)
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

    # -- WGriddingGridderBase --------------------------------------------
    def ConstantMemoryUsage(self):
        v = ctypes.c_size_t()
        _lib.check(self._lib.wg_constant_memory_usage(self._h, ctypes.byref(v)))
        return v.value

    def PerVisibilityMemoryUsage(self):
        v = ctypes.c_size_t()
        _lib.check(self._lib.wg_per_visibility_memory_usage(self._h, ctypes.byref(v)))
        return v.value

    def InitializeInversion(self):
        _lib.check(self._lib.wg_initialize_inversion(self._h))

    def AddInversionData(self, n_rows, n_chan, uvw, freq, vis):
        uvw = _f64(uvw, (n_rows, 3))
        freq = _f64(freq, (n_chan,))
        vis = _c64(vis)
        if vis.size != n_rows * n_chan:
            raise WgError("vis has %d entries, expected %d" % (vis.size, n_rows * n_chan))
        _lib.check(self._lib.wg_add_inversion_data(self._h, n_rows, n_chan, _ptr(uvw),
                                                   _ptr(freq), _ptr(vis)))

    def AddInversionDataWithCorrectionCallback(self, gain_mode, n_polarizations, n_rows, uvws,
                                               frequencies, n_channels, antennas, visibilities,
                                               time_offsets, gains, apply_forward=False,
                                               polarization=POL_I):
        """Mirror of AddInversionDataWithCorrectionCallback
        (wgridder/gridder_simple.h:29-35).  The MsGridder callback state is
        passed explicitly: antennas [n_rows, 2] antenna pairs, time_offsets
        [n_rows] rows of `gains`, gains [n_times, n_chan, n_antennas, nparms]
        complex64 (the cached solution table), visibilities [n_rows, n_chan,
        n_pol] complex64.  gain_mode: 0 xx, 1 yy, 2 trace, 3 2-vis diagonal,
        4 full (GainMode order).  polarization: the gridder's polarization
        (POL_* codes) that 4 linear correlations are converted to."""
        uvw = np.ascontiguousarray(uvws, np.float64).reshape(-1, 3)
        freq = np.ascontiguousarray(frequencies, np.float64)
        ant = np.ascontiguousarray(antennas, np.uint32).reshape(-1, 2)
        vis = np.ascontiguousarray(visibilities, np.complex64)
        toff = np.ascontiguousarray(time_offsets, np.uint64)
        g = np.ascontiguousarray(gains, np.complex64)
        n_times, n_chan_g, n_ant, nparms = g.shape
        if n_chan_g != n_channels or vis.size != n_rows * n_channels * n_polarizations:
            raise ValueError("shape mismatch")
        _lib.check(self._lib.wg_add_inversion_data_corrected(
            self._h, int(gain_mode), int(n_polarizations), int(n_rows), _ptr(uvw), _ptr(freq),
            int(n_channels), _ptr(ant), _ptr(vis), _ptr(toff), _ptr(g), int(n_times), int(n_ant),
            int(nparms), 1 if apply_forward else 0, int(polarization)))

    def FinalizeImage(self, multiplication_factor):
        _lib.check(self._lib.wg_finalize_image(self._h, float(multiplication_factor)))

    def RealImage(self):
        out = np.empty(self.width * self.height, dtype=np.float32)
        _lib.check(self._lib.wg_real_image(self._h, _ptr(out)))
        return out

    def InitializePrediction(self, image_data):
        img = np.ascontiguousarray(image_data, dtype=np.float32).ravel()
        if img.size != self.width * self.height:
            raise WgError("image has %d pixels, expected %d" % (img.size,
                                                               self.width * self.height))
        _lib.check(self._lib.wg_initialize_prediction(self._h, _ptr(img)))

    def PredictVisibilities(self, n_rows, n_chan, uvw, freq, vis):
        """Overwrites ``vis`` (complex64, n_rows * n_chan, C-contiguous)."""
        uvw = _f64(uvw, (n_rows, 3))
        freq = _f64(freq, (n_chan,))
        if not (isinstance(vis, np.ndarray) and vis.dtype == np.complex64
                and vis.flags.c_contiguous and vis.size == n_rows * n_chan):
            raise WgError("vis must be a C-contiguous complex64 array of n_rows*n_chan")
        _lib.check(self._lib.wg_predict_visibilities(self._h, n_rows, n_chan, _ptr(uvw),
                                                     _ptr(freq), _ptr(vis)))

    # -- device-resident variants (torch tensors on the GPU) ----------------
    def AddInversionDataDevice(self, uvw_dev, freq, vis_dev):
        n_rows, n_chan = int(vis_dev.shape[0]), int(vis_dev.shape[1])
        freq = _f64(freq, (n_chan,))
        _lib.check(self._lib.wg_add_inversion_data_device(
            self._h, n_rows, n_chan, ctypes.c_void_p(uvw_dev.data_ptr()), _ptr(freq),
            ctypes.c_void_p(vis_dev.data_ptr())))

    def PredictVisibilitiesDevice(self, uvw_dev, freq, vis_out_dev):
        n_rows, n_chan = int(vis_out_dev.shape[0]), int(vis_out_dev.shape[1])
        freq = _f64(freq, (n_chan,))
        _lib.check(self._lib.wg_predict_visibilities_device(
            self._h, n_rows, n_chan, ctypes.c_void_p(uvw_dev.data_ptr()), _ptr(freq),
            ctypes.c_void_p(vis_out_dev.data_ptr())))

    def RealImageDevice(self, image_dev):
        """RealImage into a device float32 tensor of width*height."""
        self._check_image(image_dev)
        _lib.check(self._lib.wg_real_image_device(self._h, ctypes.c_void_p(image_dev.data_ptr())))

    def InitializePredictionDevice(self, image_dev):
        """InitializePrediction from a device float32 tensor of width*height."""
        self._check_image(image_dev)
        _lib.check(self._lib.wg_initialize_prediction_device(
            self._h, ctypes.c_void_p(image_dev.data_ptr())))

    def ResampleDevice(self, in_dev, in_w, in_h, out_dev, out_w, out_h):
        """Fourier-interpolation resampling of a float32 device image
        [in_h, in_w] into out_dev [out_h, out_w] on this handle's stream."""
        if in_dev.numel() != in_w * in_h or out_dev.numel() != out_w * out_h:
            raise WgError("image sizes do not match the tensors")
        _lib.check(self._lib.wg_resample(self._h, ctypes.c_void_p(in_dev.data_ptr()), in_w, in_h,
                                         ctypes.c_void_p(out_dev.data_ptr()), out_w, out_h))

    def RealImageResampled(self, out_w, out_h):
        """RealImage resampled to out_w x out_h (FinishInversion's resample)."""
        out = np.empty(out_w * out_h, dtype=np.float32)
        _lib.check(self._lib.wg_real_image_resampled(self._h, _ptr(out), out_w, out_h))
        return out

    def InitializePredictionResampled(self, image, in_w, in_h):
        """InitializePrediction of a model of in_w x in_h, resampled to the
        gridder's width x height (StartPredict's resample)."""
        img = np.ascontiguousarray(image, dtype=np.float32).ravel()
        if img.size != in_w * in_h:
            raise WgError("image has %d pixels, expected %d" % (img.size, in_w * in_h))
        _lib.check(self._lib.wg_initialize_prediction_resampled(self._h, _ptr(img), in_w, in_h))

    def _check_image(self, t):
        import torch
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
                and t.numel() == self.width * self.height):
            raise WgError("image must be a contiguous float32 device tensor of width*height")

    def PrepareVisibilities(self, uvw_dev, freq, n_rows, n_chan):
        freq = _f64(freq, (n_chan,))
        self._prep_freq = freq
        _lib.check(self._lib.wg_prepare_visibilities(
            self._h, n_rows, n_chan, ctypes.c_void_p(uvw_dev.data_ptr()), _ptr(freq)))

    def AddInversionPrepared(self, vis_dev):
        _lib.check(self._lib.wg_add_inversion_prepared(self._h,
                                                       ctypes.c_void_p(vis_dev.data_ptr())))

    def PredictPrepared(self, vis_out_dev):
        _lib.check(self._lib.wg_predict_prepared(self._h,
                                                 ctypes.c_void_p(vis_out_dev.data_ptr())))

    def Autotune(self, uvw_dev, freq, n_rows, n_chan, max_sample_vis=1 << 27):
        """Measure the candidate kernel shapes and tiles on a row sample of
        the device uvw and keep the fastest; returns the trials (fastest
        first)."""
        freq = _f64(freq, (n_chan,))
        buf = ctypes.create_string_buffer(1 << 16)
        _lib.check(self._lib.wg_autotune(self._h, n_rows, n_chan,
                                         ctypes.c_void_p(uvw_dev.data_ptr()), _ptr(freq),
                                         int(max_sample_vis), buf, 1 << 16))
        return json.loads(buf.value.decode())

    def CaptureMajorCycle(self, vis_dev=None, vis_out_dev=None, factor=1.0):
        """Capture clear + invert(vis_dev) + FinalizeImage(factor) and / or
        predict(vis_out_dev) on the prepared set as a HIP graph."""
        what = (1 if vis_dev is not None else 0) | (2 if vis_out_dev is not None else 0)
        _lib.check(self._lib.wg_capture_major_cycle(
            self._h, ctypes.c_void_p(vis_dev.data_ptr() if vis_dev is not None else 0),
            ctypes.c_void_p(vis_out_dev.data_ptr() if vis_out_dev is not None else 0),
            float(factor), what))

    def ReplayMajorCycle(self, count=1):
        _lib.check(self._lib.wg_replay_major_cycle(self._h, int(count)))

    # -- introspection ------------------------------------------------------
    def plan(self):
        buf = ctypes.create_string_buffer(4096)
        _lib.check(self._lib.wg_plan_json(self._h, buf, 4096))
        s = buf.value.decode()
        return json.loads(s) if s else {}

    def set_timing(self, on=True):
        _lib.check(self._lib.wg_set_timing(self._h, 1 if on else 0))

    def stage_times(self, reset=True):
        out = (ctypes.c_double * 14)()
        _lib.check(self._lib.wg_stage_times(self._h, out, 1 if reset else 0))
        return dict(zip(["plan", "sort", "grid", "fft", "degrid", "total",
                         "grid_kernel_ms", "grid_launches", "grid_bytes",
                         "degrid_kernel_ms", "degrid_launches", "degrid_bytes",
                         "grid_passes", "degrid_passes"], list(out)))

    def set_max_chunk_vis(self, max_vis=0):
        """Test hook: largest engine pass of the drop-in calls (0: automatic)."""
        _lib.check(self._lib.wgx_set_max_chunk_vis(self._h, int(max_vis)))

    def device_bytes(self):
        """Device bytes the handle holds."""
        v = ctypes.c_size_t()
        _lib.check(self._lib.wgx_device_bytes(self._h, ctypes.byref(v)))
        return v.value

    def set_tuning(self, W=0, sigma=0.0, tile=0, sub_slab=0, slab_bytes=0):
        _lib.check(self._lib.wg_set_tuning(self._h, int(W), float(sigma), int(tile),
                                           int(sub_slab), int(slab_bytes)))

    @staticmethod
    def min_epsilon():
        """Smallest epsilon the library accepts (FP32 grids and FFT)."""
        lib = _lib.load()
        v = ctypes.c_double()
        _lib.check(lib.wg_min_epsilon(ctypes.byref(v)))
        return v.value

    def set_precision(self, split=-1):
        """Matrix-core gridding arithmetic for W <= 8: -1 split f16 pairs
        where epsilon allows, 0 FP32, 1 split whenever W <= 8."""
        _lib.check(self._lib.wg_set_precision(self._h, int(split)))

    def sort_keys(self, uvw, freq, vis=None):
        """Test hook: (unsorted bucket keys, stable sorted permutation, plan)."""
        uvw = _f64(uvw)
        freq = _f64(freq)
        n_rows, n_chan = uvw.size // 3, freq.size
        keys = np.empty(n_rows * n_chan, dtype=np.uint32)
        perm = np.empty(n_rows * n_chan, dtype=np.uint32)
        nv = ctypes.c_size_t()
        buf = ctypes.create_string_buffer(4096)
        vp = None
        if vis is not None:
            vis = _c64(vis)
            vp = _ptr(vis)
        _lib.check(self._lib.wgx_sort_keys(self._h, n_rows, n_chan, _ptr(uvw), _ptr(freq), vp,
                                           _ptr(keys), _ptr(perm), ctypes.byref(nv), buf, 4096))
        return keys, perm[: nv.value], json.loads(buf.value.decode())


def radix_sort(keys, key_bits):
    """Test hook: the device radix sort of uint32 keys by their low key_bits
    (stable); returns (sorted keys, element indices in sorted order)."""
    lib = _lib.load()
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    ko = np.empty_like(k)
    po = np.empty_like(k)
    _lib.check(lib.wgx_radix_sort(k.size, int(key_bits), _ptr(k), _ptr(ko), _ptr(po)))
    return ko, po


def fft_lines(data, inverse=False):
    """Test hook: batched LDS FFT of each row of a complex64 2-D array."""
    lib = _lib.load()
    a = np.ascontiguousarray(data, dtype=np.complex64).copy()
    nlines, n = a.shape
    _lib.check(lib.wgx_fft(n, nlines, 1 if inverse else 0, _ptr(a)))
    return a


def synth_visibilities(uvw_dev, freq, src, noise_sigma, seed, vis_dev):
    """Benchmark / test input hook (wgx_synth_visibilities): fills the
    device tensor vis_dev [n_rows, n_chan, 2] float32 with the visibilities
    of the sources src (n_src x {l, m, n - 1, S, gaussian sigma}) at the
    device uvw [n_rows, 3] (metres) plus seeded Gaussian noise."""
    lib = _lib.load()
    freq = _f64(freq)
    src = np.ascontiguousarray(src, dtype=np.float64).reshape(-1, 5)
    n_rows = int(uvw_dev.shape[0])
    _lib.check(lib.wgx_synth_visibilities(
        ctypes.c_void_p(uvw_dev.data_ptr()), n_rows, freq.size, _ptr(freq), src.shape[0],
        _ptr(src), float(noise_sigma), int(seed), ctypes.c_void_p(vis_dev.data_ptr())))
