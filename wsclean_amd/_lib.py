"""ctypes binding of libwgrid_hip.so (the C-ABI in include/wgrid_hip.h).

The library is built in-tree by ``python -m wsclean_amd.build``.  Loading
fails loudly if it is missing: there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwgrid_hip.so")

c_size = ctypes.c_size_t
c_dbl_p = ctypes.POINTER(ctypes.c_double)
c_flt_p = ctypes.POINTER(ctypes.c_float)
c_u32_p = ctypes.POINTER(ctypes.c_uint32)


class WgParams(ctypes.Structure):
    _fields_ = [
        ("width", c_size), ("height", c_size),
        ("width_t", c_size), ("height_t", c_size),
        ("pixel_size_x", ctypes.c_double), ("pixel_size_y", ctypes.c_double),
        ("l_shift", ctypes.c_double), ("m_shift", ctypes.c_double),
        ("nthreads", c_size), ("epsilon", ctypes.c_double),
        ("verbosity", c_size), ("tuning", ctypes.c_int),
    ]


class WgRequest(ctypes.Structure):
    _fields_ = [
        ("npix", c_size), ("pixel_size", ctypes.c_double), ("epsilon", ctypes.c_double),
        ("cycles", ctypes.c_int), ("n_rows", c_size), ("n_chan", c_size),
        ("uvw_dev", ctypes.c_void_p), ("freq", ctypes.c_void_p), ("vis_dev", ctypes.c_void_p),
        ("vis_out_dev", ctypes.c_void_p), ("image_host", ctypes.c_void_p),
        ("residual_dev", ctypes.c_void_p),
    ]


# (name, restype, argtypes) for every symbol declared in include/wgrid_hip.h
SIGNATURES = [
    ("wg_create", ctypes.c_int, [ctypes.POINTER(WgParams), ctypes.POINTER(ctypes.c_void_p)]),
    ("wg_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("wg_constant_memory_usage", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(c_size)]),
    ("wg_per_visibility_memory_usage", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(c_size)]),
    ("wg_initialize_inversion", ctypes.c_int, [ctypes.c_void_p]),
    ("wg_add_inversion_data", ctypes.c_int,
     [ctypes.c_void_p, c_size, c_size, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_finalize_image", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("wg_real_image", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_initialize_prediction", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_real_image_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_initialize_prediction_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_predict_visibilities", ctypes.c_int,
     [ctypes.c_void_p, c_size, c_size, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_last_error", ctypes.c_char_p, []),
    ("wg_add_inversion_data_corrected", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, c_size, c_size, ctypes.c_void_p, ctypes.c_void_p, c_size,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_size, c_size, c_size,
      ctypes.c_int, ctypes.c_int]),
    ("wg_add_inversion_data_device", ctypes.c_int,
     [ctypes.c_void_p, c_size, c_size, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_predict_visibilities_device", ctypes.c_int,
     [ctypes.c_void_p, c_size, c_size, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_prepare_visibilities", ctypes.c_int,
     [ctypes.c_void_p, c_size, c_size, ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_add_inversion_prepared", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_predict_prepared", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_autotune", ctypes.c_int,
     [ctypes.c_void_p, c_size, c_size, ctypes.c_void_p, ctypes.c_void_p, c_size,
      ctypes.c_char_p, c_size]),
    ("wg_capture_major_cycle", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int]),
    ("wg_replay_major_cycle", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("wg_plan_json", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, c_size]),
    ("wg_set_timing", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("wg_stage_times", ctypes.c_int, [ctypes.c_void_p, c_dbl_p, ctypes.c_int]),
    ("wg_set_tuning", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, c_size]),
    ("wg_set_precision", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("wg_min_epsilon", ctypes.c_int, [c_dbl_p]),
    ("wgx_sort_keys", ctypes.c_int,
     [ctypes.c_void_p, c_size, c_size, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(c_size), ctypes.c_char_p, c_size]),
    ("wgx_fft", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    ("wgx_radix_sort", ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p]),
    ("wgx_make_plan", ctypes.c_int,
     [ctypes.POINTER(WgParams), ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int,
      ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, c_size]),
    ("wg_facet_batch_create", ctypes.c_int,
     [ctypes.POINTER(WgParams), c_size, ctypes.POINTER(ctypes.c_void_p)]),
    ("wg_facet_batch_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("wg_facet_batch_prepare", ctypes.c_int,
     [ctypes.c_void_p, c_size, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p]),
    ("wg_facet_batch_initialize_inversion", ctypes.c_int, [ctypes.c_void_p]),
    ("wg_facet_batch_add_inversion", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_facet_batch_finalize_image", ctypes.c_int, [ctypes.c_void_p, c_size, ctypes.c_double]),
    ("wg_facet_batch_real_image", ctypes.c_int, [ctypes.c_void_p, c_size, ctypes.c_void_p]),
    ("wg_facet_batch_real_image_device", ctypes.c_int,
     [ctypes.c_void_p, c_size, ctypes.c_void_p]),
    ("wg_facet_batch_initialize_prediction", ctypes.c_int,
     [ctypes.c_void_p, c_size, ctypes.c_void_p]),
    ("wg_facet_batch_initialize_prediction_device", ctypes.c_int,
     [ctypes.c_void_p, c_size, ctypes.c_void_p]),
    ("wg_facet_batch_predict", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_facet_batch_virtual_count", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(c_size)]),
    ("wg_facet_batch_capture_major_cycle", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    ("wg_facet_batch_replay_major_cycle", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("wg_facet_batch_set_timing", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("wg_facet_batch_stage_times", ctypes.c_int, [ctypes.c_void_p, c_dbl_p, ctypes.c_int]),
    ("wg_facet_batch_plan_json", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, c_size]),
    ("wg_image_weights_create", ctypes.c_int,
     [ctypes.c_int, ctypes.c_double, c_size, c_size, ctypes.c_double, ctypes.c_double,
      ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_void_p)]),
    ("wg_image_weights_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("wg_image_weights_size", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(c_size), ctypes.POINTER(c_size)]),
    ("wg_image_weights_grid", ctypes.c_int,
     [ctypes.c_void_p, c_size, c_size, c_size, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p]),
    ("wg_image_weights_finish", ctypes.c_int, [ctypes.c_void_p]),
    ("wg_image_weights_set_min_uv_range", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("wg_image_weights_set_max_uv_range", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("wg_image_weights_set_tukey_taper", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    ("wg_image_weights_set_tukey_inner_taper", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    ("wg_image_weights_set_edge_taper", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("wg_image_weights_set_edge_tukey_taper", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    ("wg_image_weights_set_gaussian_taper", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    ("wg_image_weights_rank_filter", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, c_size]),
    ("wg_image_weights_get_grid", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_image_weights_sample", ctypes.c_int,
     [ctypes.c_void_p, c_size, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_prepare_rows", ctypes.c_int,
     [ctypes.c_void_p, c_size, c_size, c_size, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
      ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_prepare_rows_noalloc", ctypes.c_int,
     [ctypes.c_void_p, c_size, c_size, c_size, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
      ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_part_open", ctypes.c_int,
     [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, c_size,
      ctypes.POINTER(ctypes.c_void_p)]),
    ("wg_part_close", ctypes.c_int, [ctypes.c_void_p]),
    ("wg_part_info", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_char_p,
      c_size]),
    ("wg_part_read", ctypes.c_int,
     [ctypes.c_void_p, c_size, c_size, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_part_grid", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_size, ctypes.c_int, ctypes.c_int,
      ctypes.c_void_p, ctypes.c_void_p]),
    ("wg_resample", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, c_size, c_size, ctypes.c_void_p, c_size, c_size]),
    ("wg_real_image_resampled", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size, c_size]),
    ("wg_initialize_prediction_resampled", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, c_size, c_size]),
    ("wgx_facet_atlas", ctypes.c_int,
     [ctypes.POINTER(WgParams), c_size, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
      c_size]),
    ("wgx_set_max_chunk_vis", ctypes.c_int, [ctypes.c_void_p, c_size]),
    ("wgx_device_bytes", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(c_size)]),
    ("wgx_es_poly", ctypes.c_int,
     [ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    ("wg_server_create", ctypes.c_int,
     [ctypes.c_int, c_size, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("wg_server_submit", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(WgRequest), ctypes.POINTER(ctypes.c_uint64)]),
    ("wg_server_wait", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, c_dbl_p]),
    ("wg_server_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    ("wg_server_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("wgx_synth_visibilities", ctypes.c_int,
     [ctypes.c_void_p, c_size, c_size, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
      ctypes.c_double, ctypes.c_uint64, ctypes.c_void_p]),
]

_lib = None


class WgError(RuntimeError):
    pass


def load():
    """Load libwgrid_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise WgError("libwgrid_hip.so is not built; run python -m wsclean_amd.build")
    # PyTorch ships its own HIP runtime (libamdhip64.so, SONAME
    # libamdhip64.so.7).  Two HIP runtimes in one process do not share the
    # device, so when torch is installed it is loaded first and this
    # library's libamdhip64.so.7 dependency binds to torch's copy; device
    # pointers from torch tensors are then valid here and vice versa.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != 0:
        msg = _lib.wg_last_error()
        raise WgError(msg.decode() if msg else "wgrid error %d" % status)
