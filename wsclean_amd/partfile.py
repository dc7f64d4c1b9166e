"""Reordered measurement-set parts (WSClean's ReorderedMsProvider files,
msproviders/reordering.h:22-60) read through the C-ABI: the meta file of a
spectral window and one part's data, weight and model files.

``ReorderedPart.Read`` returns rows as numpy arrays; ``ReorderedPart.Grid``
streams the whole part into a gridder handle (disk reads overlapped with
the device's weighting and gridding of the previous chunk), as
GridMeasurementSet does over a reordered provider
(wgridder/wgriddingmsgridder.cpp:117-183).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import WgError


def _enc(path):
    return None if path is None else str(path).encode()


class ReorderedPart:
    """One reordered part.  n_pol: visibilities per row and channel in the
    files (1 for a single polarization)."""

    def __init__(self, meta_path, data_path, weight_path=None, model_path=None, n_pol=1):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.wg_part_open(_enc(meta_path), _enc(data_path), _enc(weight_path),
                                          _enc(model_path), int(n_pol), ctypes.byref(h)))
        self._h = h
        info = np.zeros(6, np.uint64)
        t = ctypes.c_double()
        name = ctypes.create_string_buffer(4096)
        _lib.check(self._lib.wg_part_info(self._h, info.ctypes.data_as(ctypes.c_void_p),
                                          ctypes.byref(t), name, 4096))
        (self.n_rows, self.n_chan, self.channel_start, self.data_desc_id, has_model,
         self.n_pol) = (int(x) for x in info)
        self.has_model = bool(has_model)
        self.start_time = t.value
        self.ms_path = name.value.decode()

    def close(self):
        if getattr(self, "_h", None):
            self._lib.wg_part_close(self._h)
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

    def Read(self, row0=0, n=None, model=False):
        """Rows [row0, row0 + n): dict of uvw [n, 3], antennas [n, 2],
        time [n], field [n], vis / model [n, chan, pol] complex64 and
        weights [n, chan, pol] float32."""
        n = self.n_rows - row0 if n is None else int(n)
        if row0 < 0 or n < 0 or row0 + n > self.n_rows:
            raise WgError("rows beyond the part")
        shape = (n, self.n_chan, self.n_pol)
        out = dict(uvw=np.empty((n, 3)), antennas=np.empty((n, 2), np.uint16),
                   time=np.empty(n), field=np.empty(n, np.uint16),
                   vis=np.empty(shape, np.complex64), weights=np.empty(shape, np.float32))
        if model:
            out["model"] = np.empty(shape, np.complex64)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        _lib.check(self._lib.wg_part_read(self._h, int(row0), n, p(out["uvw"]),
                                          p(out["antennas"]), p(out["time"]), p(out["field"]),
                                          p(out["vis"]), p(out["weights"]),
                                          p(out["model"]) if model else None))
        return out

    def Grid(self, gridder, freq, chunk_rows=0, weighting=0, polarization=1,
             image_weights=None):
        """Adds the whole part to the gridder's image; returns the weight
        sums (as weighting.prepare_rows)."""
        freq = np.ascontiguousarray(freq, np.float64)
        if freq.size != self.n_chan:
            raise WgError("one frequency per channel of the part")
        stats = np.zeros(4)
        _lib.check(self._lib.wg_part_grid(
            self._h, gridder._h, freq.ctypes.data_as(ctypes.c_void_p), int(chunk_rows),
            int(weighting), int(polarization),
            image_weights.handle if image_weights is not None else None,
            stats.ctypes.data_as(ctypes.c_void_p)))
        return dict(total_weight=float(stats[0]), visibility_weight_sum=float(stats[1]),
                    max_gridded_weight=float(stats[2]), gridded_visibility_count=int(stats[3]))
