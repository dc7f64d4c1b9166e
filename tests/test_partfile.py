"""Reordered measurement-set part files (msproviders/reordering.h:22-60)
written here with the reference's packed layouts (MetaHeader 20 bytes,
MetaRecord 38 bytes, PartHeader 21 bytes; data, weight and model rows of
channel_count x n_pol values, msproviders/reorderedfilewriter.cpp:100-125,
msproviders/reorderedmsprovider.cpp:88-93) and read back through the
library's C-ABI reader, value for value.  No GPU."""
import os
import struct

import numpy as np
import pytest

from wsclean_amd import WgError
from wsclean_amd.partfile import ReorderedPart


def write_part(directory, ms_path="/data/obs.ms", n_rows=500, n_chan=7, n_pol=1,
               channel_start=3, data_desc_id=2, start_time=5.0e9, with_weights=True,
               with_model=False, seed=0, uvw=None, vis=None):
    """Writes the meta file of spectral window data_desc_id and part 0 of
    it (file names as GetMetaFilename / GetPartPrefix with a polarization
    name "I", msproviders/reordering.cpp:35-57); returns the paths and the
    values written (uvw [rows, 3] and vis [rows, chan, pol] may be given)."""
    rng = np.random.default_rng(seed)
    prefix = os.path.join(str(directory), "obs.ms")
    meta = prefix + "-spw%d-parted-meta.tmp" % data_desc_id
    part = prefix + "-part0000-I-b%d" % data_desc_id
    uvw = rng.normal(0, 2000.0, (n_rows, 3)) if uvw is None else np.asarray(uvw, np.float64)
    time = start_time + 10.0 * np.arange(n_rows)
    ant = rng.integers(0, 64, (n_rows, 2)).astype(np.uint16)
    field = rng.integers(0, 3, n_rows).astype(np.uint16)
    if vis is None:
        vis = rng.normal(size=(n_rows, n_chan, n_pol)) + 1j * rng.normal(size=(n_rows, n_chan,
                                                                             n_pol))
    vis = np.ascontiguousarray(vis, np.complex64)
    wts = rng.uniform(0.1, 2.0, (n_rows, n_chan, n_pol)).astype(np.float32)
    model = (0.1 * vis).astype(np.complex64)
    name = ms_path.encode()
    with open(meta, "wb") as f:
        f.write(struct.pack("<dQI", start_time, n_rows, len(name)))
        f.write(name)
        for r in range(n_rows):
            f.write(struct.pack("<ddddHHH", uvw[r, 0], uvw[r, 1], uvw[r, 2], time[r],
                                ant[r, 0], ant[r, 1], field[r]))
    with open(part + ".tmp", "wb") as f:
        f.write(struct.pack("<QQI?", n_chan, channel_start, data_desc_id, with_model))
        f.write(vis.tobytes())
    paths = dict(meta=meta, data=part + ".tmp", weights=None, model=None)
    if with_weights:
        paths["weights"] = part + "-w.tmp"
        with open(paths["weights"], "wb") as f:
            f.write(wts.tobytes())
    if with_model:
        paths["model"] = part + "-m.tmp"
        with open(paths["model"], "wb") as f:
            f.write(model.tobytes())
    values = dict(uvw=uvw, time=time, antennas=ant, field=field, vis=vis, weights=wts,
                  model=model)
    return paths, values


def test_struct_sizes_match_the_reference():
    # reordering.h:63, :68, :101 static_asserts
    assert struct.calcsize("<dQI") == 20
    assert struct.calcsize("<ddddHHH") == 38
    assert struct.calcsize("<QQI?") == 21


@pytest.mark.parametrize("n_pol,with_weights,with_model", [(1, True, False), (4, True, True),
                                                           (2, False, False)])
def test_write_and_read_back(tmp_path, n_pol, with_weights, with_model):
    paths, v = write_part(tmp_path, n_pol=n_pol, with_weights=with_weights,
                          with_model=with_model, seed=n_pol)
    with ReorderedPart(paths["meta"], paths["data"], paths["weights"], paths["model"],
                       n_pol=n_pol) as p:
        assert (p.n_rows, p.n_chan, p.channel_start, p.data_desc_id, p.n_pol) == (500, 7, 3, 2,
                                                                               n_pol)
        assert p.has_model == with_model
        assert p.start_time == 5.0e9 and p.ms_path == "/data/obs.ms"
        got = p.Read(model=with_model)
        np.testing.assert_array_equal(got["uvw"], v["uvw"])
        np.testing.assert_array_equal(got["time"], v["time"])
        np.testing.assert_array_equal(got["antennas"], v["antennas"])
        np.testing.assert_array_equal(got["field"], v["field"])
        np.testing.assert_array_equal(got["vis"], v["vis"])
        want_w = v["weights"] if with_weights else np.ones_like(v["weights"])
        np.testing.assert_array_equal(got["weights"], want_w)
        if with_model:
            np.testing.assert_array_equal(got["model"], v["model"])
        # a row range in the middle
        sub = p.Read(123, 77)
        np.testing.assert_array_equal(sub["uvw"], v["uvw"][123:200])
        np.testing.assert_array_equal(sub["vis"], v["vis"][123:200])
        np.testing.assert_array_equal(sub["weights"], want_w[123:200])
        with pytest.raises(WgError):
            p.Read(450, 51)


def test_truncated_or_missing_files_are_errors(tmp_path):
    paths, _ = write_part(tmp_path, n_rows=50)
    with open(paths["data"], "r+b") as f:
        f.truncate(21 + 49 * 7 * 8)          # one row short
    with pytest.raises(WgError, match="data file"):
        ReorderedPart(paths["meta"], paths["data"], paths["weights"])
    with pytest.raises(WgError, match="cannot open"):
        ReorderedPart(paths["meta"], str(tmp_path / "none.tmp"))
