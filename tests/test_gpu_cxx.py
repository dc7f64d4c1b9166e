"""The C++ wrapper include/wgrid_hip.hpp run as WSClean's adapter would run
it (tests/cxx/wrapper_check.cpp, built with the library): its inversion,
corrected inversion and prediction equal the Python mirror's on the same
inputs (both call the same C-ABI), the inversion meets epsilon against the
FP64 direct FT, and a failed call surfaces as std::runtime_error."""
import os
import subprocess

import numpy as np
import pytest

from wsclean_amd import WGriddingGridder_Simple, synth
from wsclean_amd.gridder import POL_XX

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cxx", "wrapper_check")


def test_cpp_wrapper_matches_python_mirror(oracle, tmp_path):
    assert os.path.exists(EXE), "build the library first (__graft_entry__.build)"
    p = synth.small_problem(seed=55, n_ant=16, n_times=20, n_chan=3, npix=128, noise=0.02)
    n, eps = 128, 1e-5
    n_rows, n_chan = p["uvw"].shape[0], len(p["freq"])
    rng = np.random.default_rng(55)
    n_ant, n_times = 16, 4
    ant = np.sort(rng.integers(0, n_ant, (n_rows, 2)), axis=1).astype(np.uint32)
    toff = rng.integers(0, n_times, n_rows).astype(np.uint64)
    gains = (rng.uniform(0.8, 1.2, (n_times, n_chan, n_ant, 2))
             * np.exp(1j * rng.uniform(-0.3, 0.3, (n_times, n_chan, n_ant, 2)))).astype(np.complex64)
    model = np.zeros((n, n), np.float32)
    model[rng.integers(4, n - 4, 10), rng.integers(4, n - 4, 10)] = rng.uniform(-1, 1, 10)
    vis = np.ascontiguousarray(p["vis"], np.complex64)
    src = tmp_path / "in.bin"
    with open(src, "wb") as f:
        f.write(np.array([n, n_rows, n_chan, n_ant, n_times], np.uint64).tobytes())
        f.write(np.array([p["dx"], eps], np.float64).tobytes())
        for a in (p["uvw"].astype(np.float64), p["freq"].astype(np.float64), vis, model, ant, toff,
                  gains):
            f.write(np.ascontiguousarray(a).tobytes())
    dst = tmp_path / "out.bin"
    print("running", EXE, flush=True)
    r = subprocess.run([EXE, str(src), str(dst)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = open(dst, "rb").read()
    img = np.frombuffer(raw, np.float32, n * n, 0).reshape(n, n)
    img_c = np.frombuffer(raw, np.float32, n * n, 4 * n * n).reshape(n, n)
    out = np.frombuffer(raw, np.complex64, n_rows * n_chan, 8 * n * n).reshape(n_rows, n_chan)
    mem = np.frombuffer(raw, np.uint64, 2, 8 * n * n + 8 * n_rows * n_chan)

    g = WGriddingGridder_Simple(n, n, n, n, p["dx"], p["dx"], 0.0, 0.0, 1, eps)
    g.InitializeInversion()
    g.AddInversionData(n_rows, n_chan, p["uvw"], p["freq"], vis)
    g.FinalizeImage(1.0)
    np.testing.assert_array_equal(img, g.RealImage().reshape(n, n))
    g.InitializeInversion()
    g.AddInversionDataWithCorrectionCallback(0, 1, n_rows, p["uvw"], p["freq"], n_chan, ant,
                                             vis, toff, gains, False, POL_XX)
    g.FinalizeImage(1.0)
    np.testing.assert_array_equal(img_c, g.RealImage().reshape(n, n))
    g.InitializePrediction(model.ravel())
    ref_out = np.zeros((n_rows, n_chan), np.complex64)
    g.PredictVisibilities(n_rows, n_chan, p["uvw"], p["freq"], ref_out)
    np.testing.assert_array_equal(out, ref_out)
    assert int(mem[0]) == g.ConstantMemoryUsage() and int(mem[1]) == g.PerVisibilityMemoryUsage()
    g.close()
    ref = oracle.dft_invert(n, n, p["dx"], p["dx"], 0.0, 0.0, p["uvw"], p["freq"], vis)
    e = float(np.linalg.norm(img - ref) / np.linalg.norm(ref))
    print("C++ wrapper invert vs direct FT %.3e" % e, flush=True)
    assert e < eps
