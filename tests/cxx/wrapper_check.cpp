// Drives the C++ wrapper include/wgrid_hip.hpp as WSClean's adapter class
// would: one inversion, one corrected inversion and one prediction of a
// problem read from a file, results written to another file, and a failed
// call surfacing as std::runtime_error.
//
// Input (little-endian): u64 n, n_rows, n_chan, n_ant, n_times; f64 dx, eps;
// uvw [n_rows][3] f64; freq [n_chan] f64; vis [n_rows][n_chan] complex<f32>;
// model [n][n] f32; antennas [n_rows][2] u32; time_offsets [n_rows] u64;
// gains [n_times][n_chan][n_ant][2] complex<f32>.
// Output: image [n][n] f32, corrected image [n][n] f32, predicted
// [n_rows][n_chan] complex<f32>, u64 memory usage (constant, per vis).
#include <complex>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <vector>

#include "wgrid_hip.hpp"

template <typename T>
static void Read(std::ifstream& f, T* p, size_t n) {
  f.read(reinterpret_cast<char*>(p), std::streamsize(n * sizeof(T)));
  if (!f) throw std::runtime_error("short input file");
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: wrapper_check input output\n");
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    uint64_t hdr[5];
    double prm[2];
    Read(in, hdr, 5);
    Read(in, prm, 2);
    const size_t n = hdr[0], n_rows = hdr[1], n_chan = hdr[2], n_ant = hdr[3], n_times = hdr[4];
    std::vector<double> uvw(n_rows * 3), freq(n_chan);
    std::vector<std::complex<float>> vis(n_rows * n_chan), gains(n_times * n_chan * n_ant * 2);
    std::vector<float> model(n * n);
    std::vector<uint32_t> ant(n_rows * 2);
    std::vector<uint64_t> toff(n_rows);
    Read(in, uvw.data(), uvw.size());
    Read(in, freq.data(), freq.size());
    Read(in, vis.data(), vis.size());
    Read(in, model.data(), model.size());
    Read(in, ant.data(), ant.size());
    Read(in, toff.data(), toff.size());
    Read(in, gains.data(), gains.size());

    wgrid_hip::Gridder g(n, n, n, n, prm[0], prm[0], 0.0, 0.0, 1, prm[1]);
    g.InitializeInversion();
    g.AddInversionData(n_rows, n_chan, uvw.data(), freq.data(), vis.data());
    g.FinalizeImage(1.0);
    const std::vector<float> img = g.RealImage();

    // diagonal gains (GainMode index 3 needs two polarizations; xx on one)
    g.InitializeInversion();
    g.AddInversionDataCorrected(0, 1, n_rows, uvw.data(), freq.data(), n_chan, ant.data(),
                                vis.data(), toff.data(), gains.data(), n_times, n_ant, 2, false,
                                WG_POL_XX);
    g.FinalizeImage(1.0);
    const std::vector<float> img_c = g.RealImage();

    g.InitializePrediction(model.data());
    std::vector<std::complex<float>> out(n_rows * n_chan);
    g.PredictVisibilities(n_rows, n_chan, uvw.data(), freq.data(), out.data());
    const uint64_t mem[2] = {g.ConstantMemoryUsage(), g.PerVisibilityMemoryUsage()};

    bool threw = false;
    try {
      g.AddInversionData(n_rows, n_chan, nullptr, freq.data(), vis.data());
    } catch (const std::runtime_error&) {
      threw = true;
    }
    if (!threw) {
      std::fprintf(stderr, "a null uvw pointer did not throw\n");
      return 3;
    }
    std::ofstream o(argv[2], std::ios::binary);
    o.write(reinterpret_cast<const char*>(img.data()), std::streamsize(img.size() * 4));
    o.write(reinterpret_cast<const char*>(img_c.data()), std::streamsize(img_c.size() * 4));
    o.write(reinterpret_cast<const char*>(out.data()), std::streamsize(out.size() * 8));
    o.write(reinterpret_cast<const char*>(mem), 16);
    std::printf("wrapper_check: ok\n");
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "wrapper_check: %s\n", e.what());
    return 1;
  }
}
