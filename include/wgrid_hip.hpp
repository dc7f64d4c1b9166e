// wgrid_hip.hpp -- header-only C++ wrapper over the C-ABI of wgrid_hip.h.
//
// wgrid_hip::Gridder has the member functions of WSClean's
// WGriddingGridderBase (wgridder/gridder_simple.h:20-42) with the same
// argument meaning and layouts, and rethrows a failed call as
// std::runtime_error, as the reference's gridder throws (ducc0 MR_assert,
// wgridder/gridder_simple_implementation.h:49).  It depends on nothing but
// the C-ABI, so a WSClean-side adapter class can forward to it one call per
// member.
#ifndef WGRID_HIP_HPP_
#define WGRID_HIP_HPP_

#include <complex>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "wgrid_hip.h"

namespace wgrid_hip {

inline void Check(int status) {
  if (status != 0) throw std::runtime_error(std::string("wgrid_hip: ") + wg_last_error());
}

// Smallest epsilon the library accepts (wg_min_epsilon).
inline double MinEpsilon() {
  double e = 0.0;
  Check(wg_min_epsilon(&e));
  return e;
}

class Gridder {
 public:
  // Arguments of the WGriddingGridder_Simple constructor
  // (wgridder/gridder_simple.h:84-88), in the same order.
  Gridder(size_t width, size_t height, size_t width_t, size_t height_t, double pixel_size_x,
          double pixel_size_y, double l_shift, double m_shift, size_t nthreads,
          double epsilon = 1e-4, size_t verbosity = 0, bool tuning = false)
      : width_(width), height_(height) {
    wg_params p{};
    p.width = width;
    p.height = height;
    p.width_t = width_t;
    p.height_t = height_t;
    p.pixel_size_x = pixel_size_x;
    p.pixel_size_y = pixel_size_y;
    p.l_shift = l_shift;
    p.m_shift = m_shift;
    p.nthreads = nthreads;
    p.epsilon = epsilon;
    p.verbosity = verbosity;
    p.tuning = tuning ? 1 : 0;
    Check(wg_create(&p, &h_));
  }
  ~Gridder() {
    if (h_) wg_destroy(h_);
  }
  Gridder(const Gridder&) = delete;
  Gridder& operator=(const Gridder&) = delete;

  size_t ConstantMemoryUsage() const {
    size_t b = 0;
    Check(wg_constant_memory_usage(h_, &b));
    return b;
  }
  size_t PerVisibilityMemoryUsage() const {
    size_t b = 0;
    Check(wg_per_visibility_memory_usage(h_, &b));
    return b;
  }
  void InitializeInversion() { Check(wg_initialize_inversion(h_)); }
  void AddInversionData(size_t n_rows, size_t n_chan, const double* uvw, const double* freq,
                        const std::complex<float>* vis) {
    Check(wg_add_inversion_data(h_, n_rows, n_chan, uvw, freq,
                                reinterpret_cast<const float*>(vis)));
  }
  void FinalizeImage(double multiplication_factor) {
    Check(wg_finalize_image(h_, multiplication_factor));
  }
  std::vector<float> RealImage() {
    std::vector<float> out(width_ * height_);
    Check(wg_real_image(h_, out.data()));
    return out;
  }
  void InitializePrediction(const float* image_data) {
    Check(wg_initialize_prediction(h_, image_data));
  }
  // const in the reference (gridder_simple.h:39-41); the handle serialises
  // calls internally.
  void PredictVisibilities(size_t n_rows, size_t n_chan, const double* uvw, const double* freq,
                           std::complex<float>* vis) const {
    Check(wg_predict_visibilities(h_, n_rows, n_chan, uvw, freq, reinterpret_cast<float*>(vis)));
  }

  // AddInversionDataWithCorrectionCallback with the solutions passed
  // explicitly (wg_add_inversion_data_corrected): gains [n_times][n_chan]
  // [n_antennas][nparms], antennas [n_rows][2], vis [n_rows][n_chan][n_pol].
  void AddInversionDataCorrected(int gain_mode, size_t n_pol, size_t n_rows, const double* uvw,
                                 const double* freq, size_t n_chan, const uint32_t* antennas,
                                 const std::complex<float>* vis, const uint64_t* time_offsets,
                                 const std::complex<float>* gains, size_t n_times,
                                 size_t n_antennas, size_t nparms, bool apply_forward,
                                 int polarization) {
    Check(wg_add_inversion_data_corrected(
        h_, gain_mode, n_pol, n_rows, uvw, freq, n_chan, antennas,
        reinterpret_cast<const float*>(vis), time_offsets, reinterpret_cast<const float*>(gains),
        n_times, n_antennas, nparms, apply_forward ? 1 : 0, polarization));
  }

  wg_handle* Handle() const { return h_; }

 private:
  wg_handle* h_ = nullptr;
  size_t width_, height_;
};

// The library's served request stream (wg_server_*): Submit queues a
// request and returns its id, Wait blocks until it has ended (rethrowing its
// failure) and gives its submission, start and end times in seconds since
// the server started; the destructor waits for every submitted request.
class Server {
 public:
  explicit Server(int slots = 4, size_t reserve_bytes = size_t(1) << 30, int max_idle = 8) {
    Check(wg_server_create(slots, reserve_bytes, max_idle, &s_));
  }
  ~Server() {
    if (s_) wg_server_destroy(s_);
  }
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  uint64_t Submit(const wg_request& r) {
    uint64_t id = 0;
    Check(wg_server_submit(s_, &r, &id));
    return id;
  }
  struct Times {
    double submitted, started, ended;
  };
  Times Wait(uint64_t id) {
    double t[3] = {0, 0, 0};
    Check(wg_server_wait(s_, id, t));
    return Times{t[0], t[1], t[2]};
  }
  // completed, peak running, pool hits, pool misses, handles closed, least
  // free device bytes at an admission, waited admissions, idle handles
  std::vector<uint64_t> Stats() {
    std::vector<uint64_t> v(8);
    Check(wg_server_stats(s_, v.data()));
    return v;
  }

 private:
  wg_server* s_ = nullptr;
};

}  // namespace wgrid_hip

#endif  // WGRID_HIP_HPP_
