// C-ABI of the MI355X w-gridder (declared in include/wgrid_hip.h).  The
// handle mirrors WGriddingGridder_Simple<NumT>: it owns the trimmed image
// (fp64, x-major index x * height_t + y as in
// wgridder/gridder_simple_implementation.h:147-167) in HBM between calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <algorithm>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/wgrid_hip.h"
#include "device.hpp"
#include "engine.hpp"
#include "partfile.hpp"

#include <future>

namespace {

thread_local std::string g_last_error;

constexpr double kSigmaMax = 2.0;  // WGriddingGridder_Simple::sigma_max

__global__ void ScaleKernel(double* __restrict__ p, size_t n, double f) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] *= f;
}

// out[(i + dx) + (j + dy) * width] = img[i * height_t + j], NaN elsewhere.
__global__ void RealImageKernel(const double* __restrict__ img, float* __restrict__ out,
                                int width, int height, int width_t, int height_t) {
  const size_t n = size_t(width) * height;
  const int dx = (width - width_t) / 2, dy = (height - height_t) / 2;
  for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += size_t(gridDim.x) * blockDim.x) {
    const int x = int(k % width), y = int(k / width);
    const int i = x - dx, j = y - dy;
    out[k] = (i >= 0 && i < width_t && j >= 0 && j < height_t)
                 ? float(img[size_t(i) * height_t + j])
                 : __builtin_nanf("");
  }
}

// img[i * height_t + j] = image[(i + dx) + (j + dy) * width]
__global__ void LoadModelKernel(const float* __restrict__ image, double* __restrict__ img,
                                int width, int height, int width_t, int height_t) {
  const size_t n = size_t(width_t) * height_t;
  const int dx = (width - width_t) / 2, dy = (height - height_t) / 2;
  for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += size_t(gridDim.x) * blockDim.x) {
    const int i = int(k / height_t), j = int(k % height_t);
    img[k] = double(image[size_t(i + dx) + size_t(j + dy) * width]);
  }
}

// Device allocation owned for the duration of one entry point: freed on
// every exit path, exceptions included.
struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { WG_HIP(hipMalloc(&p, bytes ? bytes : 16)); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

unsigned Blocks(size_t n) {
  size_t b = (n + 255) / 256;
  return unsigned(b < 65536 ? (b ? b : 1) : 65536);
}

template <typename F>
int Guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 1;
  } catch (...) {
    g_last_error = "unknown error";
    return 1;
  }
}

// Runs `body` (device work on the engine stream, no host synchronisation)
// once eagerly, so that every arena buffer, twiddle table and launch
// attribute it needs exists, then captures it into an executable HIP graph.
// The body clears with fill kernels, never with memset nodes.
template <typename F>
hipGraphExec_t CaptureGraph(wg::Engine& e, F&& body) {
  hipStream_t s = e.Stream();
  body();
  WG_HIP(hipStreamSynchronize(s));
  const bool timing = e.timing;
  e.timing = false;  // no event records inside the graph
  WG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipGraph_t g = nullptr;
  try {
    body();
  } catch (...) {
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    e.timing = timing;
    throw;
  }
  e.timing = timing;
  WG_HIP(hipStreamEndCapture(s, &g));
  hipGraphExec_t exec = nullptr;
  const hipError_t err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  WG_HIP(err);
  return exec;
}

void DropGraph(hipGraphExec_t& g) {
  if (g) (void)hipGraphExecDestroy(g);
  g = nullptr;
}

// Autotuner results of this process ("wisdom"), keyed by the image
// geometry, epsilon, channel count and the visibility count rounded to a
// power of two; handles created with tuning on reuse them.
struct TuneChoice {
  int W;
  double sigma;
  int T;
};
std::mutex g_wisdom_mu;
std::map<std::string, TuneChoice> g_wisdom;

std::string WisdomKey(const wg_params& p, size_t n_rows, size_t n_chan) {
  char key[256];
  const double nv = double(n_rows) * double(n_chan);
  std::snprintf(key, sizeof(key), "%zu %zu %.9g %.9g %.9g %.9g %.3g %zu %d", p.width_t,
                p.height_t, p.pixel_size_x, p.pixel_size_y, p.l_shift, p.m_shift, p.epsilon,
                n_chan, nv > 0 ? int(std::lround(std::log2(nv))) : 0);
  return key;
}

std::string TrialsJson(const std::vector<wg::Engine::TuneTrial>& t) {
  std::string o = "[";
  char buf[512];
  for (size_t i = 0; i < t.size(); ++i) {
    std::snprintf(buf, sizeof(buf),
                  "%s{\"W\":%d,\"sigma\":%.4g,\"T\":%d,\"Sg\":%d,\"P\":%d,\"nu\":%d,"
                  "\"est_ms\":%.6g,\"vis_ms\":%.6g,\"fixed_ms\":%.6g,\"ok\":%s}",
                  i ? "," : "", t[i].W, t[i].sigma, t[i].T, t[i].Sg, t[i].P, t[i].nu,
                  t[i].est_ms < 1e300 ? t[i].est_ms : -1.0, t[i].vis_ms, t[i].fixed_ms,
                  t[i].error.empty() ? "true" : "false");
    o += buf;
  }
  return o + "]";
}

// Page-locked staging of large host transfers.  A copy from or to pageable
// memory runs at about half the PCIe rate (30 GB/s to and 18 GB/s from the
// device, against 57 GB/s both ways from page-locked memory, on an MI355X
// host), so each 64 MiB chunk goes through one of two page-locked buffers:
// several host threads copy a chunk between the caller's memory and one
// buffer while the DMA engine moves the other.  Transfers below 4 MiB go
// directly.  One stager serves one handle (its calls hold the handle lock).
class HostStager {
 public:
  HostStager() = default;
  HostStager(const HostStager&) = delete;
  HostStager& operator=(const HostStager&) = delete;
  ~HostStager() {
    for (int b = 0; b < kBufs; ++b) {
      if (ev_[b]) {
        if (used_[b]) (void)hipEventSynchronize(ev_[b]);
        (void)hipEventDestroy(ev_[b]);
      }
      if (buf_[b]) (void)hipHostFree(buf_[b]);
    }
  }

  // dst (device) = src (host); the copy is ordered on stream s, and src may
  // be reused when this returns
  void ToDevice(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes < kDirect) {
      WG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
      return;
    }
    Init();
    size_t c = 0;
    for (size_t off = 0; off < bytes; off += kChunk, ++c) {
      const int b = int(c % kBufs);
      const size_t n = std::min(kChunk, bytes - off);
      // the buffer's previous transfer has left it
      if (used_[b]) WG_HIP(hipEventSynchronize(ev_[b]));
      ParallelCopy(buf_[b], static_cast<const char*>(src) + off, n);
      WG_HIP(hipMemcpyAsync(static_cast<char*>(dst) + off, buf_[b], n, hipMemcpyHostToDevice, s));
      WG_HIP(hipEventRecord(ev_[b], s));
      used_[b] = true;
    }
  }

  // dst (host) = src (device) after the work already on stream s; complete
  // when this returns
  void ToHost(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes < kDirect) {
      WG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
      WG_HIP(hipStreamSynchronize(s));
      return;
    }
    Init();
    int pend = -1;
    size_t pend_off = 0, pend_n = 0, c = 0;
    auto drain = [&]() {
      WG_HIP(hipEventSynchronize(ev_[pend]));
      ParallelCopy(static_cast<char*>(dst) + pend_off, buf_[pend], pend_n);
    };
    for (size_t off = 0; off < bytes; off += kChunk, ++c) {
      const int b = int(c % kBufs);
      const size_t n = std::min(kChunk, bytes - off);
      WG_HIP(hipMemcpyAsync(buf_[b], static_cast<const char*>(src) + off, n,
                            hipMemcpyDeviceToHost, s));
      WG_HIP(hipEventRecord(ev_[b], s));
      used_[b] = true;
      // the chunk before this one goes to the caller while this one moves
      if (pend >= 0) drain();
      pend = b;
      pend_off = off;
      pend_n = n;
    }
    if (pend >= 0) drain();
  }

 private:
  static constexpr size_t kChunk = size_t(64) << 20;
  static constexpr size_t kDirect = size_t(4) << 20;
  static constexpr int kBufs = 2;

  void Init() {
    for (int b = 0; b < kBufs; ++b) {
      if (!buf_[b]) WG_HIP(hipHostMalloc(&buf_[b], kChunk));
      if (!ev_[b]) WG_HIP(hipEventCreateWithFlags(&ev_[b], hipEventDisableTiming));
    }
  }

  // memcpy of n bytes with up to kThreads host threads (one chunk)
  static void ParallelCopy(void* dst, const void* src, size_t n) {
    const unsigned hc = std::thread::hardware_concurrency();
    const int T = int(std::max(1u, std::min(kThreads, hc ? hc / 2 : 1u)));
    const size_t per = ((n + T - 1) / T + 4095) & ~size_t(4095);
    std::vector<std::thread> th;
    size_t done = per;  // bytes [0, per) copied by this thread
    try {
      for (int t = 1; t < T && size_t(t) * per < n; ++t) {
        const size_t a = size_t(t) * per, m = std::min(per, n - a);
        th.emplace_back([=] {
          std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, m);
        });
        done = a + m;
      }
    } catch (...) {
      // no more threads: this thread copies what none was started for
    }
    std::memcpy(dst, src, std::min(per, n));
    if (done < n)
      std::memcpy(static_cast<char*>(dst) + done, static_cast<const char*>(src) + done, n - done);
    for (auto& x : th) x.join();
  }

  static constexpr unsigned kThreads = 8;
  void* buf_[kBufs] = {};
  hipEvent_t ev_[kBufs] = {};
  bool used_[kBufs] = {};  // a transfer was recorded on the buffer's event
};

}  // namespace

struct wg_handle {
  wg_params prm{};
  std::unique_ptr<wg::Engine> eng;
  double* img = nullptr;  // [width_t][height_t] fp64
  wg::VisSet vs_call;
  wg::VisSet vs_prep;
  bool prepared = false;
  std::string plan_json;
  hipGraphExec_t cycle = nullptr;  // captured major cycle on vs_prep
  uint64_t cycle_gen = 0;          // arena generation the graph's pointers belong to
  int device = 0;                  // device current at creation (the stream's)
  size_t max_chunk_vis = 0;        // test hook: largest engine call (0: automatic)
  HostStager stage;                // host transfers of the host-buffer calls
  std::mutex mu;

  ~wg_handle() { DropGraph(cycle); }
  size_t npix() const { return prm.width_t * prm.height_t; }
  hipStream_t s() const { return eng->Stream(); }
};

struct wg_facet_batch {
  std::vector<wg_params> prm;
  std::unique_ptr<wg::FacetBatch> fb;
  std::vector<double*> img;  // per facet [width_t][height_t] fp64
  std::string plan_json;
  hipGraphExec_t cycle = nullptr;  // captured major cycle of the batch
  uint64_t cycle_gen = 0;          // arena generation the graph's pointers belong to
  std::mutex mu;

  ~wg_facet_batch() { DropGraph(cycle); }

  hipStream_t s() { return fb->Eng().Stream(); }
  size_t npix(size_t f) const { return prm[f].width_t * prm[f].height_t; }
};

struct wg_image_weights {
  int mode = 0;
  double robust = 0;
  bool as_taper = false;
  bool finished = false;
  hipStream_t stream = nullptr;
  wg::DeviceArena arena;
  wg::WeightGrid g{};
  double* total_part = nullptr;  // gridded-weight partials (accumulated)
  double* sums = nullptr;        // total, sum of squares
  std::mutex mu;

  size_t cells() const { return size_t(g.height / 2) * size_t(g.width); }
  ~wg_image_weights() {
    arena.ReleaseAll();
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }
};

// Device scratch of wg_prepare_rows_noalloc, one copy per GPU for the life
// of the process: the statistics partials of the row kernel, the four
// statistics, and the wavelengths and selection flags of at most
// kNoAllocMaxChan channels.  Host code reaches it only through the address
// hipGetSymbolAddress returns on each device.
constexpr size_t kNoAllocMaxChan = 65536;
constexpr size_t kNoAllocPartWords = 4 * 1024;  // == wg::WeightStatPartials()
constexpr size_t kNoAllocStatsOff = kNoAllocPartWords * sizeof(double);
constexpr size_t kNoAllocLambdaOff = kNoAllocStatsOff + 256;
constexpr size_t kNoAllocSelOff = kNoAllocLambdaOff + kNoAllocMaxChan * sizeof(double);
constexpr size_t kNoAllocScratchBytes = kNoAllocSelOff + kNoAllocMaxChan;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu[kNoAllocMaxDevices];
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// lambda[c] = c / freq[c] in place (the FP64 division is correctly rounded,
// as the host's)
__global__ void WavelengthsKernel(double* __restrict__ lam, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    lam[i] = wg::kSpeedOfLight / lam[i];
}

}  // namespace


struct wg_part_reader {
  std::unique_ptr<wg::PartReader> r;
};

namespace {

// Page-locked host memory owned for the duration of one entry point.
struct PinnedBuf {
  void* p = nullptr;
  explicit PinnedBuf(size_t bytes) { WG_HIP(hipHostMalloc(&p, bytes ? bytes : 16)); }
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

}  // namespace

namespace wg {
// The thread-local message of wg_last_error, for entry points defined in
// other files of the library (server.hip).
void SetLastError(const std::string& msg) { g_last_error = msg; }
}  // namespace wg

extern "C" {

const char* wg_last_error(void) { return g_last_error.c_str(); }

int wg_create(const wg_params* params, wg_handle** out) {
  return Guard([&] {
    WG_CHECK(params && out, "null argument");
    WG_CHECK(params->verbosity <= 2, "verbosity must be 0, 1, or 2");
    WG_CHECK(params->width_t <= params->width && params->height_t <= params->height,
             "trimmed size exceeds image size");
    WG_CHECK(params->width_t > 0 && params->height_t > 0, "empty image");
    auto h = std::make_unique<wg_handle>();
    h->prm = *params;
    wg::ImageGeometry g;
    g.nx = int(params->width_t);
    g.ny = int(params->height_t);
    g.dx = params->pixel_size_x;
    g.dy = params->pixel_size_y;
    g.l_shift = params->l_shift;
    g.m_shift = params->m_shift;
    g.epsilon = params->epsilon;
    wg::PlanOptions opt;
    WG_HIP(hipGetDevice(&h->device));
    h->eng = std::make_unique<wg::Engine>(g, opt);
    h->img = static_cast<double*>(h->eng->Arena().Get("image.acc", h->npix() * sizeof(double)));
    WG_HIP(hipMemsetAsync(h->img, 0, h->npix() * sizeof(double), h->s()));
    WG_HIP(hipStreamSynchronize(h->s()));
    *out = h.release();
  });
}

int wg_destroy(wg_handle* h) {
  return Guard([&] {
    if (!h) return;
    delete h;  // the engine releases the image with its arena
  });
}

// Device memory of the gridder.  Per image pixel: the FP64 image, the FP32
// dirty accumulator and correction, and the host-transfer images; plus the
// FFT slab (w planes of the occupied uv box and two line buffers, budgeted
// by the plan at slab_bytes) and tables.  Per visibility of one engine call:
// sort keys and permutation with their ping-pong copies, sorted coordinates
// and phasors, each sorted visibility's renumbered row, sorted or
// accumulated visibilities and the staging of host calls, grown by the
// arena's 1/8 rounding; per row the staged and renumbered uvw, the row sort
// and the row maxima.
constexpr size_t kDevBytesPerPixel = 28;
constexpr size_t kDevBytesTables = size_t(64) << 20;
constexpr size_t kDevBytesPerVis = 72;
constexpr size_t kDevBytesPerRow = 96;

int wg_constant_memory_usage(const wg_handle* h, size_t* bytes) {
  return Guard([&] {
    WG_CHECK(h && bytes, "null argument");
    // The reference counts its padded grid and image on the host
    // (gridder_simple_implementation.h:53-60); this library's constant
    // footprint is in device memory.
    const size_t npix = h->prm.width_t * h->prm.height_t;
    *bytes = kDevBytesPerPixel * npix + h->eng->Options().slab_bytes + kDevBytesTables;
  });
}

int wg_per_visibility_memory_usage(const wg_handle* h, size_t* bytes) {
  return Guard([&] {
    WG_CHECK(h && bytes, "null argument");
    // device bytes per visibility of one call (the reference: 8 host bytes,
    // gridder_simple_implementation.h:62-68); calls larger than the device
    // holds are split by the library
    *bytes = kDevBytesPerVis;
  });
}

int wg_initialize_inversion(wg_handle* h) {
  return Guard([&] {
    WG_CHECK(h, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    WG_HIP(hipMemsetAsync(h->img, 0, h->npix() * sizeof(double), h->s()));
    WG_HIP(hipStreamSynchronize(h->s()));
  });
}

// The tuned variant (wg_params.tuning): the autotuner's (W, sigma, tile)
// for this geometry and visibility count, measured once per process.
static void ApplyTuning(wg_handle* h, size_t n_rows, size_t n_chan, const double* uvw_dev,
                        const double* freq) {
  if (!h->prm.tuning || n_rows == 0) return;
  const std::string key = WisdomKey(h->prm, n_rows, n_chan);
  TuneChoice c{0, 0, 0};
  {
    std::lock_guard<std::mutex> wl(g_wisdom_mu);
    auto it = g_wisdom.find(key);
    if (it != g_wisdom.end()) c = it->second;
  }
  if (c.W == 0) {
    const auto trials = h->eng->Autotune(n_rows, n_chan, uvw_dev, freq, size_t(1) << 27);
    c = TuneChoice{trials.front().W, trials.front().sigma, trials.front().T};
    std::lock_guard<std::mutex> wl(g_wisdom_mu);
    g_wisdom[key] = c;
  }
  wg::PlanOptions& o = h->eng->Options();
  o.force_W = c.W;
  o.force_sigma = c.sigma;
  o.T = c.T;
}

// Rows per engine call: at most 2^31 - 1 visibilities (the index range of
// the sorted permutation) and what 80 % of the free device memory (plus what
// this handle's arena already holds, which is reused) takes at
// kDevBytesPerVis / kDevBytesPerRow; wgx_set_max_chunk_vis lowers it.
static size_t ChunkRows(wg_handle* h, size_t n_rows, size_t n_chan) {
  size_t cap = size_t(0x7FFFFFFF) / n_chan;
  if (h->max_chunk_vis) cap = std::min(cap, std::max<size_t>(1, h->max_chunk_vis / n_chan));
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
    const double budget = 0.8 * double(free_b + h->eng->Arena().Allocated());
    const double per_row = double(kDevBytesPerVis) * double(n_chan) + double(kDevBytesPerRow);
    cap = std::min(cap, std::max<size_t>(1, size_t(budget / per_row)));
  }
  WG_CHECK(cap >= 1, "no rows fit the device");
  return std::min(cap, n_rows);
}

static void AddInversion(wg_handle* h, size_t n_rows, size_t n_chan, const double* uvw_dev,
                         const double* freq, const float* vis_dev) {
  wg::Engine& e = *h->eng;
  ApplyTuning(h, n_rows, n_chan, uvw_dev, freq);
  e.Prepare(h->vs_call, "call", n_rows, n_chan, uvw_dev, freq,
            reinterpret_cast<const float2*>(vis_dev));
  h->plan_json = h->vs_call.plan.describe();
  e.Invert(h->vs_call, reinterpret_cast<const float2*>(vis_dev), h->img);
  WG_HIP(hipStreamSynchronize(h->s()));
}

int wg_add_inversion_data(wg_handle* h, size_t n_rows, size_t n_chan, const double* uvw,
                          const double* freq, const float* vis) {
  return Guard([&] {
    WG_CHECK(h && freq && (n_rows == 0 || (uvw && vis)), "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    if (n_rows == 0 || n_chan == 0) return;
    wg::DeviceArena& a = h->eng->Arena();
    // chunks of rows: the inversion of a chunk adds to the image
    const size_t step = ChunkRows(h, n_rows, n_chan);
    for (size_t r0 = 0; r0 < n_rows; r0 += step) {
      const size_t nr = std::min(step, n_rows - r0);
      double* duvw = static_cast<double*>(a.Get("h2d.uvw", nr * 3 * sizeof(double)));
      float* dvis = static_cast<float*>(a.Get("h2d.vis", nr * n_chan * 8));
      h->stage.ToDevice(duvw, uvw + 3 * r0, nr * 3 * sizeof(double), h->s());
      h->stage.ToDevice(dvis, vis + 2 * r0 * n_chan, nr * n_chan * 8, h->s());
      AddInversion(h, nr, n_chan, duvw, freq, dvis);
    }
  });
}

int wg_add_inversion_data_corrected(wg_handle* h, int gain_mode, size_t n_pol, size_t n_rows,
                                    const double* uvw, const double* freq, size_t n_chan,
                                    const uint32_t* antennas, const float* vis,
                                    const uint64_t* time_offsets, const float* gains,
                                    size_t n_times, size_t n_antennas, size_t nparms,
                                    int apply_forward, int polarization) {
  return Guard([&] {
    WG_CHECK(h && freq && (n_rows == 0 || (uvw && antennas && vis && time_offsets && gains)),
             "null argument");
    WG_CHECK(gain_mode >= 0 && gain_mode <= 4, "unknown gain mode");
    WG_CHECK(wg::ValidPolCode(polarization), "unknown polarization code");
    static const size_t kNPol[5] = {1, 1, 1, 2, 4};
    WG_CHECK(n_pol == kNPol[gain_mode], "polarization count does not match the gain mode");
    WG_CHECK(nparms == 2 || nparms == 4, "solutions have 2 or 4 values per antenna");
    WG_CHECK(n_antennas > 0 && n_times > 0, "empty solution table");
    std::lock_guard<std::mutex> lk(h->mu);
    if (n_rows == 0 || n_chan == 0) return;
    for (size_t r = 0; r < n_rows; ++r) {
      WG_CHECK(antennas[2 * r] < n_antennas && antennas[2 * r + 1] < n_antennas,
               "antenna index out of range");
      WG_CHECK(time_offsets[r] < n_times, "time offset out of range");
    }
    wg::DeviceArena& a = h->eng->Arena();
    const size_t ngain = n_times * n_chan * n_antennas * nparms;
    float2* dgain = static_cast<float2*>(a.Get("h2d.gains", ngain * sizeof(float2)));
    hipStream_t s = h->s();
    WG_HIP(hipMemcpyAsync(dgain, gains, ngain * sizeof(float2), hipMemcpyHostToDevice, s));
    // chunks of rows (the polarized input takes n_pol visibilities per row
    // and channel)
    const size_t step = std::max<size_t>(1, ChunkRows(h, n_rows, n_chan * (n_pol + 1)));
    for (size_t r0 = 0; r0 < n_rows; r0 += step) {
      const size_t nr = std::min(step, n_rows - r0);
      const size_t nvis = nr * n_chan;
      double* duvw = static_cast<double*>(a.Get("h2d.uvw", nr * 3 * sizeof(double)));
      float2* dpol = static_cast<float2*>(a.Get("h2d.polvis", nvis * n_pol * sizeof(float2)));
      uint32_t* dant = static_cast<uint32_t*>(a.Get("h2d.ant", nr * 2 * sizeof(uint32_t)));
      uint64_t* dtoff = static_cast<uint64_t*>(a.Get("h2d.toff", nr * sizeof(uint64_t)));
      float2* dvis = static_cast<float2*>(a.Get("h2d.vis", nvis * sizeof(float2)));
      h->stage.ToDevice(duvw, uvw + 3 * r0, nr * 3 * sizeof(double), s);
      h->stage.ToDevice(dpol, vis + 2 * r0 * n_chan * n_pol, nvis * n_pol * sizeof(float2), s);
      WG_HIP(hipMemcpyAsync(dant, antennas + 2 * r0, nr * 2 * sizeof(uint32_t),
                            hipMemcpyHostToDevice, s));
      WG_HIP(hipMemcpyAsync(dtoff, time_offsets + r0, nr * sizeof(uint64_t),
                            hipMemcpyHostToDevice, s));
      wg::LaunchCorrectCollapse(dpol, dant, dtoff, dgain, nr, n_chan, n_antennas, int(nparms),
                                gain_mode, int(n_pol), apply_forward != 0, polarization, dvis, s);
      AddInversion(h, nr, n_chan, duvw, freq, reinterpret_cast<const float*>(dvis));
    }
  });
}

int wg_add_inversion_data_device(wg_handle* h, size_t n_rows, size_t n_chan,
                                 const double* uvw_dev, const double* freq,
                                 const float* vis_dev) {
  return Guard([&] {
    WG_CHECK(h && freq && (n_rows == 0 || (uvw_dev && vis_dev)), "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    if (n_rows == 0 || n_chan == 0) return;
    const size_t step = ChunkRows(h, n_rows, n_chan);
    for (size_t r0 = 0; r0 < n_rows; r0 += step)
      AddInversion(h, std::min(step, n_rows - r0), n_chan, uvw_dev + 3 * r0, freq,
                   vis_dev + 2 * r0 * n_chan);
  });
}

int wg_finalize_image(wg_handle* h, double f) {
  return Guard([&] {
    WG_CHECK(h, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    ScaleKernel<<<Blocks(h->npix()), 256, 0, h->s()>>>(h->img, h->npix(), f);
    WG_HIP(hipGetLastError());
    WG_HIP(hipStreamSynchronize(h->s()));
  });
}

// Writes the width*height float image (trimmed image centred, NaN border)
// to device memory `d` on the handle's stream.
static void RealImageTo(wg_handle* h, float* d) {
  const size_t n = h->prm.width * h->prm.height;
  RealImageKernel<<<Blocks(n), 256, 0, h->s()>>>(h->img, d, int(h->prm.width),
                                               int(h->prm.height), int(h->prm.width_t),
                                               int(h->prm.height_t));
  WG_HIP(hipGetLastError());
}

int wg_real_image(wg_handle* h, float* out) {
  return Guard([&] {
    WG_CHECK(h && out, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    const size_t n = h->prm.width * h->prm.height;
    float* d = static_cast<float*>(h->eng->Arena().Get("d2h.image", n * 4));
    RealImageTo(h, d);
    h->stage.ToHost(out, d, n * 4, h->s());
  });
}

int wg_real_image_device(wg_handle* h, float* image_dev) {
  return Guard([&] {
    WG_CHECK(h && image_dev, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    RealImageTo(h, image_dev);
    WG_HIP(hipStreamSynchronize(h->s()));
  });
}

// Loads the centred trimmed region of the width*height float model `d`
// (device memory) into the fp64 image on the handle's stream.
static void LoadModelFrom(wg_handle* h, const float* d) {
  LoadModelKernel<<<Blocks(h->npix()), 256, 0, h->s()>>>(d, h->img, int(h->prm.width),
                                                       int(h->prm.height), int(h->prm.width_t),
                                                       int(h->prm.height_t));
  WG_HIP(hipGetLastError());
}

int wg_initialize_prediction(wg_handle* h, const float* image) {
  return Guard([&] {
    WG_CHECK(h && image, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    const size_t n = h->prm.width * h->prm.height;
    float* d = static_cast<float*>(h->eng->Arena().Get("h2d.image", n * 4));
    h->stage.ToDevice(d, image, n * 4, h->s());
    LoadModelFrom(h, d);
    WG_HIP(hipStreamSynchronize(h->s()));
  });
}

int wg_initialize_prediction_device(wg_handle* h, const float* image_dev) {
  return Guard([&] {
    WG_CHECK(h && image_dev, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    LoadModelFrom(h, image_dev);
    WG_HIP(hipStreamSynchronize(h->s()));
  });
}

static void Predict(wg_handle* h, size_t n_rows, size_t n_chan, const double* uvw_dev,
                    const double* freq, float* out_dev) {
  wg::Engine& e = *h->eng;
  ApplyTuning(h, n_rows, n_chan, uvw_dev, freq);
  e.Prepare(h->vs_call, "call", n_rows, n_chan, uvw_dev, freq, nullptr);
  h->plan_json = h->vs_call.plan.describe();
  e.Predict(h->vs_call, h->img, reinterpret_cast<float2*>(out_dev));
  WG_HIP(hipStreamSynchronize(h->s()));
}

int wg_predict_visibilities(wg_handle* h, size_t n_rows, size_t n_chan, const double* uvw,
                            const double* freq, float* vis_out) {
  return Guard([&] {
    WG_CHECK(h && freq && (n_rows == 0 || (uvw && vis_out)), "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    if (n_rows == 0 || n_chan == 0) return;
    wg::DeviceArena& a = h->eng->Arena();
    const size_t step = ChunkRows(h, n_rows, n_chan);
    for (size_t r0 = 0; r0 < n_rows; r0 += step) {
      const size_t nr = std::min(step, n_rows - r0);
      double* duvw = static_cast<double*>(a.Get("h2d.uvw", nr * 3 * sizeof(double)));
      float* dout = static_cast<float*>(a.Get("h2d.vis", nr * n_chan * 8));
      h->stage.ToDevice(duvw, uvw + 3 * r0, nr * 3 * sizeof(double), h->s());
      Predict(h, nr, n_chan, duvw, freq, dout);
      h->stage.ToHost(vis_out + 2 * r0 * n_chan, dout, nr * n_chan * 8, h->s());
    }
  });
}

int wg_predict_visibilities_device(wg_handle* h, size_t n_rows, size_t n_chan,
                                   const double* uvw_dev, const double* freq,
                                   float* vis_out_dev) {
  return Guard([&] {
    WG_CHECK(h && freq && (n_rows == 0 || (uvw_dev && vis_out_dev)), "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    if (n_rows == 0 || n_chan == 0) return;
    const size_t step = ChunkRows(h, n_rows, n_chan);
    for (size_t r0 = 0; r0 < n_rows; r0 += step)
      Predict(h, std::min(step, n_rows - r0), n_chan, uvw_dev + 3 * r0, freq,
              vis_out_dev + 2 * r0 * n_chan);
  });
}

int wg_prepare_visibilities(wg_handle* h, size_t n_rows, size_t n_chan, const double* uvw_dev,
                            const double* freq) {
  return Guard([&] {
    WG_CHECK(h && freq && uvw_dev, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    h->prepared = false;
    DropGraph(h->cycle);
    ApplyTuning(h, n_rows, n_chan, uvw_dev, freq);
    h->eng->Prepare(h->vs_prep, "prep", n_rows, n_chan, uvw_dev, freq, nullptr);
    h->plan_json = h->vs_prep.plan.describe();
    h->prepared = true;
  });
}

int wg_add_inversion_prepared(wg_handle* h, const float* vis_dev) {
  return Guard([&] {
    WG_CHECK(h && vis_dev, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    WG_CHECK(h->prepared, "wg_prepare_visibilities has not been called");
    h->eng->Invert(h->vs_prep, reinterpret_cast<const float2*>(vis_dev), h->img);
    WG_HIP(hipStreamSynchronize(h->s()));
  });
}

int wg_predict_prepared(wg_handle* h, float* vis_out_dev) {
  return Guard([&] {
    WG_CHECK(h && vis_out_dev, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    WG_CHECK(h->prepared, "wg_prepare_visibilities has not been called");
    h->eng->Predict(h->vs_prep, h->img, reinterpret_cast<float2*>(vis_out_dev));
    WG_HIP(hipStreamSynchronize(h->s()));
  });
}

int wg_autotune(wg_handle* h, size_t n_rows, size_t n_chan, const double* uvw_dev,
                const double* freq, size_t max_sample_vis, char* report, size_t len) {
  return Guard([&] {
    WG_CHECK(h && uvw_dev && freq, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    const auto trials = h->eng->Autotune(n_rows, n_chan, uvw_dev, freq, max_sample_vis);
    {
      std::lock_guard<std::mutex> wl(g_wisdom_mu);
      g_wisdom[WisdomKey(h->prm, n_rows, n_chan)] =
          TuneChoice{trials.front().W, trials.front().sigma, trials.front().T};
    }
    if (report && len) {
      const std::string j = TrialsJson(trials);
      std::strncpy(report, j.c_str(), len - 1);
      report[len - 1] = 0;
    }
  });
}

int wg_capture_major_cycle(wg_handle* h, const float* vis_dev, float* vis_out_dev,
                           double factor, int what) {
  return Guard([&] {
    WG_CHECK(h && (what & 3) != 0, "null handle or nothing to capture");
    WG_CHECK(!(what & 1) || vis_dev, "invert needs vis_dev");
    WG_CHECK(!(what & 2) || vis_out_dev, "predict needs vis_out_dev");
    std::lock_guard<std::mutex> lk(h->mu);
    WG_CHECK(h->prepared, "wg_prepare_visibilities has not been called");
    DropGraph(h->cycle);
    h->cycle = CaptureGraph(*h->eng, [&] {
      if (what & 1) {
        wg::LaunchFillZero(h->img, h->npix() * sizeof(double), h->s());
        h->eng->Invert(h->vs_prep, reinterpret_cast<const float2*>(vis_dev), h->img);
        ScaleKernel<<<Blocks(h->npix()), 256, 0, h->s()>>>(h->img, h->npix(), factor);
        WG_HIP(hipGetLastError());
      }
      if (what & 2)
        h->eng->Predict(h->vs_prep, h->img, reinterpret_cast<float2*>(vis_out_dev));
    });
    h->cycle_gen = h->eng->Arena().Generation();
  });
}

int wg_replay_major_cycle(wg_handle* h, int count) {
  return Guard([&] {
    WG_CHECK(h && count >= 0, "null handle or negative count");
    std::lock_guard<std::mutex> lk(h->mu);
    WG_CHECK(h->cycle, "no captured major cycle (wg_capture_major_cycle)");
    if (h->eng->Arena().Generation() != h->cycle_gen) {
      // a later call on this handle grew or released a buffer the graph
      // addresses: replaying would touch freed device memory
      DropGraph(h->cycle);
      throw wg::Error("a later call reallocated device buffers of the captured major cycle; "
                      "capture it again (wg_capture_major_cycle)");
    }
    for (int i = 0; i < count; ++i) WG_HIP(hipGraphLaunch(h->cycle, h->s()));
    WG_HIP(hipStreamSynchronize(h->s()));
  });
}

int wg_plan_json(const wg_handle* h, char* buf, size_t len) {
  return Guard([&] {
    WG_CHECK(h && buf && len > 0, "null argument");
    std::strncpy(buf, h->plan_json.c_str(), len - 1);
    buf[len - 1] = 0;
  });
}

int wg_set_timing(wg_handle* h, int enabled) {
  return Guard([&] {
    WG_CHECK(h, "null handle");
    h->eng->timing = enabled != 0;
  });
}

int wg_stage_times(wg_handle* h, double* out14, int reset) {
  return Guard([&] {
    WG_CHECK(h && out14, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    // every entry point synchronises its stream before returning, so all
    // recorded events have completed
    h->eng->CollectTimes();
    wg::StageTimes& t = h->eng->Times();
    const double v[14] = {t.plan, t.sort, t.grid, t.fft, t.degrid,
                          t.plan + t.sort + t.grid + t.fft + t.degrid,
                          t.grid_kernel_ms, t.grid_launches, t.grid_bytes,
                          t.degrid_kernel_ms, t.degrid_launches, t.degrid_bytes,
                          t.grid_passes, t.degrid_passes};
    std::copy(v, v + 14, out14);
    if (reset) t = wg::StageTimes{};
  });
}

int wg_set_tuning(wg_handle* h, int W, double sigma, int tile, int sub_slab, size_t slab_bytes) {
  return Guard([&] {
    WG_CHECK(h, "null handle");
    wg::PlanOptions& o = h->eng->Options();
    o.force_W = W;
    o.force_sigma = sigma;
    o.T = tile;
    o.Sg = sub_slab;
    if (slab_bytes) o.slab_bytes = slab_bytes;
  });
}

int wg_min_epsilon(double* epsilon) {
  return Guard([&] {
    WG_CHECK(epsilon, "null argument");
    *epsilon = wg::Fp32EpsilonFloor();
  });
}

int wg_set_precision(wg_handle* h, int split) {
  return Guard([&] {
    WG_CHECK(h, "null handle");
    WG_CHECK(split >= -1 && split <= 1, "precision mode must be -1, 0 or 1");
    h->eng->Options().split = split;
  });
}

int wgx_sort_keys(wg_handle* h, size_t n_rows, size_t n_chan, const double* uvw,
                  const double* freq, const float* vis, uint32_t* keys_out, uint32_t* perm_out,
                  size_t* n_valid, char* plan_json, size_t plan_len) {
  return Guard([&] {
    WG_CHECK(h && uvw && freq && keys_out && perm_out && n_valid, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    wg::Engine& e = *h->eng;
    wg::DeviceArena& a = e.Arena();
    const size_t nvis = n_rows * n_chan;
    double* duvw = static_cast<double*>(a.Get("h2d.uvw", n_rows * 3 * sizeof(double)));
    WG_HIP(hipMemcpyAsync(duvw, uvw, n_rows * 3 * sizeof(double), hipMemcpyHostToDevice,
                          e.Stream()));
    float* dvis = nullptr;
    if (vis) {
      dvis = static_cast<float*>(a.Get("h2d.vis", nvis * 8));
      WG_HIP(hipMemcpyAsync(dvis, vis, nvis * 8, hipMemcpyHostToDevice, e.Stream()));
    }
    e.Prepare(h->vs_call, "call", n_rows, n_chan, duvw, freq, reinterpret_cast<float2*>(dvis));
    // unsorted keys: recompute into a scratch array
    uint32_t* k = static_cast<uint32_t*>(a.Get("x.keys", nvis * 4));
    uint32_t* v = static_cast<uint32_t*>(a.Get("x.vals", nvis * 4));
    wg::LaunchKeys(h->vs_call.dp, duvw, h->vs_call.scale, n_rows, n_chan,
                   reinterpret_cast<float2*>(dvis), k, v, e.Stream());
    WG_HIP(hipMemcpyAsync(keys_out, k, nvis * 4, hipMemcpyDeviceToHost, e.Stream()));
    // the sorted permutation without the flip flags
    uint32_t* pm = static_cast<uint32_t*>(a.Get("x.perm", h->vs_call.n_valid * 4 + 4));
    wg::LaunchMaskBits(h->vs_call.perm, h->vs_call.n_valid, ~wg::kPermFlip, pm, e.Stream());
    WG_HIP(hipMemcpyAsync(perm_out, pm, h->vs_call.n_valid * 4, hipMemcpyDeviceToHost,
                          e.Stream()));
    WG_HIP(hipStreamSynchronize(e.Stream()));
    *n_valid = h->vs_call.n_valid;
    h->plan_json = h->vs_call.plan.describe();
    if (plan_json && plan_len) {
      std::strncpy(plan_json, h->plan_json.c_str(), plan_len - 1);
      plan_json[plan_len - 1] = 0;
    }
  });
}

int wgx_make_plan(const wg_params* params, const double* e, long long nvis, int W, double sigma,
                  int tile, int sub_slab, char* plan_json, size_t plan_len) {
  return Guard([&] {
    WG_CHECK(params && e && plan_json && plan_len, "null argument");
    wg::ImageGeometry g;
    g.nx = int(params->width_t);
    g.ny = int(params->height_t);
    g.dx = params->pixel_size_x;
    g.dy = params->pixel_size_y;
    g.l_shift = params->l_shift;
    g.m_shift = params->m_shift;
    g.epsilon = params->epsilon;
    wg::PlanOptions o;
    o.force_W = W;
    o.force_sigma = sigma;
    o.T = tile;
    o.Sg = sub_slab;
    o.nvis_hint = nvis;
    const wg::Plan p = wg::MakePlan(g, wg::VisExtent{e[0], e[1], e[2], e[3], e[4], e[5]}, o);
    std::strncpy(plan_json, p.describe().c_str(), plan_len - 1);
    plan_json[plan_len - 1] = 0;
  });
}

int wgx_set_max_chunk_vis(wg_handle* h, size_t max_vis) {
  return Guard([&] {
    WG_CHECK(h, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    h->max_chunk_vis = max_vis;
  });
}

int wgx_device_bytes(wg_handle* h, size_t* bytes) {
  return Guard([&] {
    WG_CHECK(h && bytes, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    *bytes = h->eng->Arena().Allocated();
  });
}

int wgx_es_poly(int W, double beta, float* coef, int* degree) {
  return Guard([&] {
    WG_CHECK(coef && degree, "null argument");
    float t[wg::kEsPolyTaps][wg::kEsPolyCoefs];
    wg::EsPolyCoefficients(W, beta, t);
    std::memcpy(coef, t, sizeof(t));
    *degree = wg::EsPolyDegree(W);
  });
}

int wgx_facet_atlas(const wg_params* facets, size_t n_facets, const double* extents6,
                    const long long* nvis, char* json, size_t len) {
  return Guard([&] {
    WG_CHECK(facets && extents6 && nvis && json && len && n_facets, "null argument");
    std::vector<wg::ImageGeometry> geoms(n_facets);
    std::vector<wg::VisExtent> exts(n_facets);
    std::vector<long long> nv(nvis, nvis + n_facets);
    double eps = facets[0].epsilon;
    for (size_t f = 0; f < n_facets; ++f) eps = std::min(eps, facets[f].epsilon);
    for (size_t f = 0; f < n_facets; ++f) {
      geoms[f].nx = int(facets[f].width_t);
      geoms[f].ny = int(facets[f].height_t);
      geoms[f].dx = facets[f].pixel_size_x;
      geoms[f].dy = facets[f].pixel_size_y;
      geoms[f].l_shift = facets[f].l_shift;
      geoms[f].m_shift = facets[f].m_shift;
      geoms[f].epsilon = eps;
      const double* e = extents6 + 6 * f;
      exts[f] = wg::VisExtent{e[0], e[1], e[2], e[3], e[4], e[5]};
    }
    const wg::FacetLayout lay = wg::PlanFacets(geoms, exts, nv, wg::PlanOptions{});
    std::string out = "{\"atlas\":" + lay.atlas.describe() + ",\"facets\":[";
    for (size_t f = 0; f < n_facets; ++f) {
      if (f) out += ",";
      out += "{\"tu0\":" + std::to_string(lay.tu0[f]) + ",\"tv0\":" +
             std::to_string(lay.tv0[f]) + ",\"plan\":" + lay.plans[f].describe() + "}";
    }
    out += "]}";
    WG_CHECK(out.size() < len, "JSON buffer too small");
    std::strncpy(json, out.c_str(), len - 1);
    json[len - 1] = 0;
  });
}

int wgx_fft(int n, int nlines, int inverse, float* data) {
  return Guard([&] {
    WG_CHECK(data && n > 0 && nlines > 0 && n <= wg::kMaxFFT, "bad arguments");
    const int R = wg::LongFFTRadix(n);
    WG_CHECK(R > 0, "FFT length not split by 2..5 into LDS FFT lines");
    const int m = n / R;
    const std::vector<float2> tw = wg::FFTTwiddleTable(m);
    const size_t bytes = size_t(n) * nlines * sizeof(float2);
    DevBuf dtw(tw.size() * sizeof(float2)), dd(bytes), tmp(R > 1 ? bytes : 16);
    WG_HIP(hipMemcpy(dtw.p, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    WG_HIP(hipMemcpy(dd.p, data, bytes, hipMemcpyHostToDevice));
    if (R > 1)
      wg::LaunchLineFFTLong(dd.as<float2>(), nlines, n, R, wg::MakeFFTDesc(m, dtw.as<float2>()),
                            tmp.as<float2>(), inverse != 0, nullptr);
    else
      wg::LaunchLineFFT(dd.as<float2>(), nlines, wg::MakeFFTDesc(n, dtw.as<float2>()),
                        inverse != 0, nullptr);
    WG_HIP(hipDeviceSynchronize());
    WG_HIP(hipMemcpy(data, dd.p, bytes, hipMemcpyDeviceToHost));
  });
}

int wgx_radix_sort(size_t n, int key_bits, const uint32_t* keys_in, uint32_t* keys_out,
                   uint32_t* perm_out) {
  return Guard([&] {
    WG_CHECK(keys_in && keys_out && perm_out, "null argument");
    WG_CHECK(key_bits >= 1 && key_bits <= 32 && n < 0x7FFFFFFFull, "bad arguments");
    if (n == 0) return;
    const size_t bytes = n * sizeof(uint32_t);
    DevBuf k(bytes), v(bytes), k2(bytes), v2(bytes);
    DevBuf scr(wg::RadixSortScratchWords(n) * sizeof(uint32_t));
    WG_HIP(hipMemcpy(k.p, keys_in, bytes, hipMemcpyHostToDevice));
    const bool alt = wg::DeviceRadixSort(k.as<uint32_t>(), v.as<uint32_t>(), k2.as<uint32_t>(),
                                         v2.as<uint32_t>(), n, key_bits, scr.as<uint32_t>(),
                                         nullptr, true);
    WG_HIP(hipDeviceSynchronize());
    WG_HIP(hipMemcpy(keys_out, alt ? k2.p : k.p, bytes, hipMemcpyDeviceToHost));
    WG_HIP(hipMemcpy(perm_out, alt ? v2.p : v.p, bytes, hipMemcpyDeviceToHost));
  });
}

int wgx_synth_visibilities(const double* uvw_dev, size_t n_rows, size_t n_chan,
                           const double* freq, int n_src, const double* src, double noise_sigma,
                           uint64_t seed, float* vis_dev) {
  return Guard([&] {
    WG_CHECK(uvw_dev && freq && vis_dev && (n_src == 0 || src), "null argument");
    std::vector<double> scale(n_chan);
    for (size_t c = 0; c < n_chan; ++c) scale[c] = freq[c] / wg::kSpeedOfLight;
    DevBuf dscale(n_chan * sizeof(double)), dsrc((size_t(n_src) * 5 + 1) * sizeof(double));
    WG_HIP(hipMemcpy(dscale.p, scale.data(), n_chan * sizeof(double), hipMemcpyHostToDevice));
    if (n_src)
      WG_HIP(hipMemcpy(dsrc.p, src, size_t(n_src) * 5 * sizeof(double), hipMemcpyHostToDevice));
    wg::LaunchSynth(uvw_dev, n_rows, n_chan, dscale.as<double>(), dsrc.as<double>(), n_src,
                    noise_sigma, seed, reinterpret_cast<float2*>(vis_dev), nullptr);
    WG_HIP(hipDeviceSynchronize());
  });
}

// ------------------------------------------------------------ facet batches

int wg_facet_batch_create(const wg_params* facets, size_t n_facets, wg_facet_batch** out) {
  return Guard([&] {
    WG_CHECK(facets && out && n_facets > 0, "null argument or no facets");
    auto b = std::make_unique<wg_facet_batch>();
    std::vector<wg::ImageGeometry> geoms(n_facets);
    double eps = facets[0].epsilon;
    for (size_t f = 0; f < n_facets; ++f) {
      const wg_params& p = facets[f];
      WG_CHECK(p.width_t <= p.width && p.height_t <= p.height,
               "trimmed size exceeds image size");
      WG_CHECK(p.width_t > 0 && p.height_t > 0, "empty image");
      eps = std::min(eps, p.epsilon);
      geoms[f].nx = int(p.width_t);
      geoms[f].ny = int(p.height_t);
      geoms[f].dx = p.pixel_size_x;
      geoms[f].dy = p.pixel_size_y;
      geoms[f].l_shift = p.l_shift;
      geoms[f].m_shift = p.m_shift;
    }
    for (auto& g : geoms) g.epsilon = eps;
    b->prm.assign(facets, facets + n_facets);
    b->fb = std::make_unique<wg::FacetBatch>(geoms, wg::PlanOptions{});
    b->img.resize(n_facets);
    for (size_t f = 0; f < n_facets; ++f) {
      b->img[f] = static_cast<double*>(b->fb->Eng().Arena().Get(
          "fb.image" + std::to_string(f), b->npix(f) * sizeof(double)));
      WG_HIP(hipMemsetAsync(b->img[f], 0, b->npix(f) * sizeof(double), b->s()));
    }
    WG_HIP(hipStreamSynchronize(b->s()));
    *out = b.release();
  });
}

int wg_facet_batch_destroy(wg_facet_batch* b) {
  return Guard([&] { delete b; });
}

int wg_facet_batch_prepare(wg_facet_batch* b, size_t n_chan, const double* freq,
                           const double* uvw_dev, const size_t* row_offset,
                           const size_t* n_rows) {
  return Guard([&] {
    WG_CHECK(b && freq && uvw_dev && row_offset && n_rows, "null argument");
    std::lock_guard<std::mutex> lk(b->mu);
    DropGraph(b->cycle);
    b->fb->Prepare(n_chan, freq, uvw_dev, row_offset, n_rows);
    b->plan_json = b->fb->Describe();
    WG_HIP(hipStreamSynchronize(b->s()));
  });
}

int wg_facet_batch_initialize_inversion(wg_facet_batch* b) {
  return Guard([&] {
    WG_CHECK(b, "null handle");
    std::lock_guard<std::mutex> lk(b->mu);
    for (size_t f = 0; f < b->img.size(); ++f)
      WG_HIP(hipMemsetAsync(b->img[f], 0, b->npix(f) * sizeof(double), b->s()));
    WG_HIP(hipStreamSynchronize(b->s()));
  });
}

int wg_facet_batch_add_inversion(wg_facet_batch* b, const float* vis_dev) {
  return Guard([&] {
    WG_CHECK(b && vis_dev, "null argument");
    std::lock_guard<std::mutex> lk(b->mu);
    WG_CHECK(b->fb->Prepared(), "wg_facet_batch_prepare has not been called");
    b->fb->Invert(reinterpret_cast<const float2*>(vis_dev), b->img.data());
    WG_HIP(hipStreamSynchronize(b->s()));
  });
}

int wg_facet_batch_finalize_image(wg_facet_batch* b, size_t facet, double factor) {
  return Guard([&] {
    WG_CHECK(b && facet < b->img.size(), "null handle or facet out of range");
    std::lock_guard<std::mutex> lk(b->mu);
    ScaleKernel<<<Blocks(b->npix(facet)), 256, 0, b->s()>>>(b->img[facet], b->npix(facet),
                                                          factor);
    WG_HIP(hipGetLastError());
    WG_HIP(hipStreamSynchronize(b->s()));
  });
}

static void FacetRealImage(wg_facet_batch* b, size_t f, float* d) {
  const wg_params& p = b->prm[f];
  RealImageKernel<<<Blocks(p.width * p.height), 256, 0, b->s()>>>(
      b->img[f], d, int(p.width), int(p.height), int(p.width_t), int(p.height_t));
  WG_HIP(hipGetLastError());
}

int wg_facet_batch_real_image(wg_facet_batch* b, size_t facet, float* out) {
  return Guard([&] {
    WG_CHECK(b && out && facet < b->img.size(), "null argument or facet out of range");
    std::lock_guard<std::mutex> lk(b->mu);
    const size_t n = b->prm[facet].width * b->prm[facet].height;
    float* d = static_cast<float*>(b->fb->Eng().Arena().Get("d2h.image", n * 4));
    FacetRealImage(b, facet, d);
    WG_HIP(hipMemcpyAsync(out, d, n * 4, hipMemcpyDeviceToHost, b->s()));
    WG_HIP(hipStreamSynchronize(b->s()));
  });
}

int wg_facet_batch_real_image_device(wg_facet_batch* b, size_t facet, float* image_dev) {
  return Guard([&] {
    WG_CHECK(b && image_dev && facet < b->img.size(), "null argument or facet out of range");
    std::lock_guard<std::mutex> lk(b->mu);
    FacetRealImage(b, facet, image_dev);
    WG_HIP(hipStreamSynchronize(b->s()));
  });
}

static void FacetLoadModel(wg_facet_batch* b, size_t f, const float* d) {
  const wg_params& p = b->prm[f];
  LoadModelKernel<<<Blocks(b->npix(f)), 256, 0, b->s()>>>(
      d, b->img[f], int(p.width), int(p.height), int(p.width_t), int(p.height_t));
  WG_HIP(hipGetLastError());
}

int wg_facet_batch_initialize_prediction(wg_facet_batch* b, size_t facet, const float* image) {
  return Guard([&] {
    WG_CHECK(b && image && facet < b->img.size(), "null argument or facet out of range");
    std::lock_guard<std::mutex> lk(b->mu);
    const size_t n = b->prm[facet].width * b->prm[facet].height;
    float* d = static_cast<float*>(b->fb->Eng().Arena().Get("h2d.image", n * 4));
    WG_HIP(hipMemcpyAsync(d, image, n * 4, hipMemcpyHostToDevice, b->s()));
    FacetLoadModel(b, facet, d);
    WG_HIP(hipStreamSynchronize(b->s()));
  });
}

int wg_facet_batch_initialize_prediction_device(wg_facet_batch* b, size_t facet,
                                                const float* image_dev) {
  return Guard([&] {
    WG_CHECK(b && image_dev && facet < b->img.size(), "null argument or facet out of range");
    std::lock_guard<std::mutex> lk(b->mu);
    FacetLoadModel(b, facet, image_dev);
    WG_HIP(hipStreamSynchronize(b->s()));
  });
}

int wg_facet_batch_predict(wg_facet_batch* b, float* vis_out_dev) {
  return Guard([&] {
    WG_CHECK(b && vis_out_dev, "null argument");
    std::lock_guard<std::mutex> lk(b->mu);
    WG_CHECK(b->fb->Prepared(), "wg_facet_batch_prepare has not been called");
    std::vector<const double*> img(b->img.begin(), b->img.end());
    b->fb->Predict(img.data(), reinterpret_cast<float2*>(vis_out_dev));
    WG_HIP(hipStreamSynchronize(b->s()));
  });
}

int wg_facet_batch_capture_major_cycle(wg_facet_batch* b, const float* vis_dev,
                                       float* vis_out_dev, const double* factors, int what) {
  return Guard([&] {
    WG_CHECK(b && (what & 3) != 0, "null handle or nothing to capture");
    WG_CHECK(!(what & 1) || (vis_dev && factors), "invert needs vis_dev and factors");
    WG_CHECK(!(what & 2) || vis_out_dev, "predict needs vis_out_dev");
    std::lock_guard<std::mutex> lk(b->mu);
    WG_CHECK(b->fb->Prepared(), "wg_facet_batch_prepare has not been called");
    DropGraph(b->cycle);
    std::vector<double> fac;
    if (what & 1) fac.assign(factors, factors + b->img.size());
    std::vector<const double*> cimg(b->img.begin(), b->img.end());
    b->cycle = CaptureGraph(b->fb->Eng(), [&] {
      if (what & 1) {
        for (size_t f = 0; f < b->img.size(); ++f)
          wg::LaunchFillZero(b->img[f], b->npix(f) * sizeof(double), b->s());
        b->fb->Invert(reinterpret_cast<const float2*>(vis_dev), b->img.data());
        for (size_t f = 0; f < b->img.size(); ++f) {
          ScaleKernel<<<Blocks(b->npix(f)), 256, 0, b->s()>>>(b->img[f], b->npix(f), fac[f]);
          WG_HIP(hipGetLastError());
        }
      }
      if (what & 2) b->fb->Predict(cimg.data(), reinterpret_cast<float2*>(vis_out_dev));
    });
    b->cycle_gen = b->fb->Eng().Arena().Generation();
  });
}

int wg_facet_batch_replay_major_cycle(wg_facet_batch* b, int count) {
  return Guard([&] {
    WG_CHECK(b && count >= 0, "null handle or negative count");
    std::lock_guard<std::mutex> lk(b->mu);
    WG_CHECK(b->cycle, "no captured major cycle (wg_facet_batch_capture_major_cycle)");
    if (b->fb->Eng().Arena().Generation() != b->cycle_gen) {
      DropGraph(b->cycle);
      throw wg::Error("a later call reallocated device buffers of the captured major cycle; "
                      "capture it again (wg_facet_batch_capture_major_cycle)");
    }
    for (int i = 0; i < count; ++i) WG_HIP(hipGraphLaunch(b->cycle, b->s()));
    WG_HIP(hipStreamSynchronize(b->s()));
  });
}

int wg_facet_batch_set_timing(wg_facet_batch* b, int enabled) {
  return Guard([&] {
    WG_CHECK(b, "null handle");
    b->fb->Eng().timing = enabled != 0;
  });
}

int wg_facet_batch_stage_times(wg_facet_batch* b, double* out14, int reset) {
  return Guard([&] {
    WG_CHECK(b && out14, "null argument");
    std::lock_guard<std::mutex> lk(b->mu);
    wg::Engine& e = b->fb->Eng();
    e.CollectTimes();
    wg::StageTimes& t = e.Times();
    const double v[14] = {t.plan, t.sort, t.grid, t.fft, t.degrid,
                          t.plan + t.sort + t.grid + t.fft + t.degrid,
                          t.grid_kernel_ms, t.grid_launches, t.grid_bytes,
                          t.degrid_kernel_ms, t.degrid_launches, t.degrid_bytes,
                          t.grid_passes, t.degrid_passes};
    std::copy(v, v + 14, out14);
    if (reset) t = wg::StageTimes{};
  });
}

int wg_facet_batch_virtual_count(wg_facet_batch* b, size_t* n) {
  return Guard([&] {
    WG_CHECK(b && n, "null argument");
    *n = b->fb->NumVirtual();
  });
}

int wg_facet_batch_plan_json(wg_facet_batch* b, char* buf, size_t len) {
  return Guard([&] {
    WG_CHECK(b && buf && len > 0, "null argument");
    std::lock_guard<std::mutex> lk(b->mu);
    std::strncpy(buf, b->plan_json.c_str(), len - 1);
    buf[len - 1] = 0;
  });
}

// ------------------------------------------------------------ imaging weights

int wg_image_weights_create(int mode, double briggs_robustness, size_t width, size_t height,
                            double pixel_scale_x, double pixel_scale_y, int weights_as_taper,
                            double super_weight, wg_image_weights** out) {
  return Guard([&] {
    WG_CHECK(out, "null argument");
    WG_CHECK(mode >= 0 && mode <= 2, "mode: 0 natural, 1 uniform, 2 Briggs");
    WG_CHECK(super_weight > 0 && pixel_scale_x > 0 && pixel_scale_y > 0, "bad scale");
    auto w = std::make_unique<wg_image_weights>();
    w->mode = mode;
    w->robust = briggs_robustness;
    w->as_taper = weights_as_taper != 0;
    // imageweights.cpp:33-48: sizes divided by the super weight, made even
    size_t iw = size_t(std::llround(double(width) / super_weight));
    size_t ih = size_t(std::llround(double(height) / super_weight));
    if (iw % 2) ++iw;
    if (ih % 2) ++ih;
    WG_CHECK(iw >= 2 && ih >= 2 && iw < (1u << 30) && ih < (1u << 30), "bad weight grid size");
    WG_HIP(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    w->arena.SetStream(w->stream);
    w->g.width = int(iw);
    w->g.height = int(ih);
    w->g.psx = pixel_scale_x;
    w->g.psy = pixel_scale_y;
    w->g.grid = static_cast<double*>(w->arena.Get("grid", w->cells() * sizeof(double)));
    w->total_part =
        static_cast<double*>(w->arena.Get("total", wg::WeightStatPartials() * sizeof(double)));
    w->sums = static_cast<double*>(w->arena.Get("sums", 2 * sizeof(double)));
    wg::LaunchFillZero(w->g.grid, w->cells() * sizeof(double), w->stream);
    wg::LaunchFillZero(w->total_part, wg::WeightStatPartials() * sizeof(double), w->stream);
    wg::LaunchFillZero(w->sums, 2 * sizeof(double), w->stream);
    WG_HIP(hipStreamSynchronize(w->stream));
    *out = w.release();
  });
}

int wg_image_weights_destroy(wg_image_weights* w) {
  return Guard([&] { delete w; });
}

int wg_image_weights_size(const wg_image_weights* w, size_t* width, size_t* height) {
  return Guard([&] {
    WG_CHECK(w && width && height, "null argument");
    *width = size_t(w->g.width);
    *height = size_t(w->g.height);
  });
}

static std::vector<double> Wavelengths(const double* freq, size_t n_chan) {
  std::vector<double> lam(n_chan);
  for (size_t c = 0; c < n_chan; ++c) {
    WG_CHECK(freq[c] > 0, "frequencies must be positive");
    lam[c] = wg::kSpeedOfLight / freq[c];
  }
  return lam;
}

int wg_image_weights_grid(wg_image_weights* w, size_t n_rows, size_t n_chan, size_t n_pol,
                          const double* uvw_dev, const double* freq, const float* weights_dev) {
  return Guard([&] {
    WG_CHECK(w && freq && (n_rows == 0 || (uvw_dev && weights_dev)), "null argument");
    WG_CHECK(n_pol >= 1 && n_pol <= 4, "1 to 4 polarizations");
    std::lock_guard<std::mutex> lk(w->mu);
    WG_CHECK(!w->finished, "Grid() after FinishGridding()");
    const std::vector<double> lam = Wavelengths(freq, n_chan);
    double* dl = static_cast<double*>(w->arena.Get("lambda", n_chan * sizeof(double)));
    WG_HIP(hipMemcpyAsync(dl, lam.data(), n_chan * sizeof(double), hipMemcpyHostToDevice,
                          w->stream));
    wg::LaunchGridWeights(w->g, uvw_dev, dl, n_rows, n_chan, n_pol, weights_dev, w->as_taper,
                          w->total_part, w->stream);
    WG_HIP(hipStreamSynchronize(w->stream));
  });
}

int wg_image_weights_finish(wg_image_weights* w) {
  return Guard([&] {
    WG_CHECK(w, "null handle");
    std::lock_guard<std::mutex> lk(w->mu);
    WG_CHECK(!w->finished, "FinishGridding() called twice");
    w->finished = true;
    double* scratch =
        static_cast<double*>(w->arena.Get("scratch", wg::WeightStatPartials() * sizeof(double)));
    wg::LaunchFinishWeights(w->g, w->mode, w->robust, w->total_part, w->sums, scratch, w->stream);
    WG_HIP(hipStreamSynchronize(w->stream));
  });
}

static int Taper(wg_image_weights* w, int kind, double a, double b) {
  return Guard([&] {
    WG_CHECK(w, "null handle");
    std::lock_guard<std::mutex> lk(w->mu);
    wg::LaunchWeightTaper(w->g, kind, a, b, w->stream);
    WG_HIP(hipStreamSynchronize(w->stream));
  });
}

int wg_image_weights_set_min_uv_range(wg_image_weights* w, double min_uv) {
  return Taper(w, 0, min_uv, 0);
}
int wg_image_weights_set_max_uv_range(wg_image_weights* w, double max_uv) {
  return Taper(w, 1, max_uv, 0);
}
int wg_image_weights_set_tukey_taper(wg_image_weights* w, double transition, double max_uv) {
  return Taper(w, 2, transition, max_uv);
}
int wg_image_weights_set_tukey_inner_taper(wg_image_weights* w, double transition,
                                           double min_uv) {
  return Taper(w, 3, transition, min_uv);
}
int wg_image_weights_set_edge_taper(wg_image_weights* w, double size) {
  return Taper(w, 4, size, 0);
}
int wg_image_weights_set_edge_tukey_taper(wg_image_weights* w, double transition, double edge) {
  return Taper(w, 5, transition, edge);
}
int wg_image_weights_set_gaussian_taper(wg_image_weights* w, double beam_size) {
  // imageweights.cpp:333-341: -2 (half-power uv * 2 sqrt(2 ln 2))^2
  const double half_power_uv = 1.0 / (beam_size * 2.0 * M_PI);
  const long double sigma_to_hp = 2.0L * sqrtl(2.0L * logl(2.0L));
  double m2s = half_power_uv * double(sigma_to_hp);
  m2s *= -2.0 * m2s;
  return Taper(w, 6, m2s, 0);
}

int wg_image_weights_rank_filter(wg_image_weights* w, double rank_limit, size_t window) {
  return Guard([&] {
    WG_CHECK(w && window > 0, "null handle or empty window");
    std::lock_guard<std::mutex> lk(w->mu);
    double* src = static_cast<double*>(w->arena.Get("rank.src", w->cells() * sizeof(double)));
    WG_HIP(hipMemcpyAsync(src, w->g.grid, w->cells() * sizeof(double), hipMemcpyDeviceToDevice,
                          w->stream));
    wg::LaunchRankFilter(w->g, src, rank_limit, int(window), w->stream);
    WG_HIP(hipStreamSynchronize(w->stream));
    w->arena.Release("rank.src");
  });
}

int wg_image_weights_get_grid(wg_image_weights* w, double* image) {
  return Guard([&] {
    WG_CHECK(w && image, "null argument");
    std::lock_guard<std::mutex> lk(w->mu);
    const size_t n = size_t(w->g.width) * size_t(w->g.height);
    double* d = static_cast<double*>(w->arena.Get("image", n * sizeof(double)));
    wg::LaunchGetWeightGrid(w->g, d, w->stream);
    WG_HIP(hipMemcpyAsync(image, d, n * sizeof(double), hipMemcpyDeviceToHost, w->stream));
    WG_HIP(hipStreamSynchronize(w->stream));
    w->arena.Release("image");
  });
}

int wg_image_weights_sample(wg_image_weights* w, size_t n, const double* u, const double* v,
                            double* out) {
  return Guard([&] {
    WG_CHECK(w && (n == 0 || (u && v && out)), "null argument");
    std::lock_guard<std::mutex> lk(w->mu);
    if (!n) return;
    double* d = static_cast<double*>(w->arena.Get("sample", 3 * n * sizeof(double)));
    WG_HIP(hipMemcpyAsync(d, u, n * sizeof(double), hipMemcpyHostToDevice, w->stream));
    WG_HIP(hipMemcpyAsync(d + n, v, n * sizeof(double), hipMemcpyHostToDevice, w->stream));
    wg::LaunchSampleWeights(w->g, d, d + n, n, d + 2 * n, w->stream);
    WG_HIP(hipMemcpyAsync(out, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost, w->stream));
    WG_HIP(hipStreamSynchronize(w->stream));
  });
}

int wg_prepare_rows(wg_handle* h, size_t n_rows, size_t n_chan, size_t n_pol,
                    const double* uvw_dev, const double* freq, const float* vis_dev,
                    const float* weights_dev, const float* model_dev, const uint8_t* selected,
                    int psf, double psf_dl, double psf_dm, int weighting, int polarization,
                    wg_image_weights* image_weights, float* out_dev, double* stats4) {
  return Guard([&] {
    WG_CHECK(wg::ValidPolCode(polarization), "unknown polarization code");
    WG_CHECK(h && freq && stats4, "null argument");
    WG_CHECK(n_rows == 0 || (uvw_dev && weights_dev && out_dev && (psf || vis_dev)),
             "null argument");
    WG_CHECK(n_pol == 1 || n_pol == 2 || n_pol == 4, "1, 2 or 4 polarizations");
    WG_CHECK(weighting >= 0 && weighting <= 2, "weighting: 0 normal, 1 squared, 2 unit");
    std::lock_guard<std::mutex> lk(h->mu);
    wg::DeviceArena& a = h->eng->Arena();
    const std::vector<double> lam = Wavelengths(freq, n_chan);
    double* dl = static_cast<double*>(a.Get("prep.lambda", n_chan * sizeof(double)));
    WG_HIP(hipMemcpyAsync(dl, lam.data(), n_chan * sizeof(double), hipMemcpyHostToDevice,
                          h->s()));
    uint8_t* dsel = nullptr;
    if (selected) {
      dsel = static_cast<uint8_t*>(a.Get("prep.selected", n_chan));
      WG_HIP(hipMemcpyAsync(dsel, selected, n_chan, hipMemcpyHostToDevice, h->s()));
    }
    double* part =
        static_cast<double*>(a.Get("prep.part", wg::WeightStatPartials() * sizeof(double)));
    double* dstats = static_cast<double*>(a.Get("prep.stats", 4 * sizeof(double)));
    wg::PrepArgs args{};
    args.n_rows = n_rows;
    args.n_chan = n_chan;
    args.n_pol = int(n_pol);
    args.psf = psf != 0;
    args.psf_dl = psf_dl;
    args.psf_dm = psf_dm;
    args.psf_dn = std::sqrt(1.0 - psf_dl * psf_dl - psf_dm * psf_dm) - 1.0;
    args.weighting = weighting;
    args.polarization = polarization;
    std::unique_lock<std::mutex> wl;
    const wg::WeightGrid* g = nullptr;
    if (image_weights) {
      wl = std::unique_lock<std::mutex>(image_weights->mu);
      WG_CHECK(image_weights->finished, "imaging weights used before FinishGridding()");
      // the grid is written on the weights' own stream: finished and synchronised
      g = &image_weights->g;
    }
    wg::LaunchPrepareRows(args, uvw_dev, dl, reinterpret_cast<const float2*>(vis_dev),
                          weights_dev, reinterpret_cast<const float2*>(model_dev), dsel, g,
                          reinterpret_cast<float2*>(out_dev), part, dstats, h->s());
    WG_HIP(hipMemcpyAsync(stats4, dstats, 4 * sizeof(double), hipMemcpyDeviceToHost, h->s()));
    WG_HIP(hipStreamSynchronize(h->s()));
  });
}

// ------------------------------------------------ allocation-free row prep

int wg_prepare_rows_noalloc(wg_handle* h, size_t n_rows, size_t n_chan, size_t n_pol,
                            const double* uvw_dev, const double* freq, const float* vis_dev,
                            const float* weights_dev, const float* model_dev,
                            const uint8_t* selected, int psf, double psf_dl, double psf_dm,
                            int weighting, int polarization, wg_image_weights* image_weights,
                            float* out_dev, double* stats4) {
  return Guard([&] {
    WG_CHECK(h && freq && stats4, "null argument");
    WG_CHECK(n_rows == 0 || (uvw_dev && weights_dev && out_dev && (psf || vis_dev)),
             "null argument");
    WG_CHECK(n_pol == 1 || n_pol == 2 || n_pol == 4, "1, 2 or 4 polarizations");
    WG_CHECK(weighting >= 0 && weighting <= 2, "weighting: 0 normal, 1 squared, 2 unit");
    WG_CHECK(wg::ValidPolCode(polarization), "unknown polarization code");
    WG_CHECK(n_chan >= 1 && n_chan <= kNoAllocMaxChan,
             "wg_prepare_rows_noalloc takes 1 to 65536 channels (its fixed device scratch)");
    WG_CHECK(wg::WeightStatPartials() == kNoAllocPartWords, "statistics partials size changed");
    for (size_t c = 0; c < n_chan; ++c) WG_CHECK(freq[c] > 0, "frequencies must be positive");
    int dev = -1;
    WG_HIP(hipGetDevice(&dev));
    WG_CHECK(dev == h->device, "the current device is not the handle's device");
    WG_CHECK(dev >= 0 && dev < kNoAllocMaxDevices, "device index out of range");
    // one call at a time per device: calls share the device's scratch
    std::lock_guard<std::mutex> sl(g_scratch_mu[dev]);
    if (!g_scratch_ptr[dev]) {
      void* p = nullptr;
      WG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      WG_CHECK(p, "no device address for the row-preparation scratch");
      g_scratch_ptr[dev] = p;
    }
    unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    double* part = reinterpret_cast<double*>(base);
    double* dstats = reinterpret_cast<double*>(base + kNoAllocStatsOff);
    double* dl = reinterpret_cast<double*>(base + kNoAllocLambdaOff);
    uint8_t* dsel = selected ? base + kNoAllocSelOff : nullptr;
    std::lock_guard<std::mutex> lk(h->mu);
    hipStream_t s = h->s();
    WG_HIP(hipMemcpyAsync(dl, freq, n_chan * sizeof(double), hipMemcpyHostToDevice, s));
    WavelengthsKernel<<<1, 256, 0, s>>>(dl, n_chan);
    WG_HIP(hipGetLastError());
    if (selected) WG_HIP(hipMemcpyAsync(dsel, selected, n_chan, hipMemcpyHostToDevice, s));
    wg::PrepArgs args{};
    args.n_rows = n_rows;
    args.n_chan = n_chan;
    args.n_pol = int(n_pol);
    args.psf = psf != 0;
    args.psf_dl = psf_dl;
    args.psf_dm = psf_dm;
    args.psf_dn = std::sqrt(1.0 - psf_dl * psf_dl - psf_dm * psf_dm) - 1.0;
    args.weighting = weighting;
    args.polarization = polarization;
    std::unique_lock<std::mutex> wl;
    const wg::WeightGrid* g = nullptr;
    if (image_weights) {
      wl = std::unique_lock<std::mutex>(image_weights->mu);
      WG_CHECK(image_weights->finished, "imaging weights used before FinishGridding()");
      g = &image_weights->g;
    }
    // the row kernel writes every partial and the statistics kernel all
    // four statistics, so nothing of the previous call's contents is read
    wg::LaunchPrepareRows(args, uvw_dev, dl, reinterpret_cast<const float2*>(vis_dev),
                          weights_dev, reinterpret_cast<const float2*>(model_dev), dsel, g,
                          reinterpret_cast<float2*>(out_dev), part, dstats, s);
    WG_HIP(hipMemcpyAsync(stats4, dstats, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    WG_HIP(hipStreamSynchronize(s));
  });
}

// ------------------------------------------------------------ reordered parts

int wg_part_open(const char* meta_path, const char* data_path, const char* weight_path,
                 const char* model_path, size_t n_pol, wg_part_reader** out) {
  return Guard([&] {
    WG_CHECK(meta_path && data_path && out, "null argument");
    auto pr = std::make_unique<wg_part_reader>();
    pr->r = std::make_unique<wg::PartReader>(meta_path, data_path, weight_path ? weight_path : "",
                                             model_path ? model_path : "", n_pol);
    *out = pr.release();
  });
}

int wg_part_close(wg_part_reader* p) {
  return Guard([&] { delete p; });
}

int wg_part_info(const wg_part_reader* p, uint64_t* info6, double* start_time, char* ms_path,
                 size_t ms_path_len) {
  return Guard([&] {
    WG_CHECK(p && info6, "null argument");
    const wg::PartReader& r = *p->r;
    const uint64_t v[6] = {r.Rows(), r.Channels(), r.ChannelStart(), r.DataDescId(),
                           r.HasModel() ? 1u : 0u, r.Polarizations()};
    std::copy(v, v + 6, info6);
    if (start_time) *start_time = r.StartTime();
    if (ms_path && ms_path_len) {
      std::strncpy(ms_path, r.MsPath().c_str(), ms_path_len - 1);
      ms_path[ms_path_len - 1] = 0;
    }
  });
}

int wg_part_read(const wg_part_reader* p, size_t row0, size_t n, double* uvw, uint16_t* antennas,
                 double* times, uint16_t* fields, float* vis, float* weights, float* model) {
  return Guard([&] {
    WG_CHECK(p, "null argument");
    const wg::PartReader& r = *p->r;
    WG_CHECK(row0 <= r.Rows() && n <= r.Rows() - row0, "rows beyond the part");
    if (uvw || antennas || times || fields) r.ReadMeta(row0, n, uvw, antennas, times, fields);
    if (vis) r.ReadData(row0, n, vis);
    if (weights) r.ReadWeights(row0, n, weights);
    if (model) r.ReadModel(row0, n, model);
  });
}

int wg_part_grid(wg_part_reader* p, wg_handle* h, const double* freq, size_t chunk_rows,
                 int weighting, int polarization, wg_image_weights* image_weights,
                 double* stats4) {
  return Guard([&] {
    WG_CHECK(p && h && freq && stats4, "null argument");
    WG_CHECK(weighting >= 0 && weighting <= 2, "weighting: 0 normal, 1 squared, 2 unit");
    WG_CHECK(wg::ValidPolCode(polarization), "unknown polarization code");
    const wg::PartReader& r = *p->r;
    const size_t R = r.Rows(), C = r.Channels(), P = r.Polarizations();
    std::fill(stats4, stats4 + 4, 0.0);
    for (size_t c = 0; c < C; ++c) WG_CHECK(freq[c] > 0, "frequencies must be positive");
    std::lock_guard<std::mutex> lk(h->mu);
    if (R == 0) return;
    std::unique_lock<std::mutex> wl;
    const wg::WeightGrid* g = nullptr;
    if (image_weights) {
      wl = std::unique_lock<std::mutex>(image_weights->mu);
      WG_CHECK(image_weights->finished, "imaging weights used before FinishGridding()");
      g = &image_weights->g;
    }
    // rows per chunk: at most 2^24 visibilities (the page-locked staging of
    // two chunks), and what one engine pass takes
    size_t step = chunk_rows ? chunk_rows : std::max<size_t>(1, (size_t(1) << 24) / (C * P));
    step = std::min({step, R, ChunkRows(h, R, C * (P + 1))});
    // two page-locked staging sets: the disk read of the next chunk runs on
    // a worker thread while the device grids the current one
    struct Stage {
      std::unique_ptr<PinnedBuf> rec, uvw, vis, wts;
    } st[2];
    for (Stage& b : st) {
      b.rec = std::make_unique<PinnedBuf>(step * wg::kMetaRecordBytes);
      b.uvw = std::make_unique<PinnedBuf>(step * 3 * sizeof(double));
      b.vis = std::make_unique<PinnedBuf>(step * C * P * 8);
      b.wts = std::make_unique<PinnedBuf>(step * C * P * 4);
    }
    PinnedBuf hstats(4 * sizeof(double));
    auto read_chunk = [&r, &st, step, R](size_t k) {
      const size_t r0 = k * step, nr = std::min(step, R - r0);
      Stage& b = st[k & 1];
      r.ReadMetaRaw(r0, nr, b.rec->p);
      wg::PartReader::UnpackUvw(b.rec->p, nr, b.uvw->as<double>());
      r.ReadData(r0, nr, b.vis->as<float>());
      r.ReadWeights(r0, nr, b.wts->as<float>());
    };
    wg::DeviceArena& a = h->eng->Arena();
    hipStream_t s = h->s();
    const std::vector<double> lam = Wavelengths(freq, C);
    double* dlam = static_cast<double*>(a.Get("part.lambda", C * sizeof(double)));
    WG_HIP(hipMemcpyAsync(dlam, lam.data(), C * sizeof(double), hipMemcpyHostToDevice, s));
    WG_HIP(hipStreamSynchronize(s));
    double* dpart =
        static_cast<double*>(a.Get("part.partials", wg::WeightStatPartials() * sizeof(double)));
    double* dstats = static_cast<double*>(a.Get("part.stats", 4 * sizeof(double)));
    double* duvw = static_cast<double*>(a.Get("part.uvw", step * 3 * sizeof(double)));
    float2* dvis = static_cast<float2*>(a.Get("part.vis", step * C * P * sizeof(float2)));
    float* dwts = static_cast<float*>(a.Get("part.wts", step * C * P * sizeof(float)));
    float2* dout = static_cast<float2*>(a.Get("part.out", step * C * sizeof(float2)));
    const size_t nchunks = (R + step - 1) / step;
    std::future<void> next = std::async(std::launch::async, read_chunk, size_t(0));
    for (size_t k = 0; k < nchunks; ++k) {
      next.get();  // chunk k is in staging set k & 1 (rethrows a read error)
      const size_t r0 = k * step, nr = std::min(step, R - r0);
      const Stage& b = st[k & 1];
      WG_HIP(hipMemcpyAsync(duvw, b.uvw->p, nr * 3 * sizeof(double), hipMemcpyHostToDevice, s));
      WG_HIP(hipMemcpyAsync(dvis, b.vis->p, nr * C * P * 8, hipMemcpyHostToDevice, s));
      WG_HIP(hipMemcpyAsync(dwts, b.wts->p, nr * C * P * 4, hipMemcpyHostToDevice, s));
      WG_HIP(hipStreamSynchronize(s));  // staging set k & 1 is free again
      if (k + 1 < nchunks) next = std::async(std::launch::async, read_chunk, k + 1);
      wg::PrepArgs args{};
      args.n_rows = nr;
      args.n_chan = C;
      args.n_pol = int(P);
      args.weighting = weighting;
      args.polarization = polarization;
      wg::LaunchPrepareRows(args, duvw, dlam, dvis, dwts, nullptr, nullptr, g, dout, dpart, dstats,
                            s);
      WG_HIP(hipMemcpyAsync(hstats.p, dstats, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
      AddInversion(h, nr, C, duvw, freq, reinterpret_cast<const float*>(dout));
      const double* q = hstats.as<double>();
      stats4[0] += q[0];
      stats4[1] += q[1];
      stats4[2] = std::max(stats4[2], q[2]);
      stats4[3] += q[3];
    }
  });
}

// ------------------------------------------------------------ resampling

int wg_resample(wg_handle* h, const float* in_dev, size_t in_w, size_t in_h, float* out_dev,
                size_t out_w, size_t out_h) {
  return Guard([&] {
    WG_CHECK(h && in_dev && out_dev, "null argument");
    WG_CHECK(in_w <= 16384 && in_h <= 16384 && out_w <= 16384 && out_h <= 16384,
             "image side above 16384");
    std::lock_guard<std::mutex> lk(h->mu);
    h->eng->Resample(in_dev, int(in_w), int(in_h), out_dev, int(out_w), int(out_h));
  });
}

int wg_real_image_resampled(wg_handle* h, float* out, size_t out_w, size_t out_h) {
  return Guard([&] {
    WG_CHECK(h && out, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    const size_t n = h->prm.width * h->prm.height;
    float* d = static_cast<float*>(h->eng->Arena().Get("d2h.image", n * 4));
    float* r = static_cast<float*>(h->eng->Arena().Get("resampled.image", out_w * out_h * 4));
    RealImageTo(h, d);
    h->eng->Resample(d, int(h->prm.width), int(h->prm.height), r, int(out_w), int(out_h));
    WG_HIP(hipMemcpyAsync(out, r, out_w * out_h * 4, hipMemcpyDeviceToHost, h->s()));
    WG_HIP(hipStreamSynchronize(h->s()));
  });
}

int wg_initialize_prediction_resampled(wg_handle* h, const float* image, size_t in_w,
                                       size_t in_h) {
  return Guard([&] {
    WG_CHECK(h && image, "null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    const size_t n = h->prm.width * h->prm.height;
    float* src = static_cast<float*>(h->eng->Arena().Get("resample.src", in_w * in_h * 4));
    float* d = static_cast<float*>(h->eng->Arena().Get("h2d.image", n * 4));
    WG_HIP(hipMemcpyAsync(src, image, in_w * in_h * 4, hipMemcpyHostToDevice, h->s()));
    h->eng->Resample(src, int(in_w), int(in_h), d, int(h->prm.width), int(h->prm.height));
    LoadModelFrom(h, d);
    WG_HIP(hipStreamSynchronize(h->s()));
  });
}

}  // extern "C"
