// Served request stream on one GPU, in the library: continuous batching of
// w-gridding requests with admission by measured free HBM.
//
// A request is one imaging job: an image geometry, device-resident uvw and
// visibilities, and a number of major cycles.  Each cycle inverts the
// residual visibilities into a dirty image, takes a model from it (0.1 x the
// pixels at or above half the image's peak: a stand-in for the
// deconvolution), predicts the model's visibilities and forms the next
// residual -- WSClean's major cycle (main/wsclean.cpp:1519-1632) on the
// gridder calls of this library (wg_prepare_visibilities, then per cycle
// wg_initialize_inversion, wg_add_inversion_prepared, wg_finalize_image,
// wg_real_image_device, wg_initialize_prediction_device,
// wg_predict_prepared).  The sorted visibility set, the grids and the images
// stay in HBM for every cycle; only the last dirty image goes to the host.
//
// Requests are queued at submission and admitted in order: the head of the
// queue starts as soon as a worker slot is free and its device memory
// (estimate less what its pooled handle already holds), plus what the
// running requests have still to allocate (their estimates less what they
// reported after preparing and after their first cycle), plus a reserve,
// fits in the device memory free at that moment (hipMemGetInfo).  Short of
// memory, idle pooled handles are closed first, least recently used first;
// a request that does not fit even alone runs when nothing else does.
// Later arrivals wait behind the head, so large requests are not starved.
//
// Each admitted request runs on a gridder handle of its geometry taken from
// a pool of idle handles (so the correction table, FFT twiddles and device
// arena of earlier requests are reused) and on its worker's HIP stream;
// every handle owns its own stream and stream-ordered arena, so admitted
// requests overlap on the device.  A request's results do not depend on what
// runs beside it: every kernel of the path is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/wgrid_hip.h"
#include "common.hpp"

namespace wg {
void SetLastError(const std::string& msg);  // capi.hip
}

namespace {

using Clock = std::chrono::steady_clock;

// Largest pixel value, NaN pixels ignored, as an order-preserving unsigned
// key (atomicMax): the key of x is bits(x) ^ 0x80000000 for x >= 0 and
// ~bits(x) below; 0 (below every non-NaN key) when no pixel is a number.
__device__ __forceinline__ uint32_t OrderKey(float x) {
  const uint32_t b = __float_as_uint(x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float FromOrderKey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void __launch_bounds__(256) PeakKernel(const float* __restrict__ img, size_t n,
                                                  uint32_t* __restrict__ peak) {
  uint32_t best = 0u;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const float x = img[i];
    if (x == x) best = max(best, OrderKey(x));
  }
  for (int m = 32; m >= 1; m >>= 1) best = max(best, uint32_t(__shfl_xor(int(best), m)));
  if ((threadIdx.x & 63) == 0) atomicMax(peak, best);
}

// model = 0.1 img where img >= peak / 2, else 0 (NaN pixels give 0).
__global__ void __launch_bounds__(256) ModelKernel(const float* __restrict__ img, size_t n,
                                                   const uint32_t* __restrict__ peak,
                                                   float* __restrict__ model) {
  const uint32_t k = *peak;
  const float half = k ? 0.5f * FromOrderKey(k) : -__builtin_inff();
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const float x = img[i];
    model[i] = x >= half ? x * 0.1f : 0.f;
  }
}

// residual = vis - out, element by element (complex64 as float pairs).
__global__ void __launch_bounds__(256) ResidualKernel(const float* __restrict__ vis,
                                                      const float* __restrict__ out, size_t n,
                                                      float* __restrict__ residual) {
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256)
    residual[i] = vis[i] - out[i];
}

unsigned GridOf(size_t n) {
  return unsigned(std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 65536)));
}

void Check(int status) {
  if (status != 0) throw wg::Error(wg_last_error());
}

// Device bytes a request allocates beyond its inputs: per visibility the
// sorted set (keys and permutation with their ping-pong copies, coordinates,
// phasors, gathered and accumulated visibilities, grown by the arena's 1/8
// rounding) and the residual; per pixel the FP64 image, the dirty /
// correction / transfer images and the worker's dirty image and model; the
// FFT slab (w planes of at most (2 npix)^2 uv cells plus two line buffers,
// capped by the planner's slab budget); and 64 MiB of plan tables.
size_t Estimate(const wg_request& r) {
  const size_t n_vis = r.n_rows * r.n_chan, n2 = r.npix * r.npix;
  const size_t slab = std::min<size_t>(size_t(8) << 30, 64 * n2 * 32);
  return 72 * n_vis + 24 * r.n_rows + 32 * n2 + slab + (size_t(64) << 20);
}

using Key = std::tuple<size_t, double, double>;  // image side, pixel size, epsilon

struct Job {
  uint64_t id = 0;
  wg_request req{};
  std::vector<double> freq;
  size_t need = 0;
  size_t allocated = 0;  // reported by the worker
  wg_handle* handle = nullptr;
  double t_submit = 0, t_start = 0, t_done = 0;
  bool done = false;
  std::string error;
};

}  // namespace

struct wg_server {
  int slots = 1;
  size_t reserve = 0;
  int max_idle = 8;
  int device = 0;
  Clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<Job>> queue;
  std::vector<std::shared_ptr<Job>> running;
  std::map<uint64_t, std::shared_ptr<Job>> jobs;
  std::vector<std::pair<Key, wg_handle*>> idle;  // least recently returned first
  std::vector<std::thread> workers;
  uint64_t next_id = 1;
  bool stopping = false;
  // statistics
  uint64_t completed = 0, peak_running = 0, hits = 0, misses = 0, evictions = 0, waits = 0;
  size_t min_free = SIZE_MAX;

  double Now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  // Under mu: true once the head job may start, its handle (the pooled one
  // of its geometry or a new one) in j->handle; false to wait.
  bool TryAdmit(const std::shared_ptr<Job>& j) {
    if (int(running.size()) >= slots) return false;
    const Key key{j->req.npix, j->req.pixel_size, j->req.epsilon};
    for (;;) {
      // the most recently returned idle handle of the job's geometry
      int mine = -1;
      for (int i = int(idle.size()) - 1; i >= 0; --i)
        if (idle[size_t(i)].first == key) {
          mine = i;
          break;
        }
      size_t have = 0;
      if (mine >= 0) {
        size_t b = 0;
        if (wgx_device_bytes(idle[size_t(mine)].second, &b) == 0) have = b;
      }
      size_t free_b = 0, total_b = 0;
      WG_HIP(hipMemGetInfo(&free_b, &total_b));
      size_t to_come = 0;
      for (const auto& r : running) to_come += r->need > r->allocated ? r->need - r->allocated : 0;
      const size_t want = (j->need > have ? j->need - have : 0) + to_come + reserve;
      const bool fits = want <= free_b;
      // short of memory: close the least recently used idle handle other
      // than the job's own and measure again
      const int victim = (mine == 0) ? 1 : 0;
      if (!fits && victim < int(idle.size())) {
        wg_destroy(idle[size_t(victim)].second);
        idle.erase(idle.begin() + victim);
        ++evictions;
        continue;
      }
      if (!fits && !running.empty()) return false;
      min_free = std::min(min_free, free_b);
      if (mine >= 0) {
        j->handle = idle[size_t(mine)].second;
        idle.erase(idle.begin() + mine);
        ++hits;
      } else {
        wg_params p{};
        p.width = p.height = p.width_t = p.height_t = j->req.npix;
        p.pixel_size_x = p.pixel_size_y = j->req.pixel_size;
        p.nthreads = 1;
        p.epsilon = j->req.epsilon;
        wg_handle* h = nullptr;
        if (wg_create(&p, &h) != 0) throw wg::Error(wg_last_error());
        j->handle = h;
        ++misses;
      }
      j->allocated = have;
      return true;
    }
  }

  void Release(const Key& key, wg_handle* h, bool ok) {
    if (!h) return;
    if (!ok) {
      wg_destroy(h);
      return;
    }
    idle.emplace_back(key, h);
    while (int(idle.size()) > max_idle) {
      wg_destroy(idle.front().second);
      idle.erase(idle.begin());
      ++evictions;
    }
  }

  void Worker();
};

namespace {

// Grow-only device buffers of one worker, on its stream.
struct WorkerBuffers {
  hipStream_t s = nullptr;
  float* residual = nullptr;
  float* img = nullptr;
  float* model = nullptr;
  uint32_t* peak = nullptr;
  size_t n_res = 0, n_img = 0;

  void Ensure(size_t res_floats, size_t img_floats) {
    if (res_floats > n_res) {
      if (residual) WG_HIP(hipFreeAsync(residual, s));
      residual = nullptr;
      WG_HIP(hipMallocAsync(reinterpret_cast<void**>(&residual), res_floats * 4, s));
      n_res = res_floats;
    }
    if (img_floats > n_img) {
      if (img) WG_HIP(hipFreeAsync(img, s));
      if (model) WG_HIP(hipFreeAsync(model, s));
      img = model = nullptr;
      WG_HIP(hipMallocAsync(reinterpret_cast<void**>(&img), img_floats * 4, s));
      WG_HIP(hipMallocAsync(reinterpret_cast<void**>(&model), img_floats * 4, s));
      n_img = img_floats;
    }
    if (!peak) WG_HIP(hipMallocAsync(reinterpret_cast<void**>(&peak), 4, s));
  }
  size_t Bytes() const { return 4 * (n_res + 2 * n_img) + (peak ? 4 : 0); }
  ~WorkerBuffers() {
    if (!s) return;
    for (void* p : {static_cast<void*>(residual), static_cast<void*>(img),
                    static_cast<void*>(model), static_cast<void*>(peak)})
      if (p) (void)hipFreeAsync(p, s);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
};

// One request on its handle: prepare, then `cycles` major cycles.
void RunJob(Job& j, WorkerBuffers& wb, std::mutex& mu) {
  const wg_request& r = j.req;
  wg_handle* h = j.handle;
  const size_t n_vis = r.n_rows * r.n_chan, n2 = r.npix * r.npix;
  auto report = [&] {
    size_t b = 0;
    Check(wgx_device_bytes(h, &b));
    std::lock_guard<std::mutex> lk(mu);
    j.allocated = b + wb.Bytes();
  };
  Check(wg_prepare_visibilities(h, r.n_rows, r.n_chan, r.uvw_dev, j.freq.data()));
  wb.Ensure(2 * n_vis, n2);
  WG_HIP(hipMemcpyAsync(wb.residual, r.vis_dev, n_vis * 8, hipMemcpyDeviceToDevice, wb.s));
  WG_HIP(hipStreamSynchronize(wb.s));
  report();
  for (int c = 0; c < r.cycles; ++c) {
    Check(wg_initialize_inversion(h));
    Check(wg_add_inversion_prepared(h, wb.residual));
    Check(wg_finalize_image(h, 1.0 / double(n_vis)));
    Check(wg_real_image_device(h, wb.img));
    // the peak word is cleared, then PeakKernel and ModelKernel run in
    // order on the worker's stream
    WG_HIP(hipMemsetAsync(wb.peak, 0, 4, wb.s));
    PeakKernel<<<GridOf(n2), 256, 0, wb.s>>>(wb.img, n2, wb.peak);
    ModelKernel<<<GridOf(n2), 256, 0, wb.s>>>(wb.img, n2, wb.peak, wb.model);
    WG_HIP(hipGetLastError());
    WG_HIP(hipStreamSynchronize(wb.s));
    Check(wg_initialize_prediction_device(h, wb.model));
    Check(wg_predict_prepared(h, r.vis_out_dev));
    if (c == 0) report();
    if (c + 1 < r.cycles) {
      ResidualKernel<<<GridOf(2 * n_vis), 256, 0, wb.s>>>(r.vis_dev, r.vis_out_dev, 2 * n_vis,
                                                          wb.residual);
      WG_HIP(hipGetLastError());
      WG_HIP(hipStreamSynchronize(wb.s));
    }
  }
  if (r.residual_dev)
    WG_HIP(hipMemcpyAsync(r.residual_dev, wb.residual, n_vis * 8, hipMemcpyDeviceToDevice, wb.s));
  if (r.image_host)
    WG_HIP(hipMemcpyAsync(r.image_host, wb.img, n2 * 4, hipMemcpyDeviceToHost, wb.s));
  WG_HIP(hipStreamSynchronize(wb.s));
}

}  // namespace

void wg_server::Worker() {
  (void)hipSetDevice(device);
  WorkerBuffers wb;
  if (hipStreamCreate(&wb.s) != hipSuccess) wb.s = nullptr;
  for (;;) {
    std::shared_ptr<Job> j;
    {
      std::unique_lock<std::mutex> lk(mu);
      for (;;) {
        if (!queue.empty()) {
          bool ok = false;
          try {
            ok = wb.s && TryAdmit(queue.front());
          } catch (const std::exception& e) {
            // the head cannot be admitted (a handle could not be created):
            // it fails, the queue goes on
            auto f = queue.front();
            queue.pop_front();
            f->error = e.what();
            f->t_start = f->t_done = Now();
            f->done = true;
            cv.notify_all();
            continue;
          }
          if (ok) {
            j = queue.front();
            queue.pop_front();
            running.push_back(j);
            peak_running = std::max<uint64_t>(peak_running, running.size());
            break;
          }
          ++waits;
        } else if (stopping) {
          return;
        }
        // re-measure free memory at least every 10 ms while the head waits
        cv.wait_for(lk, std::chrono::milliseconds(10));
      }
      j->t_start = Now();
    }
    bool ok = true;
    try {
      RunJob(*j, wb, mu);
    } catch (const std::exception& e) {
      ok = false;
      j->error = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      j->t_done = Now();
      running.erase(std::find(running.begin(), running.end(), j));
      Release(Key{j->req.npix, j->req.pixel_size, j->req.epsilon}, j->handle, ok);
      j->handle = nullptr;
      j->done = true;
      ++completed;
    }
    cv.notify_all();
  }
}

namespace wg {

template <typename F>
int ServerGuard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    SetLastError(e.what());
    return 1;
  }
}

}  // namespace wg

extern "C" {

int wg_server_create(int slots, size_t reserve_bytes, int max_idle_handles, wg_server** out) {
  return wg::ServerGuard([&] {
    WG_CHECK(out && slots >= 1 && slots <= 64 && max_idle_handles >= 0,
             "wg_server_create: slots must be in [1, 64] and max_idle_handles >= 0");
    auto s = std::make_unique<wg_server>();
    s->slots = slots;
    s->reserve = reserve_bytes;
    s->max_idle = max_idle_handles;
    WG_HIP(hipGetDevice(&s->device));
    s->t0 = Clock::now();
    for (int i = 0; i < slots; ++i) s->workers.emplace_back([p = s.get()] { p->Worker(); });
    *out = s.release();
  });
}

int wg_server_submit(wg_server* s, const wg_request* req, uint64_t* id) {
  return wg::ServerGuard([&] {
    WG_CHECK(s && req && id, "null argument");
    const wg_request& r = *req;
    WG_CHECK(r.npix >= 8 && r.npix % 2 == 0 && r.pixel_size > 0 && r.epsilon > 0 && r.cycles >= 1,
             "request: npix must be even and >= 8, pixel_size and epsilon positive, cycles >= 1");
    WG_CHECK(r.n_rows > 0 && r.n_chan > 0 && r.uvw_dev && r.freq && r.vis_dev && r.vis_out_dev,
             "request: rows, channels, uvw, frequencies and visibility buffers are required");
    auto j = std::make_shared<Job>();
    j->req = r;
    j->freq.assign(r.freq, r.freq + r.n_chan);
    j->need = Estimate(r);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      WG_CHECK(!s->stopping, "the server is shutting down");
      j->id = s->next_id++;
      j->t_submit = s->Now();
      s->jobs[j->id] = j;
      s->queue.push_back(j);
      *id = j->id;
    }
    s->cv.notify_all();
  });
}

int wg_server_wait(wg_server* s, uint64_t id, double* times3) {
  return wg::ServerGuard([&] {
    WG_CHECK(s, "null server");
    std::shared_ptr<Job> j;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      auto it = s->jobs.find(id);
      WG_CHECK(it != s->jobs.end(), "unknown request id (or already waited for)");
      j = it->second;
      s->cv.wait(lk, [&] { return j->done; });
      s->jobs.erase(it);
    }
    if (times3) {
      times3[0] = j->t_submit;
      times3[1] = j->t_start;
      times3[2] = j->t_done;
    }
    WG_CHECK(j->error.empty(), "request " + std::to_string(id) + " failed: " + j->error);
  });
}

int wg_server_stats(wg_server* s, uint64_t* out8) {
  return wg::ServerGuard([&] {
    WG_CHECK(s && out8, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out8[0] = s->completed;
    out8[1] = s->peak_running;
    out8[2] = s->hits;
    out8[3] = s->misses;
    out8[4] = s->evictions;
    out8[5] = s->min_free == SIZE_MAX ? 0 : s->min_free;
    out8[6] = s->waits;
    out8[7] = s->idle.size();
  });
}

int wg_server_destroy(wg_server* s) {
  return wg::ServerGuard([&] {
    if (!s) return;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stopping = true;
    }
    s->cv.notify_all();
    for (auto& t : s->workers) t.join();
    for (auto& e : s->idle) wg_destroy(e.second);
    s->idle.clear();
    delete s;
  });
}

}  // extern "C"
