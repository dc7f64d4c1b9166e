// Engine: plan -> keys -> stable radix sort -> bucket offsets -> sorted
// coordinates -> work items (Prepare), and the plane-slab loops of the
// invert and predict pipelines.  All device work is issued on the engine's
// own stream (default priority).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "device.hpp"
#include "engine.hpp"

namespace wg {

void BuildItems(const DevPlan& dp, const uint32_t* off, int Sg, int nsub, uint32_t* counts,
                uint32_t* pos, uint32_t* scan_scratch, WorkItem* items,
                unsigned long long* sub_vis, hipStream_t s, bool emit);

DevPlan ToDevPlan(const Plan& p) {
  DevPlan d{};
  d.ufac = p.ufac;
  d.vfac = p.vfac;
  d.half_nu = 0.5 * p.nu;
  d.half_nv = 0.5 * p.nv;
  d.w_lo = p.w_lo;
  d.inv_dw = p.inv_dw;
  d.wofs = double(p.W / 2 - 1);
  d.l_shift = p.img.l_shift;
  d.m_shift = p.img.m_shift;
  d.zc = p.zc;
  d.dw = p.dw;
  d.dx_rad = p.img.dx;
  d.dy_rad = p.img.dy;
  EsPolyCoefficients(p.W, p.beta, d.es_poly);
  d.W = p.W;
  d.T = p.T;
  d.tshift = (p.T & (p.T - 1)) == 0 ? __builtin_ctz(unsigned(p.T)) : -1;
  d.ntu = p.ntu;
  d.ntv = p.ntv;
  d.P0 = p.P0;
  d.P = p.P;
  d.nu = p.nu;
  d.nv = p.nv;
  d.box_u0 = p.box_u0;
  d.box_v0 = p.box_v0;
  d.wrap_u = p.wrap_u ? 1 : 0;
  d.wrap_v = p.wrap_v ? 1 : 0;
  d.BU = p.BU;
  d.BV = p.BV;
  d.box_pitch = size_t(p.BU);
  d.box_plane = size_t(p.BU) * size_t(p.BV);
  d.nx = p.img.nx;
  d.ny = p.img.ny;
  d.nkeys = p.nkeys;
  d.nbuckets = p.nbuckets;
  d.key_shift = p.key_shift;
  d.split = p.split ? 1 : 0;
  d.key_T = p.key_T;
  return d;
}

// ------------------------------------------------------------ launch queries

namespace {
std::mutex g_launch_mu;
std::map<const void*, size_t> g_lds_allowed;
std::map<std::tuple<const void*, int, size_t>, int> g_resident;
}  // namespace

void AllowDynamicLds(const void* kernel, size_t bytes) {
  if (bytes <= 48 * 1024) return;
  std::lock_guard<std::mutex> lk(g_launch_mu);
  size_t& have = g_lds_allowed[kernel];
  if (have >= bytes) return;
  WG_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
  have = bytes;
}

int ResidentBlocksPerCU(const void* kernel, int block, size_t lds) {
  std::lock_guard<std::mutex> lk(g_launch_mu);
  auto key = std::make_tuple(kernel, block, lds);
  auto it = g_resident.find(key);
  if (it != g_resident.end()) return it->second;
  int per_cu = 0;
  WG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds));
  g_resident[key] = per_cu;
  return per_cu;
}

int DeviceCUs() {
  static const int n = [] {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return cus;
  }();
  return n;
}

// ------------------------------------------------------------ DeviceArena

DeviceArena::~DeviceArena() { ReleaseAll(); }

void DeviceArena::SetStream(hipStream_t s) {
  stream_ = s;
  int dev = 0, pools = 0;
  async_ = hipGetDevice(&dev) == hipSuccess &&
           hipDeviceGetAttribute(&pools, hipDeviceAttributeMemoryPoolsSupported, dev) ==
               hipSuccess &&
           pools != 0;
}

void DeviceArena::Free(void* p) {
  if (!p) return;
  if (async_)
    (void)hipFreeAsync(p, stream_);
  else
    (void)hipFree(p);
}

void* DeviceArena::Get(const std::string& name, size_t bytes) {
  Slot& s = slots_[name];
  if (bytes == 0) bytes = 256;
  if (s.bytes < bytes) {
    if (s.ptr) ++generation_;  // a new slot leaves existing pointers valid
    Free(s.ptr);
    s.ptr = nullptr;
    s.bytes = 0;
    // round up so that slowly growing requests do not reallocate every call
    size_t b = bytes + bytes / 8 + 256;
    if (async_)
      WG_HIP(hipMallocAsync(&s.ptr, b, stream_));
    else
      WG_HIP(hipMalloc(&s.ptr, b));
    s.bytes = b;
  }
  return s.ptr;
}

void DeviceArena::Release(const std::string& name) {
  auto it = slots_.find(name);
  if (it == slots_.end()) return;
  ++generation_;
  Free(it->second.ptr);
  slots_.erase(it);
}

void DeviceArena::ReleaseAll() {
  if (!slots_.empty()) ++generation_;
  for (auto& kv : slots_) Free(kv.second.ptr);
  slots_.clear();
}

size_t DeviceArena::Allocated() const {
  size_t t = 0;
  for (const auto& kv : slots_) t += kv.second.bytes;
  return t;
}

// ------------------------------------------------------------ Engine

Engine::Engine(const ImageGeometry& g, const PlanOptions& opt) : geom_(g), opt_(opt) {
  WG_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  arena_.SetStream(stream_);
}

Engine::~Engine() {
  if (stream_) {
    arena_.ReleaseAll();  // stream-ordered frees, before the stream goes
    (void)hipStreamSynchronize(stream_);
    (void)hipStreamDestroy(stream_);
  }
}

const float2* Engine::Twiddles(int n) {
  auto it = twiddles_.find(n);
  if (it != twiddles_.end()) return it->second;
  const std::vector<float2> h = FFTTwiddleTable(n);
  float2* d = static_cast<float2*>(
      arena_.Get("twiddles." + std::to_string(n), h.size() * sizeof(float2)));
  WG_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice, stream_));
  WG_HIP(hipStreamSynchronize(stream_));
  twiddles_[n] = d;
  return d;
}

void Engine::CollectTimes() {
  for (PendingTime& t : pending_) {
    float ms = 0;
    if (hipEventSynchronize(t.b) == hipSuccess && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess)
      *t.dst += ms;
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  pending_.clear();
}

void Engine::Prepare(VisSet& vs, const std::string& tag, size_t n_rows, size_t n_chan,
                     const double* uvw_dev, const double* freq_host, const float2* vis_dev) {
  EventTimer tm(*this, &times_.plan);
  WG_CHECK(n_chan > 0, "n_chan must be positive");
  const size_t n_vis = n_rows * n_chan;
  // flat indices live in the low 31 bits of the sorted permutation; the
  // drop-in calls split larger inputs into row chunks below this
  WG_CHECK(n_vis < 0x80000000ull,
           "a prepared visibility set holds at most 2^31 - 1 visibilities "
           "(wg_add_inversion_data / wg_predict_visibilities split larger calls)");
  vs.tag = tag;
  vs.n_rows = n_rows;
  vs.n_chan = n_chan;
  vs.n_vis = n_vis;
  vs.skip_zero = vis_dev != nullptr;
  vs.uvw = const_cast<double*>(uvw_dev);

  std::vector<double> scale(n_chan);
  double smin = 1e300, smax = -1e300;
  for (size_t c = 0; c < n_chan; ++c) {
    WG_CHECK(freq_host[c] > 0, "frequencies must be positive");
    scale[c] = freq_host[c] / kSpeedOfLight;
    smin = std::min(smin, scale[c]);
    smax = std::max(smax, scale[c]);
  }
  vs.scale = static_cast<double*>(arena_.Get(tag + ".scale", n_chan * sizeof(double)));
  WG_HIP(hipMemcpyAsync(vs.scale, scale.data(), n_chan * sizeof(double),
                        hipMemcpyHostToDevice, stream_));

  VisExtent ext{0, 0, 0, 0, 0, 0};
  if (n_rows > 0) {
    double* part = static_cast<double*>(arena_.Get("extents", (1024 * 6 + 6) * sizeof(double)));
    LaunchExtents(uvw_dev, n_rows, smin, smax, part, part + 1024 * 6, stream_);
    double e[6];
    WG_HIP(hipMemcpyAsync(e, part + 1024 * 6, sizeof(e), hipMemcpyDeviceToHost, stream_));
    WG_HIP(hipStreamSynchronize(stream_));
    ext = VisExtent{e[0], e[1], e[2], e[3], e[4], e[5]};
    for (double v : e) WG_CHECK(std::isfinite(v), "non-finite uvw coordinates");
  }
  PlanOptions o = opt_;
  o.nvis_hint = (long long)n_vis;
  vs.plan = MakePlan(geom_, ext, o);
  WG_CHECK(vs.plan.P0 < 32768, "too many w planes");
  vs.dp = ToDevPlan(vs.plan);
  const DevPlan& dp = vs.dp;

  // ---- row renumbering ----
  // Rows are renumbered in the order of their key at the middle channel
  // (stable), i.e. in uv order, and copied padded to 32 bytes with the
  // caller row in the pad word: the coordinates kernel then reads the rows
  // of neighbouring sorted visibilities from neighbouring addresses instead
  // of at random, and each row's caller index from the same cache line.
  // the renumbered rows' caller rows stay with the set (the row maxima of
  // its inversions read the rows through them)
  uint32_t* rperm = static_cast<uint32_t*>(arena_.Get(tag + ".rows", n_rows * 4));
  double* uvw4 = static_cast<double*>(arena_.Get("uvw4", n_rows * 4 * sizeof(double)));
  uint32_t* scratch = static_cast<uint32_t*>(
      arena_.Get("sort.scratch", RadixSortScratchWords(std::max(n_vis, n_rows)) * 4));
  if (n_rows) {
    uint32_t* rk = static_cast<uint32_t*>(arena_.Get("rows.keys", n_rows * 4));
    uint32_t* rk2 = static_cast<uint32_t*>(arena_.Get("rows.keys2", n_rows * 4));
    uint32_t* rp2 = static_cast<uint32_t*>(arena_.Get(tag + ".rows2", n_rows * 4));
    LaunchKeys(dp, uvw_dev, vs.scale + n_chan / 2, n_rows, 1, nullptr, rk, nullptr, stream_);
    if (DeviceRadixSort(rk, rperm, rk2, rp2, n_rows, vs.plan.key_bits, scratch, stream_, true))
      rperm = rp2;
    LaunchPadUvw(uvw_dev, rperm, rperm, n_rows, uvw4, stream_);
  }
  vs.rows = rperm;

  // ---- keys + stable sort; the sort's last pass writes the sorted
  // coordinates and phasors and turns renumbered elements into the caller's
  // flat indices with the flip bit ----
  uint32_t* keys = static_cast<uint32_t*>(arena_.Get(tag + ".keys", n_vis * 4));
  uint32_t* vals = static_cast<uint32_t*>(arena_.Get(tag + ".perm", n_vis * 4));
  vs.coord = static_cast<float4*>(arena_.Get(tag + ".coord", n_vis * sizeof(float4)));
  vs.phasor = static_cast<float2*>(arena_.Get(tag + ".phasor", n_vis * sizeof(float2)));
  // sets that the matrix-core kernel grids in one sub-slab keep each sorted
  // visibility's renumbered row for the per-item bounds
  const bool fused = dp.W <= 8 && dp.P <= vs.plan.Sg;
  vs.vrow = fused ? static_cast<uint32_t*>(arena_.Get(tag + ".vrow", n_vis * 4)) : nullptr;
  vs.rowmax_of = nullptr;
  if (n_vis) {
    // the sort's first pass generates the index values itself.  When the
    // keys kernel reads the visibilities (zero skipping), it also leaves
    // the max |V| of every row for the Invert that follows
    uint32_t* rowmax_bits = nullptr;
    if (vis_dev) {
      rowmax_bits = static_cast<uint32_t*>(arena_.Get("grid.rowmax", n_rows * sizeof(float)));
      vs.rowmax_of = vis_dev;
    }
    LaunchKeysRenumbered(dp, uvw4, vs.scale, n_rows, n_chan, vis_dev, keys, stream_,
                         rowmax_bits);
    uint32_t* keys2 = static_cast<uint32_t*>(arena_.Get(tag + ".keys2", n_vis * 4));
    uint32_t* vals2 = static_cast<uint32_t*>(arena_.Get(tag + ".perm2", n_vis * 4));
    const SortCoords co{dp,        uvw4,      vs.scale, n_chan, 1.0 / double(n_chan),
                        vs.coord, vs.phasor, vs.vrow};
    if (DeviceRadixSort(keys, vals, keys2, vals2, n_vis, vs.plan.key_bits, scratch, stream_,
                        true, &co)) {
      keys = keys2;
      vals = vals2;
    }
  }
  vs.perm = vals;
  vs.off = static_cast<uint32_t*>(arena_.Get(tag + ".off", (size_t(dp.nbuckets) + 1) * 4));
  DeviceBucketOffsets(keys, n_vis, dp.nbuckets, uint32_t(dp.key_T) * uint32_t(dp.key_T), vs.off,
                      stream_);
  uint32_t nvalid = 0;
  WG_HIP(hipMemcpyAsync(&nvalid, vs.off + dp.nbuckets, 4, hipMemcpyDeviceToHost, stream_));
  WG_HIP(hipStreamSynchronize(stream_));
  vs.n_valid = nvalid;
  BuildWorkItems(vs);
  BuildCorrection(vs);
}

void Engine::BuildWorkItems(VisSet& vs) {
  const DevPlan& dp = vs.dp;
  const int Sg = vs.plan.Sg;
  vs.nsub = (dp.P + Sg - 1) / Sg;
  const size_t n = size_t(dp.ntu) * dp.ntv * vs.nsub;
  uint32_t* counts = static_cast<uint32_t*>(arena_.Get("items.counts", (n + 1) * 4));
  vs.item_pos = static_cast<uint32_t*>(arena_.Get(vs.tag + ".item_pos", (n + 1) * 4));
  uint32_t* scr = static_cast<uint32_t*>(arena_.Get("items.scan", ScanScratchWords(n) * 4));
  // visibility passes per sub-slab (for the kernels' byte accounting)
  unsigned long long* dvisp =
      static_cast<unsigned long long*>(arena_.Get("items.visp", vs.nsub * 8));
  LaunchFillZero(dvisp, vs.nsub * 8, stream_);
  BuildItems(dp, vs.off, Sg, vs.nsub, counts, vs.item_pos, scr, nullptr, dvisp, stream_, false);
  // item counts at the sub-slab boundaries (every ntiles-th scan entry),
  // gathered on the device and copied with a plain 1-D copy
  const size_t ntiles = size_t(dp.ntu) * dp.ntv;
  std::vector<uint32_t> bounds(vs.nsub + 1);
  std::vector<unsigned long long> visp(vs.nsub);
  uint32_t* dbounds = static_cast<uint32_t*>(arena_.Get("items.bounds", (vs.nsub + 1) * 4));
  LaunchStridedGather(vs.item_pos, ntiles, vs.nsub + 1, dbounds, stream_);
  WG_HIP(hipMemcpyAsync(bounds.data(), dbounds, (vs.nsub + 1) * 4, hipMemcpyDeviceToHost,
                        stream_));
  WG_HIP(hipMemcpyAsync(visp.data(), dvisp, vs.nsub * 8, hipMemcpyDeviceToHost, stream_));
  WG_HIP(hipStreamSynchronize(stream_));
  vs.sub_vis.assign(visp.begin(), visp.end());
  const uint32_t total = bounds[vs.nsub];
  vs.n_items = total;
  vs.max_items_per_sub = 0;
  vs.sub_items.resize(vs.nsub);
  for (int g = 0; g < vs.nsub; ++g) {
    vs.sub_items[g] = bounds[g + 1] - bounds[g];
    vs.max_items_per_sub = std::max<size_t>(vs.max_items_per_sub, vs.sub_items[g]);
  }
  vs.items = static_cast<WorkItem*>(arena_.Get(vs.tag + ".items", (total + 1) * sizeof(WorkItem)));
  if (total)
    BuildItems(dp, vs.off, Sg, vs.nsub, counts, vs.item_pos, scr, vs.items, nullptr, stream_,
               true);
}

void Engine::BuildCorrection(VisSet& vs) {
  const Plan& p = vs.plan;
  const int nx = p.img.nx, ny = p.img.ny;
  // The correction depends only on the image geometry, the kernel shape, the
  // padded grid and the w axis scale/centre: reuse it while those match.
  char key[512];
  std::snprintf(key, sizeof(key), "%d %d %.17g %.17g %.17g %.17g %d %.17g %d %d %.17g %.17g",
                nx, ny, p.img.dx, p.img.dy, p.img.l_shift, p.img.m_shift, p.W, p.beta, p.nu,
                p.nv, p.dw, p.zc);
  vs.corr = static_cast<float*>(arena_.Get(vs.tag + ".corr", size_t(nx) * ny * 4));
  if (vs.corr_key == key) return;
  // Phi(t) = sum_q a_q cos(b_q t): Gauss-Legendre in theta with x = sin(theta)
  const int nq = 64;
  std::vector<double> gx(nq), gw(nq);
  for (int i = 0; i < (nq + 1) / 2; ++i) {
    // nodes by Newton iteration on the Legendre polynomial
    double z = std::cos(M_PI * (i + 0.75) / (nq + 0.5)), pp = 0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1, p2 = 0;
      for (int j = 1; j <= nq; ++j) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = nq * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1e-15) break;
    }
    gx[i] = -z;
    gx[nq - 1 - i] = z;
    gw[i] = gw[nq - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
  }
  std::vector<double> q(2 * nq);
  for (int k = 0; k < nq; ++k) {
    const double th = 0.5 * M_PI * gx[k];
    const double c = std::cos(th);
    q[k] = 0.5 * p.W * 0.5 * M_PI * gw[k] * std::exp(p.beta * (c - 1.0)) * c;
    q[nq + k] = M_PI * p.W * std::sin(th);
  }
  double* dq = static_cast<double*>(arena_.Get(vs.tag + ".quad", 2 * nq * 8));
  double* dpu = static_cast<double*>(arena_.Get(vs.tag + ".phiu", nx * 8));
  double* dpv = static_cast<double*>(arena_.Get(vs.tag + ".phiv", ny * 8));
  WG_HIP(hipMemcpyAsync(dq, q.data(), 2 * nq * 8, hipMemcpyHostToDevice, stream_));
  LaunchPhiTables(vs.dp, dq, dq + nq, nq, dpu, dpv, stream_);
  LaunchCorrection(vs.dp, dpu, dpv, dq, dq + nq, nq, vs.corr, stream_);
  // the host vector must outlive the async copy
  WG_HIP(hipStreamSynchronize(stream_));
  vs.corr_key = key;
}

Engine::LineFFTs Engine::PlaneFFTs(const DevPlan& dp, int nplanes) {
  LineFFTs f;
  f.ru = LongFFTRadix(dp.nu);
  f.rv = LongFFTRadix(dp.nv);
  WG_CHECK(f.ru > 0 && f.rv > 0, "padded grid side not split by 2..5 into LDS FFT lines");
  f.fu = MakeFFTDesc(dp.nu / f.ru, Twiddles(dp.nu / f.ru));
  f.fv = MakeFFTDesc(dp.nv / f.rv, Twiddles(dp.nv / f.rv));
  if (f.ru > 1 || f.rv > 1) {
    // subsequence spectra of one slab: the row pass's BV lines of nu or the
    // column pass's nx lines of nv, per plane
    const size_t per_plane = std::max(size_t(dp.BV) * size_t(dp.nu), size_t(dp.nx) * size_t(dp.nv));
    f.tmp = static_cast<float2*>(arena_.Get("fft.long", size_t(nplanes) * per_plane * sizeof(float2)));
  }
  return f;
}

void Engine::Invert(VisSet& vs, const float2* vis_dev, double* img_dev) {
  const DevPlan& dp = vs.dp;
  const Plan& p = vs.plan;
  if (vs.n_valid == 0) return;
  float2* vsorted = nullptr;
  // work-item queue of the gridding kernel
  uint32_t* words = static_cast<uint32_t*>(arena_.Get("grid.words", 4 * sizeof(uint32_t)));
  // one gridding pass per visibility (a single sub-slab): the matrix-core
  // kernel gathers through the permutation itself
  const bool fused = dp.W <= 8 && vs.nsub == 1;
  // the matrix-core gridders' per-item scales
  float2* bounds = nullptr;
  {
    EventTimer tm(*this, &times_.sort);
    if (dp.W <= 8)
      bounds = static_cast<float2*>(arena_.Get("grid.bounds", vs.n_items * sizeof(float2)));
    if (fused) {
      // the row maxima of these visibilities are already there when the
      // Prepare just before read them
      float* rowmax = static_cast<float*>(arena_.Get("grid.rowmax", vs.n_rows * sizeof(float)));
      if (vs.rowmax_of != vis_dev) {
        // in caller row order, streaming through the visibilities, then
        // gathered into renumbered row order
        float* rowmax_c =
            static_cast<float*>(arena_.Get("grid.rowmax_c", vs.n_rows * sizeof(float)));
        LaunchRowMax(vis_dev, nullptr, vs.n_rows, vs.n_chan,
                     reinterpret_cast<uint32_t*>(rowmax_c), stream_);
        LaunchGatherRows(rowmax_c, vs.rows, vs.n_rows, rowmax, stream_);
      }
      LaunchItemBounds(vs.items, vs.n_items, nullptr, vs.vrow, rowmax, bounds, stream_);
    } else {
      vsorted = static_cast<float2*>(arena_.Get("vis_sorted", vs.n_valid * sizeof(float2)));
      LaunchGatherVis(vs.perm, vs.phasor, vs.n_valid, vis_dev, vsorted, stream_);
      if (bounds)
        LaunchItemBounds(vs.items, vs.n_items, vsorted, nullptr, nullptr, bounds, stream_);
    }
  }
  vs.rowmax_of = nullptr;
  const size_t npix = size_t(dp.nx) * dp.ny;
  float* dirty = static_cast<float*>(arena_.Get("dirty", npix * 4));
  LaunchFillZero(dirty, npix * 4, stream_);
  const LineFFTs ft = PlaneFFTs(dp, p.Sf);
  const size_t plane_box = size_t(dp.BU) * dp.BV;
  const size_t plane_lines = size_t(dp.BV) * dp.nx;
  float2* box = static_cast<float2*>(arena_.Get("box", p.Sf * plane_box * sizeof(float2)));
  float2* rbuf = static_cast<float2*>(arena_.Get("rbuf", p.Sf * plane_lines * sizeof(float2)));
  float2* cbuf = static_cast<float2*>(arena_.Get("cbuf", p.Sf * plane_lines * sizeof(float2)));
  float2* scratch = static_cast<float2*>(
      arena_.Get("grid.scratch", GridScratchBytes(dp, p.Sg, vs.max_items_per_sub)));
  for (int f0 = 0; f0 < p.P; f0 += p.Sf) {
    const int np = std::min(p.Sf, p.P - f0);
    {
      // every box cell of the slab's planes is written by the gather pass
      EventTimer tm(*this, &times_.grid);
      const size_t E = size_t(dp.T + dp.W - 1);
      for (int g = f0 / p.Sg; g * p.Sg < f0 + np; ++g) {
        const int sub_lo = g * p.Sg;
        const int sub_n = std::min(p.Sg, p.P - sub_lo);
        {
          EventTimer tk(*this, &times_.grid_kernel_ms);
          LaunchGrid(dp, vs, vsorted, scratch, words, bounds, g, sub_lo, sub_n, stream_,
                     fused ? vis_dev : nullptr, fused ? vs.perm : nullptr);
        }
        if (timing) {
          // coordinates (16 B) and visibility (8 B) per pass -- fused: also
          // the permutation word (4 B) and phasor (8 B) -- and one subgrid
          // store per work item
          times_.grid_launches += 1;
          times_.grid_passes += vs.sub_vis[g];
          times_.grid_bytes += (fused ? 36.0 : 24.0) * vs.sub_vis[g] +
                               8.0 * double(sub_n) * E * E * double(vs.sub_items[g]);
        }
        LaunchGatherBox(dp, vs, box, scratch, g, sub_lo, sub_n, f0, stream_);
      }
    }
    EventTimer tm(*this, &times_.fft);
    if (ft.ru > 1)
      LaunchRowFFTInvertLong(dp, ft.fu, dp.nu, ft.ru, box, ft.tmp, rbuf, np, stream_);
    else
      LaunchRowFFTInvert(dp, ft.fu, box, rbuf, np, stream_);
    LaunchTranspose(rbuf, cbuf, dp.BV, dp.nx, np, stream_);
    if (ft.rv > 1)
      LaunchColFFTInvertLong(dp, ft.fv, dp.nv, ft.rv, cbuf, ft.tmp, dirty, f0, np, stream_);
    else
      LaunchColFFTInvert(dp, ft.fv, cbuf, dirty, f0, np, stream_);
  }
  LaunchAccumulateImage(dirty, vs.corr, img_dev, npix, stream_);
}

void Engine::Predict(VisSet& vs, const double* img_dev, float2* vis_out_dev) {
  const DevPlan& dp = vs.dp;
  const Plan& p = vs.plan;
  if (vs.n_valid == 0) return;
  // one degridding pass per visibility (a single sub-slab): the
  // matrix-core kernel writes the caller's output through the permutation
  const bool fused = dp.W <= 8 && vs.nsub == 1;
  float2* acc = vis_out_dev;
  if (!fused) {
    acc = static_cast<float2*>(arena_.Get("vis_acc", vs.n_valid * sizeof(float2)));
    LaunchFillZero(acc, vs.n_valid * sizeof(float2), stream_);
  }
  const LineFFTs ft = PlaneFFTs(dp, p.Sf);
  const size_t plane_box = size_t(dp.BU) * dp.BV;
  const size_t plane_lines = size_t(dp.BV) * dp.nx;
  float2* box = static_cast<float2*>(arena_.Get("box", p.Sf * plane_box * sizeof(float2)));
  float2* rbuf = static_cast<float2*>(arena_.Get("rbuf", p.Sf * plane_lines * sizeof(float2)));
  float2* cbuf = static_cast<float2*>(arena_.Get("cbuf", p.Sf * plane_lines * sizeof(float2)));
  uint32_t* queue = static_cast<uint32_t*>(arena_.Get("grid.words", 4 * sizeof(uint32_t)));
  for (int f0 = 0; f0 < p.P; f0 += p.Sf) {
    const int np = std::min(p.Sf, p.P - f0);
    {
      EventTimer tm(*this, &times_.fft);
      if (ft.rv > 1)
        LaunchColFFTPredictLong(dp, ft.fv, dp.nv, ft.rv, img_dev, vs.corr, ft.tmp, cbuf, f0, np,
                                stream_);
      else
        LaunchColFFTPredict(dp, ft.fv, img_dev, vs.corr, cbuf, f0, np, stream_);
      LaunchTranspose(cbuf, rbuf, dp.nx, dp.BV, np, stream_);
      if (ft.ru > 1)
        LaunchRowFFTPredictLong(dp, ft.fu, dp.nu, ft.ru, rbuf, ft.tmp, box, np, stream_);
      else
        LaunchRowFFTPredict(dp, ft.fu, rbuf, box, np, stream_);
    }
    EventTimer tm(*this, &times_.degrid);
    const size_t E = size_t(dp.T + dp.W - 1);
    for (int g = f0 / p.Sg; g * p.Sg < f0 + np; ++g) {
      const int sub_lo = g * p.Sg;
      const int sub_n = std::min(p.Sg, p.P - sub_lo);
      {
        EventTimer tk(*this, &times_.degrid_kernel_ms);
        LaunchDegrid(dp, vs, box, acc, queue, g, sub_lo, sub_n, f0, stream_,
                     fused ? vs.perm : nullptr);
      }
      if (timing) {
        // coordinates (16 B) read and accumulator (8 B) read + written per
        // pass -- fused: permutation word (4 B) and phasor (8 B) read and
        // the output (8 B) written -- and one subgrid load per work item
        times_.degrid_launches += 1;
        times_.degrid_passes += vs.sub_vis[g];
        times_.degrid_bytes += (fused ? 36.0 : 32.0) * vs.sub_vis[g] +
                               8.0 * double(sub_n) * E * E * double(vs.sub_items[g]);
      }
    }
  }
  if (!fused) {
    EventTimer tm(*this, &times_.sort);
    LaunchScatterVis(vs.perm, vs.phasor, vs.n_valid, acc, vis_out_dev, stream_);
  }
}

// ------------------------------------------------------------ autotuner

std::vector<Engine::TuneTrial> Engine::Autotune(size_t n_rows, size_t n_chan,
                                                const double* uvw_dev, const double* freq_host,
                                                size_t max_sample_vis) {
  WG_CHECK(n_rows > 0 && n_chan > 0, "autotune needs visibilities");
  const size_t n_vis = n_rows * n_chan;
  const size_t want = std::max<size_t>(n_chan, std::min(n_vis, std::max<size_t>(max_sample_vis, 1)));
  const size_t stride = std::max<size_t>(1, n_vis / want);
  const size_t ns = (n_rows + stride - 1) / stride;
  const double scale_up = double(n_rows) / double(ns);
  // strided row sample of uvw; the visibility values do not change the work
  double* suvw = static_cast<double*>(arena_.Get("tune.uvw", ns * 3 * sizeof(double)));
  LaunchStrideRows(uvw_dev, stride, ns, 3, suvw, stream_);
  float2* svis = static_cast<float2*>(arena_.Get("tune.vis", ns * n_chan * sizeof(float2)));
  LaunchFillZero(svis, ns * n_chan * sizeof(float2), stream_);
  float2* sout = static_cast<float2*>(arena_.Get("tune.out", ns * n_chan * sizeof(float2)));
  double* simg = static_cast<double*>(
      arena_.Get("tune.img", size_t(geom_.nx) * geom_.ny * sizeof(double)));
  LaunchFillZero(simg, size_t(geom_.nx) * geom_.ny * sizeof(double), stream_);

  // candidates: per support W, the smallest oversampling meeting epsilon
  // and two larger ones (fewer w planes, larger grids)
  std::vector<TuneTrial> cand;
  {
    std::map<int, std::vector<double>> by_w;
    for (const KernelShape& k : KernelTable()) {
      if (k.sigma < opt_.sigma_min - 1e-9 || k.sigma > opt_.sigma_max + 1e-9) continue;
      if (ShapeErrorEstimate(k) > geom_.epsilon) continue;
      by_w[k.W].push_back(k.sigma);
    }
    for (auto& kv : by_w) {
      std::sort(kv.second.begin(), kv.second.end());
      const double s0 = kv.second.front();
      for (double sg : kv.second)
        if (sg == s0 || std::fabs(sg - (s0 + 0.15)) < 1e-9 || std::fabs(sg - (s0 + 0.3)) < 1e-9) {
          TuneTrial t;
          t.W = kv.first;
          t.sigma = sg;
          t.T = 8;
          cand.push_back(t);
        }
    }
  }
  WG_CHECK(!cand.empty(), "no kernel shape reaches epsilon");

  const PlanOptions saved = opt_;
  const bool saved_timing = timing;
  const StageTimes saved_times = times_;
  auto run = [&](TuneTrial& t) {
    try {
      opt_ = saved;
      opt_.force_W = t.W;
      opt_.force_sigma = t.sigma;
      opt_.T = t.T;
      opt_.Sg = t.Sg;
      VisSet vs;
      timing = false;
      Prepare(vs, "tune", ns, n_chan, suvw, freq_host, nullptr);
      Invert(vs, svis, simg);
      Predict(vs, simg, sout);
      WG_HIP(hipStreamSynchronize(stream_));
      timing = true;
      times_ = StageTimes{};
      Prepare(vs, "tune", ns, n_chan, suvw, freq_host, nullptr);
      Invert(vs, svis, simg);
      Predict(vs, simg, sout);
      WG_HIP(hipStreamSynchronize(stream_));
      CollectTimes();
      const StageTimes& st = times_;
      t.fixed_ms = st.fft;
      t.vis_ms = st.plan + st.sort + st.grid + st.degrid;
      t.est_ms = t.vis_ms * scale_up + t.fixed_ms;
      t.Sg = vs.plan.Sg;
      t.P = vs.plan.P;
      t.nu = vs.plan.nu;
    } catch (const std::exception& e) {
      (void)hipStreamSynchronize(stream_);
      pending_.clear();
      t.est_ms = 1e300;
      t.error = e.what();
    }
  };
  for (TuneTrial& t : cand) run(t);
  std::stable_sort(cand.begin(), cand.end(),
                   [](const TuneTrial& a, const TuneTrial& b) { return a.est_ms < b.est_ms; });
  // the best two shapes also with 16-cell tiles
  const size_t base = cand.size();
  for (size_t i = 0; i < std::min<size_t>(2, base); ++i) {
    if (cand[i].est_ms >= 1e300) continue;
    TuneTrial t;
    t.W = cand[i].W;
    t.sigma = cand[i].sigma;
    t.T = 16;
    run(t);
    cand.push_back(t);
  }
  std::stable_sort(cand.begin(), cand.end(),
                   [](const TuneTrial& a, const TuneTrial& b) { return a.est_ms < b.est_ms; });
  timing = saved_timing;
  times_ = saved_times;
  opt_ = saved;
  WG_CHECK(cand.front().est_ms < 1e300, "every autotune candidate failed: " + cand.front().error);
  opt_.force_W = cand.front().W;
  opt_.force_sigma = cand.front().sigma;
  opt_.T = cand.front().T;
  opt_.Sg = 0;
  // the trials' buffers; the shared plane buffers are re-sized by the next call
  for (const char* n : {"box", "rbuf", "cbuf", "grid.scratch",
                        "tune.uvw", "tune.vis", "tune.out", "tune.img", "tune.keys", "tune.perm",
                        "tune.keys2", "tune.perm2", "tune.coord", "tune.phasor", "tune.off",
                        "tune.items", "tune.item_pos", "tune.corr", "tune.scale", "tune.quad",
                        "tune.phiu", "tune.phiv"})
    arena_.Release(n);
  return cand;
}

}  // namespace wg
