// Device engine: owns a HIP stream and a cache of device buffers, and runs
// the invert (grid) and predict (degrid) pipelines for one gridder handle.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <map>
#include <set>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "device.hpp"

namespace wg {

// ---- primitives implemented in sort.hip ----
void DeviceExclusiveScan(const uint32_t* in, uint32_t* out, size_t n,
                         uint32_t* scratch, hipStream_t s, uint32_t* total);
size_t ScanScratchWords(size_t n);
size_t RadixSortScratchWords(size_t n);
// Stable sort of (keys, vals) by the low key_bits; keys_alt / vals_alt are
// the ping-pong buffers (all four 16-byte aligned).  index_vals: the input
// values are the element indices and vals is not read.  Returns true when
// the result is in the alt pair.
// Last-pass epilogue of the visibility sort (DeviceRadixSort's coords):
// for sorted entry p of renumbered element e = r * n_chan + ch (row r of
// the padded renumbered uvw4), coord[p] = (fu, fv, fw, packed cell / plane
// / flip), phasor[p] = the shift phasor exp(-2 pi i (u l + v m + w zc)), and
// the value becomes the caller's flat index | kPermFlip.  coord and phasor
// hold n entries (those of invalid keys, sorted last, are not used).
struct SortCoords {
  DevPlan dp;
  const double* uvw4;
  const double* scale;
  size_t n_chan;
  double inv_chan;  // 1 / n_chan
  float4* coord;
  float2* phasor;
  uint32_t* vrow;  // renumbered row of each sorted element (or null)
};
bool DeviceRadixSort(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                     size_t n, int key_bits, uint32_t* scratch, hipStream_t s,
                     bool index_vals = false, const SortCoords* coords = nullptr);
// off[b] = first sorted position with key >= b * stride, b in [0, nbuckets].
void DeviceBucketOffsets(const uint32_t* sorted_keys, size_t n, uint32_t nbuckets,
                         uint32_t stride, uint32_t* off, hipStream_t s);

// One LDS FFT line length: Stockham stage radices and the stage index
// divisors, with the device twiddle table.
struct FFTDesc {
  int n;
  int nst;
  int rad[16];
  uint32_t magic[16];  // floor(2^32 / Ns) + 1 per stage (Ns: product of earlier radices)
  const float2* tw;    // device copy of FFTTwiddleTable(n)
};

// A work item: the visibilities [begin, end) (sorted order) of one uv tile
// that touch the planes of one gridding sub-slab.
struct WorkItem {
  uint32_t tile;
  uint32_t begin;
  uint32_t end;
};

// Grow-only cache of named device allocations, stream-ordered on the
// owning engine's stream (hipMallocAsync / hipFreeAsync): growing or
// releasing a slot never synchronizes the device, so handles that serve
// requests side by side do not stall each other.  Every use of an arena
// buffer is on that same stream.
class DeviceArena {
 public:
  ~DeviceArena();
  void SetStream(hipStream_t s);
  void* Get(const std::string& name, size_t bytes);
  void Release(const std::string& name);
  void ReleaseAll();
  size_t Allocated() const;
  // Bumped whenever a slot is freed or moved to a new allocation: device
  // pointers taken from the arena (for example by a captured HIP graph)
  // stay valid exactly while the generation is unchanged.
  uint64_t Generation() const { return generation_; }

 private:
  struct Slot {
    void* ptr = nullptr;
    size_t bytes = 0;
  };
  void Free(void* p);
  std::map<std::string, Slot> slots_;
  hipStream_t stream_ = nullptr;
  bool async_ = false;  // device supports stream-ordered allocation
  uint64_t generation_ = 0;
};

// Per-call timing of the pipeline stages (ms, measured with HIP events).
struct StageTimes {
  double plan = 0, sort = 0, grid = 0, fft = 0, degrid = 0, total = 0;
  // the gridding and degridding kernels alone: time, launches and the
  // algorithmic HBM bytes they move (coordinates + visibilities read,
  // subgrids written / read)
  double grid_kernel_ms = 0, grid_launches = 0, grid_bytes = 0;
  double degrid_kernel_ms = 0, degrid_launches = 0, degrid_bytes = 0;
  // visibility passes (a visibility counts once per sub-slab it touches)
  double grid_passes = 0, degrid_passes = 0;
};

// HIP events recorded on the engine stream; read back only after the
// stream has been synchronised, so timing adds no synchronisation.
struct PendingTime {
  double* dst;
  hipEvent_t a, b;
};

// Visibility set prepared on the device: plan, sort and sorted coordinates.
// Invert and predict over the same uvw/freq reuse it (a major cycle).
struct VisSet {
  Plan plan;
  DevPlan dp;
  size_t n_rows = 0, n_chan = 0, n_vis = 0, n_valid = 0;
  bool skip_zero = false;  // zero visibilities were left out (invert only)
  // Prepare read these visibilities and left the max |V| of every row in
  // the arena's "grid.rowmax": the Invert that follows in the same library
  // call skips its row-maxima pass (cleared by that Invert)
  const float2* rowmax_of = nullptr;
  // device arrays (owned by the arena under names prefixed with `tag`)
  std::string tag;
  double* uvw = nullptr;      // [n_rows][3]
  double* scale = nullptr;    // [n_chan] freq / c
  uint32_t* perm = nullptr;   // [n_valid] sorted -> flat index | kPermFlip
  uint32_t* rows = nullptr;   // [n_rows] caller row of renumbered row r
  // [n_valid] renumbered row of each sorted visibility (sets gridded with
  // the fused gather only: the per-item bounds read the row maxima by it)
  uint32_t* vrow = nullptr;
  float2* phasor = nullptr;   // [n_valid] exp(-2 pi i (u l_s + v m_s + w zc))
  uint32_t* off = nullptr;    // [nkeys + 1] bucket offsets
  float4* coord = nullptr;    // [n_valid] fu, fv, fw, packed cell/plane/flip
  WorkItem* items = nullptr;  // work items of all gridding sub-slabs
  uint32_t* item_pos = nullptr;  // [nsub * ntiles + 1] scan of item counts
  int nsub = 0;
  std::vector<double> sub_vis;     // visibility passes of each sub-slab
  std::vector<uint32_t> sub_items;  // work items of each sub-slab
  size_t n_items = 0;
  size_t max_items_per_sub = 0;  // largest work-item count of one sub-slab
  float* corr = nullptr;      // [nx][ny] 1 / (Phi_u Phi_v Phi_w) with sign
  std::string corr_key;       // parameters the correction was built for
};

class Engine {
 public:
  Engine(const ImageGeometry& g, const PlanOptions& opt);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  const ImageGeometry& Geometry() const { return geom_; }
  PlanOptions& Options() { return opt_; }
  hipStream_t Stream() const { return stream_; }

  // Build (or rebuild) the visibility set `vs` from device uvw / host freq.
  // With `vis` non-null, zero visibilities are excluded from the sort.
  void Prepare(VisSet& vs, const std::string& tag, size_t n_rows, size_t n_chan,
               const double* uvw_dev, const double* freq_host,
               const float2* vis_dev);

  // dirty_acc[x][y] += corrected image of the visibilities (device, fp64).
  void Invert(VisSet& vs, const float2* vis_dev, double* img_dev);
  // vis_out[row][chan] = prediction from model img_dev ([x][y], fp64).
  void Predict(VisSet& vs, const double* img_dev, float2* vis_out_dev);

  // Autotuner: times prepare + invert + predict of every candidate shape
  // (support W and oversampling meeting epsilon, tile edge) on a strided
  // row sample of at most max_sample_vis visibilities and extrapolates to
  // the full set (the per-visibility stages scale with the count, the plane
  // FFTs do not).  Leaves the fastest candidate in Options() and returns
  // every trial, fastest first.
  struct TuneTrial {
    int W = 0;
    double sigma = 0;
    int T = 0, Sg = 0, P = 0, nu = 0;
    double est_ms = 0, vis_ms = 0, fixed_ms = 0;
    std::string error;
  };
  std::vector<TuneTrial> Autotune(size_t n_rows, size_t n_chan, const double* uvw_dev,
                                  const double* freq_host, size_t max_sample_vis);

  // Fourier-interpolation resampling of a float image [hi][wi] (device) to
  // [ho][wo] (device), on the engine stream (resample.hip).
  void Resample(const float* in, int wi, int hi, float* out, int wo, int ho);
  // In-place unnormalised FFT of nlines contiguous lines of length n (any
  // n <= 16384 that is 2^a 3^b 5^c, any other n <= 8192 by Bluestein).
  void LineFFTAny(float2* data, size_t nlines, int n, bool inverse);

  DeviceArena& Arena() { return arena_; }
  StageTimes& Times() { return times_; }
  // Accumulates the recorded event pairs into Times(); the caller must have
  // synchronised the stream (every public entry point does).
  void CollectTimes();
  bool timing = false;

 private:
  friend class FacetBatch;
  const float2* Twiddles(int n);
  // Line transforms of a plane slab: descriptors of the LDS FFT lines and,
  // for sides longer than the LDS FFT (ru, rv > 1 subsequences), the
  // temporary of the two-pass transform.
  struct LineFFTs {
    FFTDesc fu{}, fv{};
    int ru = 1, rv = 1;
    float2* tmp = nullptr;
  };
  LineFFTs PlaneFFTs(const DevPlan& dp, int nplanes);
  void BuildCorrection(VisSet& vs);
  void BuildWorkItems(VisSet& vs);

  ImageGeometry geom_;
  PlanOptions opt_;
  hipStream_t stream_ = nullptr;
  DeviceArena arena_;
  std::map<int, float2*> twiddles_;
  std::set<int> bluestein_ready_;  // line lengths whose chirp tables are built
  StageTimes times_;

 public:
  std::vector<PendingTime> pending_;
};

// Facets gridded and degridded together (facets.hip): one sort, one set of
// work items and one persistent launch per sub-slab for all facets; the
// plane FFTs, w screens and corrections per facet.  Facet f's rows are
// [row_off[f], row_off[f] + n_rows[f]) of the shared uvw / visibility
// buffers (rows may be shared between facets).  Predict writes the facets'
// visibilities one after another ("virtual" layout: facet f's block starts
// at the sum of the earlier facets' n_rows * n_chan).
// Launch-configuration queries, cached per kernel so that launches issued
// during a HIP graph capture make no attribute or occupancy calls once the
// same configuration has run eagerly: opt a kernel into `bytes` of dynamic
// LDS (above the 48 KiB default), workgroups of `block` threads resident per
// CU with `lds` bytes each, and the device's CU count.
void AllowDynamicLds(const void* kernel, size_t bytes);
int ResidentBlocksPerCU(const void* kernel, int block, size_t lds);
int DeviceCUs();

// Records an event pair around a scope on the engine stream when timing is
// on; Engine::CollectTimes() adds the elapsed time to *dst.
struct EventTimer {
  Engine* e;
  double* dst;
  hipEvent_t a = nullptr;
  EventTimer(Engine& eng, double* out) : e(&eng), dst(out) {
    if (!e->timing) return;
    (void)hipEventCreate(&a);
    (void)hipEventRecord(a, e->Stream());
  }
  ~EventTimer() {
    if (!e->timing || !a) return;
    hipEvent_t b = nullptr;
    (void)hipEventCreate(&b);
    (void)hipEventRecord(b, e->Stream());
    e->pending_.push_back(PendingTime{dst, a, b});
  }
};

// Host-only facet planning: per-facet plans sharing one kernel shape, tile
// and sub-slab depth (those of the largest facet), the atlas tile origin of
// every facet (shelf packing) and the atlas plan (tile grid, P0, P, keys,
// FFT slab).
struct FacetLayout {
  std::vector<Plan> plans;
  std::vector<int> tu0, tv0;
  Plan atlas;
};
FacetLayout PlanFacets(const std::vector<ImageGeometry>& geoms,
                       const std::vector<VisExtent>& exts, const std::vector<long long>& nvis,
                       const PlanOptions& opt);

class FacetBatch {
 public:
  FacetBatch(const std::vector<ImageGeometry>& facets, const PlanOptions& opt);
  ~FacetBatch();
  FacetBatch(const FacetBatch&) = delete;
  FacetBatch& operator=(const FacetBatch&) = delete;
  int NumFacets() const { return int(geoms_.size()); }
  const ImageGeometry& Geometry(int f) const { return geoms_[f]; }
  Engine& Eng() { return eng_; }
  void Prepare(size_t n_chan, const double* freq_host, const double* uvw_dev,
               const size_t* row_off, const size_t* n_rows);
  bool Prepared() const { return prepared_; }
  // img_dev[f] += corrected dirty image of facet f ([x][y], fp64)
  void Invert(const float2* vis_dev, double* const* img_dev);
  void Predict(const double* const* img_dev, float2* vis_out_dev);
  size_t NumVirtual() const { return atlas_.n_vis; }
  std::string Describe() const;

 private:
  float2* Region(float2* box, int f) const;

  Engine eng_;
  std::vector<ImageGeometry> geoms_;
  std::vector<VisSet> fvs_;  // per facet: plan, device plan, correction
  VisSet atlas_;             // the shared sorted set and work items
  std::vector<int> tu0_, tv0_;
  std::vector<size_t> row_off_, n_rows_;
  std::vector<unsigned long long> vbase_;
  size_t n_chan_ = 0;
  double* scale_ = nullptr;
  uint32_t* perm_phys_ = nullptr;
  uint32_t* rphys_ = nullptr;  // physical row of each renumbered virtual row
  size_t n_vrows_ = 0;
  bool prepared_ = false;
  // Consecutive facets' line FFTs alternate between the engine stream and
  // this one, each with its own row and column buffers, so one facet's
  // column pass overlaps the next facet's transforms (fork / join events
  // around the loop).
  hipStream_t aux_ = nullptr;
  hipEvent_t fork_ = nullptr, join_ = nullptr;
};

// ---- FFT / plane kernels (fft.hip) ----
FFTDesc MakeFFTDesc(int n, const float2* tw);
// Two-level twiddle table of the LDS FFT (host values, FP64-rounded to FP32).
std::vector<float2> FFTTwiddleTable(int n);

void LaunchRowFFTInvert(const DevPlan& dp, const FFTDesc& fu, const float2* box,
                        float2* rbuf, int nplanes, hipStream_t s);
void LaunchColFFTInvert(const DevPlan& dp, const FFTDesc& fv, const float2* cbuf,
                        float* dirty, int plane_lo, int nplanes, hipStream_t s);
void LaunchColFFTPredict(const DevPlan& dp, const FFTDesc& fv, const double* img,
                         const float* corr, float2* cbuf, int plane_lo,
                         int nplanes, hipStream_t s);
void LaunchRowFFTPredict(const DevPlan& dp, const FFTDesc& fu, const float2* rbuf,
                         float2* box, int nplanes, hipStream_t s);
// Lines longer than the LDS FFT (kMaxLdsFFT) as N = R M, M <= kMaxLdsFFT:
// LongFFTRadix(N) is R (1 for N <= kMaxLdsFFT, 0 when no R in 2..5
// splits N).  The long launchers take fm = MakeFFTDesc(M) and a temporary
// of R M complex values per line ([line][r][M]); results as their one-pass
// counterparts.
int LongFFTRadix(int n);
void LaunchRowFFTInvertLong(const DevPlan& dp, const FFTDesc& fm, int N, int R,
                            const float2* box, float2* tmp, float2* rbuf, int nplanes,
                            hipStream_t s);
void LaunchColFFTInvertLong(const DevPlan& dp, const FFTDesc& fm, int N, int R,
                            const float2* cbuf, float2* tmp, float* dirty, int plane_lo,
                            int nplanes, hipStream_t s);
void LaunchColFFTPredictLong(const DevPlan& dp, const FFTDesc& fm, int N, int R,
                             const double* img, const float* corr, float2* tmp, float2* cbuf,
                             int plane_lo, int nplanes, hipStream_t s);
void LaunchRowFFTPredictLong(const DevPlan& dp, const FFTDesc& fm, int N, int R,
                             const float2* rbuf, float2* tmp, float2* box, int nplanes,
                             hipStream_t s);
void LaunchLineFFTLong(float2* data, int nlines, int N, int R, const FFTDesc& fm, float2* tmp,
                       bool inverse, hipStream_t s);
void LaunchTranspose(const float2* in, float2* out, int rows, int cols,
                     int nplanes, hipStream_t s);
void LaunchLineFFT(float2* data, int nlines, const FFTDesc& d, bool inverse,
                   hipStream_t s);

// ---- gridding kernels (grid.hip) ----
// Grids one sub-slab: every work item writes its subgrid of the planes
// [sub_lo, sub_lo + sub_n) into its scratch slot (GridScratchBytes); then
// LaunchGatherBox writes every box cell of those planes.
// `queue` is one device word, the work-item counter of the persistent
// matrix-core kernel (cleared by the launch).  `bounds` (matrix-core
// gridders, W <= 8) holds per work item the (max, sum) of a bound on the
// components of its visibilities (LaunchItemBounds), from which each item
// takes its operand and fixed-point scales.
// With vis_raw (matrix-core gridder only), the gather is fused: the kernel
// reads vis_raw[gather_perm[i]] (flip bit in the word) times phasor[i]
// instead of vis_sorted[i].
void LaunchGrid(const DevPlan& dp, const VisSet& vs, const float2* vis_sorted,
                float2* scratch, uint32_t* queue, const float2* bounds, int sub, int sub_lo,
                int sub_n, hipStream_t s, const float2* vis_raw = nullptr,
                const uint32_t* gather_perm = nullptr);
void LaunchGatherBox(const DevPlan& dp, const VisSet& vs, float2* box,
                     float2* scratch, int sub, int sub_lo, int sub_n,
                     int slab_lo, hipStream_t s);
// Split-precision (f16 pairs) matrix-core gridding for W <= 8, same
// arguments and results as LaunchGrid's matrix-core path (grid_split.hip);
// LaunchGrid calls it when dp.split is set.
void LaunchGridSplit(const DevPlan& dp, const VisSet& vs, const float2* vis_sorted,
                     float2* scratch, uint32_t* queue, const float2* bounds, int sub,
                     int sub_lo, int sub_n, hipStream_t s, const float2* vis_raw,
                     const uint32_t* gather_perm);
size_t SplitGridLdsBytes(const DevPlan& dp, int sub_n);
// Split-precision matrix-core degridding for W <= 8 (grid_split.hip);
// LaunchDegrid calls it when dp.split is set.
void LaunchDegridSplit(const DevPlan& dp, const VisSet& vs, const float2* box, float2* acc,
                       uint32_t* queue, int sub, int sub_lo, int sub_n, int slab_lo,
                       hipStream_t s, const uint32_t* scatter_perm);
size_t SplitDegridLdsBytes(const DevPlan& dp, int sub_n);
size_t GridScratchBytes(const DevPlan& dp, int Sg, size_t max_items_per_sub);
// `queue`: one device word, the work-item counter of the persistent
// matrix-core degridder (cleared by the launch).
// With scatter_perm (matrix-core degridder only; every visibility in one
// sub-slab), the scatter is fused: acc is the caller's output, written at
// scatter_perm[i] with the conjugate phasor and the flip applied.
void LaunchDegrid(const DevPlan& dp, const VisSet& vs, const float2* box,
                  float2* acc, uint32_t* queue, int sub, int sub_lo, int sub_n,
                  int slab_lo, hipStream_t s, const uint32_t* scatter_perm = nullptr);
size_t GridLdsBytes(const DevPlan& dp, int sub_n);

// ---- per-visibility kernels (vis.hip) ----
// One facet of a facet batch on the device: its plan, the offset of its
// rows in the shared uvw / visibility buffers, and the atlas tile where its
// tile (0, 0) sits.
struct FacetDev {
  DevPlan dp;
  unsigned long long row_off;
  int tu0, tv0;
};
// keys[v] (and vals[v] = v when vals is not null) for every virtual index v
// of the batch (facet f owns [vbase[f], vbase[f + 1])).
void LaunchFacetKeys(const FacetDev* fac, const unsigned long long* vbase, int F,
                     const double* uvw, const double* scale, size_t n_chan, size_t n_virtual,
                     int atlas_ntu, int atlas_P0, int key_shift, int key_T, uint32_t* keys,
                     uint32_t* vals, hipStream_t s);
// Keys in renumbered virtual row order (element r * n_chan + ch = channel
// ch of virtual row rperm[r]; facet f owns virtual rows [vrbase[f],
// vrbase[f + 1])); uvw4 = the renumbered rows padded to 32 bytes.
void LaunchFacetKeysRenumbered(const FacetDev* fac, const unsigned long long* vrbase, int F,
                               const double* uvw4, const double* scale,
                               size_t n_vrows, size_t n_chan, int atlas_ntu, int atlas_P0,
                               int key_shift, int key_T, uint32_t* keys, hipStream_t s);
// rphys[r] = physical row and rfac[r] = facet of renumbered virtual row r.
void LaunchFacetRowMap(const FacetDev* fac, const unsigned long long* vrbase, int F,
                       const uint32_t* rperm, size_t n_vrows, uint32_t* rphys, uint32_t* rfac,
                       hipStream_t s);
// Sorted coordinates / phasors; perm enters with sorted renumbered elements
// and leaves with the virtual index, perm_phys gets the physical index
// (both with the flip bit); vrow (if not null) the renumbered virtual row.
void LaunchFacetCoords(const FacetDev* fac, const unsigned long long* vrbase, int F,
                       const double* uvw4, const double* scale,
                       size_t n_chan, uint32_t* perm, uint32_t* perm_phys, size_t n_valid,
                       float4* coord, float2* phasor, hipStream_t s, uint32_t* vrow = nullptr);
void LaunchExtents(const double* uvw, size_t n_rows, double smin, double smax,
                   double* partial, double* out, hipStream_t s);
// Unsorted sort keys of every visibility (nkeys for zero ones); vals[i] = i
// when vals is not null.
void LaunchKeys(const DevPlan& dp, const double* uvw, const double* scale,
                size_t n_rows, size_t n_chan, const float2* vis, uint32_t* keys,
                uint32_t* vals, hipStream_t s);
// uvw4[r][4] = (u, v, w, tag) of row rfetch[r] (renumbered rows), the
// fourth word holding the bits of rtag[r] (the row number the coordinate and
// key kernels report for renumbered row r).
void LaunchPadUvw(const double* uvw, const uint32_t* rfetch, const uint32_t* rtag, size_t n_rows,
                  double* uvw4, hipStream_t s, const uint32_t* rhi = nullptr);
// Keys of the visibilities in renumbered row order (element r * n_chan + ch
// = channel ch of caller row tag(r)); zero visibilities get nkeys.  With
// vis and rowmax_bits ([n_rows] words), also the bits of max |V| of every
// renumbered row (the values LaunchRowMax gives with the rows' caller
// map), from the same reads.
void LaunchKeysRenumbered(const DevPlan& dp, const double* uvw4, const double* scale,
                          size_t n_rows, size_t n_chan, const float2* vis,
                          uint32_t* keys, hipStream_t s, uint32_t* rowmax_bits = nullptr);
// out[i] = in[i] & mask
void LaunchMaskBits(const uint32_t* in, size_t n, uint32_t mask, uint32_t* out,
                    hipStream_t s);
// vis_sorted[i] = flip(vis[perm[i]]) * phasor[i].
void LaunchGatherVis(const uint32_t* perm, const float2* phasor, size_t n_valid,
                     const float2* vis, float2* vis_sorted, hipStream_t s);
// rowmax_bits[r] = bits of max |V| over the n_chan visibilities of row
// row_map[r] of vis ([rows][n_chan]), for r < n_rows: a non-negative float
// (the launch clears the words first).
void LaunchRowMax(const float2* vis, const uint32_t* row_map, size_t n_rows, size_t n_chan,
                  uint32_t* rowmax_bits, hipStream_t s);
// out[r] = in[rows[r]] for r < n.
void LaunchGatherRows(const float* in, const uint32_t* rows, size_t n, float* out,
                      hipStream_t s);
// bounds[it] = (max, sum) over the visibilities of work item it (of all
// n_items) of a bound on |Re| and |Im| of each rotated visibility: with
// vis_sorted, its own larger component; else max |V| of its row,
// rowmax[vrow[i]].  The matrix-core gridders take each item's scales from
// these.
void LaunchItemBounds(const WorkItem* items, size_t n_items, const float2* vis_sorted,
                      const uint32_t* vrow, const float* rowmax, float2* bounds, hipStream_t s);
// vis_out[perm[i]] = flip(acc[i] * conj(phasor[i])).
void LaunchScatterVis(const uint32_t* perm, const float2* phasor, size_t n_valid,
                      const float2* acc, float2* vis_out, hipStream_t s);
void LaunchCorrection(const DevPlan& dp, const double* phi_u, const double* phi_v,
                      const double* quad_a, const double* quad_b, int nquad,
                      float* corr, hipStream_t s);
void LaunchPhiTables(const DevPlan& dp, const double* quad_a, const double* quad_b, int nquad,
                     double* phi_u, double* phi_v, hipStream_t s);
void LaunchAccumulateImage(const float* dirty, const float* corr, double* img,
                           size_t n, hipStream_t s);
void LaunchFillZero(void* p, size_t bytes, hipStream_t s);
// ---- imaging weights and row preparation (weights.hip) ----
// Half-plane weight grid of ImageWeights: [height / 2][width] FP64.
struct WeightGrid {
  double* grid;
  int width, height;
  double psx, psy;
};
struct PrepArgs {
  size_t n_rows, n_chan;
  int n_pol;       // 1, 2 or 4 (collapsed to one visibility per channel)
  int psf;         // visibilities of a point source instead of vis
  double psf_dl, psf_dm, psf_dn;  // its position (dn = sqrt(1 - dl^2 - dm^2) - 1)
  int weighting;   // 0 normal, 1 squared, 2 unit visibility weights
  int polarization;  // PolCode the 4 linear correlations are converted to
};
// Number of FP64 block partials of the weight statistics / gridded total.
size_t WeightStatPartials();
// grid[cell] += w for every (row, channel, polarization) weight;
// total_partial[block] accumulates the gridded sum (kept across calls)
void LaunchGridWeights(const WeightGrid& g, const double* uvw, const double* lambda,
                       size_t n_rows, size_t n_chan, size_t n_pol, const float* weights,
                       bool as_taper, double* total_partial, hipStream_t s);
// FinishGridding; mode 0 natural, 1 uniform, 2 Briggs.  sums: 2 doubles
// (total, sum of squares), scratch: WeightStatPartials() doubles.
void LaunchFinishWeights(const WeightGrid& g, int mode, double robust, const double* total_partial,
                         double* sums, double* scratch, hipStream_t s);
// kind: 0 min uv (a), 1 max uv (a), 2 Tukey (a transition, b max uv),
// 3 inner Tukey (a transition, b min uv), 4 edge (a), 5 edge Tukey
// (a transition, b edge), 6 Gaussian (a = -2 (half-power uv * 2 sqrt(2 ln 2))^2)
void LaunchWeightTaper(const WeightGrid& g, int kind, double a, double b, hipStream_t s);
void LaunchRankFilter(const WeightGrid& g, const double* src, double rank_limit, int win,
                      hipStream_t s);
void LaunchGetWeightGrid(const WeightGrid& g, double* img, hipStream_t s);
void LaunchSampleWeights(const WeightGrid& g, const double* u, const double* v, size_t n,
                         double* out, hipStream_t s);
// out[row][chan]: weighted, collapsed visibilities; stats (device, 4
// doubles): total weight, visibility weight sum, max gridded weight, gridded
// count.  part: WeightStatPartials() doubles.  g may be null (weights 1).
void LaunchPrepareRows(const PrepArgs& a, const double* uvw, const double* lambda,
                       const float2* vis, const float* weights, const float2* model,
                       const uint8_t* selected, const WeightGrid* g, float2* out, double* part,
                       double* stats, hipStream_t s);

// out row r = in row r * stride, for rows of `row_words` 8-byte words
void LaunchStrideRows(const void* in, size_t stride, size_t n_out, size_t row_words, void* out,
                      hipStream_t s);
// out[row][ch] = collapse of the conjugate-gain (after the forward gain
// when `forward`) corrected vis[row][ch][0..npol): 2 polarizations summed,
// 4 converted to polarization `pol` (PolCode).  gains: cached solutions
// [n_times][n_chan][n_ant][nparms]; mode in GainMode order (xx, yy, trace,
// 2-vis diagonal, full).
void LaunchCorrectCollapse(const float2* vis, const uint32_t* antennas,
                           const uint64_t* time_offsets, const float2* gains, size_t n_rows,
                           size_t n_chan, size_t n_ant, int nparms, int mode, int npol,
                           bool forward, int pol, float2* out, hipStream_t s);
// out[k] = in[k * stride] for k in [0, count)
void LaunchStridedGather(const uint32_t* in, size_t stride, size_t count, uint32_t* out,
                         hipStream_t s);
void LaunchSynth(const double* uvw, size_t n_rows, size_t n_chan, const double* scale,
                 const double* src, int n_src, double noise, uint64_t seed, float2* vis,
                 hipStream_t s);

}  // namespace wg
