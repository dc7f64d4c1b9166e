// Facet batch: many facet images gridded and degridded together.
//
// WSClean's BatchInvert (gridding/msgriddermanager.cpp:227-396) reads a
// measurement-set chunk once and hands it to one gridder per facet
// (wgridder/wgriddingmsgridder.cpp:97-115); each facet has its own image
// size and phase-centre shift.  Here the facets of a batch share one sorted
// visibility set and one set of work items: every facet keeps its own plan
// (grid size, w planes, uv box, shifts), and its uv box is placed as a
// region of one "atlas" tile grid.  The sort key of a visibility is its
// atlas tile, first plane and cell, so a single radix sort orders every
// facet's visibilities, and each gridding / degridding sub-slab is ONE
// persistent launch over the work items of all facets.  The plane FFTs,
// w screens and grid corrections then run per facet on its region of the
// atlas box.
//
// All facets of a batch use one kernel shape (support W, oversampling,
// beta), one tile size and one sub-slab depth, taken from the plan of the
// largest facet; their uv boxes are packed into the atlas on shelves, with
// the W - 1 overhang of every box inside its own region so that no work
// item's subgrid reaches another facet.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <sstream>

#include "device.hpp"
#include "engine.hpp"

namespace wg {

FacetLayout PlanFacets(const std::vector<ImageGeometry>& geoms,
                       const std::vector<VisExtent>& exts, const std::vector<long long>& nvis,
                       const PlanOptions& opt) {
  const int F = int(geoms.size());
  WG_CHECK(F > 0 && int(exts.size()) == F && int(nvis.size()) == F, "facet count mismatch");
  FacetLayout lay;
  // the largest facet picks the kernel shape, the tile and the sub-slab depth
  int big = 0;
  for (int f = 1; f < F; ++f)
    if (double(geoms[f].nx) * geoms[f].ny > double(geoms[big].nx) * geoms[big].ny) big = f;
  PlanOptions o = opt;
  o.nvis_hint = nvis[big];
  const Plan lead = MakePlan(geoms[big], exts[big], o);
  o.force_W = lead.W;
  o.force_sigma = lead.sigma;
  o.T = lead.T;
  o.Sg = lead.Sg;
  lay.plans.resize(F);
  for (int f = 0; f < F; ++f) {
    o.nvis_hint = nvis[f];
    lay.plans[f] = MakePlan(geoms[f], exts[f], o);
    const Plan& p = lay.plans[f];
    WG_CHECK(p.W == lead.W && p.T == lead.T && p.Sg == lead.Sg && p.beta == lead.beta,
             "facet plans disagree on the kernel shape");
    WG_CHECK(p.nu <= kMaxLdsFFT && p.nv <= kMaxLdsFFT,
             "a facet's padded grid is larger than " + std::to_string(kMaxLdsFFT) +
                 ": image it with its own gridder");
  }

  // shelf packing of the facets' tile regions: facet f takes ntu + gap by
  // ntv + gap tiles, the gap holding the W - 1 overhang of its box
  const int T = lead.T, W = lead.W;
  const int gap = (W - 1 + T - 1) / T;
  std::vector<int> tw(F), th(F), order(F);
  double area = 0;
  int max_tw = 0;
  for (int f = 0; f < F; ++f) {
    tw[f] = lay.plans[f].ntu + gap;
    th[f] = lay.plans[f].ntv + gap;
    area += double(tw[f]) * th[f];
    max_tw = std::max(max_tw, tw[f]);
  }
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return th[a] > th[b]; });
  const int shelf_w = std::max(max_tw, int(std::ceil(std::sqrt(area))));
  lay.tu0.assign(F, 0);
  lay.tv0.assign(F, 0);
  int x = 0, y = 0, shelf_h = 0, used_w = 0;
  for (int f : order) {
    if (x > 0 && x + tw[f] > shelf_w) {
      y += shelf_h;
      x = 0;
      shelf_h = 0;
    }
    lay.tu0[f] = x;
    lay.tv0[f] = y;
    x += tw[f];
    used_w = std::max(used_w, x);
    shelf_h = std::max(shelf_h, th[f]);
  }
  Plan pa = lead;
  pa.ntu = used_w;
  pa.ntv = y + shelf_h;
  pa.BU = pa.ntu * T + W - 1;
  pa.BV = pa.ntv * T + W - 1;
  pa.box_u0 = pa.box_v0 = 0;
  pa.wrap_u = pa.wrap_v = false;
  pa.P0 = 1;
  pa.P = 1;
  for (const Plan& p : lay.plans) {
    pa.P0 = std::max(pa.P0, p.P0);
    pa.P = std::max(pa.P, p.P);
  }
  PlanSortKeys(pa);
  {
    // the atlas takes as many planes as its deepest facet; when they fit
    // one sub-slab of the full LDS, one sub-slab lets the slab gather and
    // scatter fuse into the FFT passes (the same rule as a single plan)
    const size_t E = size_t(pa.T + pa.W - 1);
    const int max_sg = int((160 * 1024 - kGridLdsReserve) / (E * E * 8));
    if (opt.Sg <= 0 && pa.P > pa.Sg && pa.P <= max_sg && pa.P <= kMaxSubSlab) pa.Sg = pa.P;
  }
  // FFT slab: whole sub-slabs whose atlas box and largest facet line
  // buffers fit the slab budget
  size_t lines = 0;
  for (const Plan& p : lay.plans) lines = std::max(lines, size_t(p.BV) * p.img.nx);
  const double per_plane = 8.0 * (double(pa.BU) * pa.BV + 2.0 * double(lines));
  int sf = int(std::max(1.0, std::floor(double(opt.slab_bytes) / per_plane)));
  sf = std::max(pa.Sg, (sf / pa.Sg) * pa.Sg);
  pa.Sf = std::min(sf, ((pa.P + pa.Sg - 1) / pa.Sg) * pa.Sg);
  lay.atlas = pa;
  return lay;
}

FacetBatch::FacetBatch(const std::vector<ImageGeometry>& facets, const PlanOptions& opt)
    : eng_(facets.empty() ? ImageGeometry{} : facets[0], opt), geoms_(facets) {
  WG_CHECK(!facets.empty(), "a facet batch needs at least one facet");
  for (const ImageGeometry& g : facets)
    WG_CHECK(g.nx > 0 && g.ny > 0, "empty facet image");
}

FacetBatch::~FacetBatch() {
  if (aux_) {
    (void)hipStreamSynchronize(aux_);
    (void)hipStreamDestroy(aux_);
  }
  if (fork_) (void)hipEventDestroy(fork_);
  if (join_) (void)hipEventDestroy(join_);
}

void FacetBatch::Prepare(size_t n_chan, const double* freq_host, const double* uvw_dev,
                         const size_t* row_off, const size_t* n_rows) {
  const int F = NumFacets();
  hipStream_t s = eng_.Stream();
  DeviceArena& A = eng_.Arena();
  EventTimer tm(eng_, &eng_.Times().plan);
  if (!aux_) {
    WG_HIP(hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking));
    WG_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
    WG_HIP(hipEventCreateWithFlags(&join_, hipEventDisableTiming));
  }
  prepared_ = false;
  WG_CHECK(n_chan > 0, "n_chan must be positive");
  row_off_.assign(row_off, row_off + F);
  n_rows_.assign(n_rows, n_rows + F);
  n_chan_ = n_chan;

  // ---- shared channel scales ----
  std::vector<double> scale(n_chan);
  double smin = 1e300, smax = -1e300;
  for (size_t c = 0; c < n_chan; ++c) {
    WG_CHECK(freq_host[c] > 0, "frequencies must be positive");
    scale[c] = freq_host[c] / kSpeedOfLight;
    smin = std::min(smin, scale[c]);
    smax = std::max(smax, scale[c]);
  }
  scale_ = static_cast<double*>(A.Get("fb.scale", n_chan * sizeof(double)));
  WG_HIP(hipMemcpyAsync(scale_, scale.data(), n_chan * sizeof(double), hipMemcpyHostToDevice, s));

  // ---- per-facet extents (one readback for all facets) ----
  double* part = static_cast<double*>(A.Get("extents", (1024 * 6 + 6) * sizeof(double)));
  double* ext_dev = static_cast<double*>(A.Get("fb.extents", size_t(F) * 6 * sizeof(double)));
  LaunchFillZero(ext_dev, size_t(F) * 6 * sizeof(double), s);
  for (int f = 0; f < F; ++f)
    if (n_rows_[f] > 0)
      LaunchExtents(uvw_dev + 3 * row_off_[f], n_rows_[f], smin, smax, part, ext_dev + 6 * f, s);
  std::vector<double> ext(size_t(F) * 6);
  WG_HIP(hipMemcpyAsync(ext.data(), ext_dev, ext.size() * sizeof(double), hipMemcpyDeviceToHost,
                        s));
  WG_HIP(hipStreamSynchronize(s));
  for (double v : ext) WG_CHECK(std::isfinite(v), "non-finite uvw coordinates");

  std::vector<VisExtent> exts(F);
  std::vector<long long> nvis(F);
  for (int f = 0; f < F; ++f) {
    const double* e = &ext[6 * size_t(f)];
    exts[f] = VisExtent{e[0], e[1], e[2], e[3], e[4], e[5]};
    nvis[f] = (long long)(n_rows_[f] * n_chan);
  }
  FacetLayout lay = PlanFacets(geoms_, exts, nvis, eng_.Options());
  const Plan& pa = lay.atlas;
  const int T = pa.T;
  tu0_ = lay.tu0;
  tv0_ = lay.tv0;
  // keep the corrections of the previous prepare: they are rebuilt only
  // when a facet's geometry, kernel, grid or w axis changed
  std::vector<std::string> corr_keys(F);
  for (int f = 0; f < F && f < int(fvs_.size()); ++f) corr_keys[f] = fvs_[f].corr_key;
  fvs_.assign(F, VisSet{});
  for (int f = 0; f < F; ++f) {
    fvs_[f].plan = lay.plans[f];
    fvs_[f].dp = ToDevPlan(lay.plans[f]);
    fvs_[f].tag = "facet" + std::to_string(f);
    fvs_[f].corr_key = corr_keys[f];
  }
  atlas_ = VisSet{};
  atlas_.plan = pa;
  atlas_.dp = ToDevPlan(pa);
  atlas_.tag = "atlas";
  atlas_.n_chan = n_chan;
  atlas_.scale = scale_;
  atlas_.uvw = const_cast<double*>(uvw_dev);

  // ---- device facet table ----
  std::vector<FacetDev> fd(F);
  vbase_.assign(F + 1, 0);
  for (int f = 0; f < F; ++f) {
    fd[f].dp = fvs_[f].dp;
    // the facet's FFT stages address its region of the atlas box
    fd[f].dp.box_pitch = size_t(pa.BU);
    fd[f].dp.box_plane = size_t(pa.BU) * size_t(pa.BV);
    fd[f].row_off = row_off_[f];
    fd[f].tu0 = tu0_[f];
    fd[f].tv0 = tv0_[f];
    fvs_[f].dp = fd[f].dp;
    vbase_[f + 1] = vbase_[f] + (unsigned long long)(n_rows_[f] * n_chan);
  }
  const size_t n_virtual = size_t(vbase_[F]);
  WG_CHECK(n_virtual < 0x80000000ull, "at most 2^31 - 1 visibilities per facet batch");
  FacetDev* dfac = static_cast<FacetDev*>(A.Get("fb.facets", size_t(F) * sizeof(FacetDev)));
  WG_HIP(hipMemcpyAsync(dfac, fd.data(), fd.size() * sizeof(FacetDev), hipMemcpyHostToDevice, s));

  // ---- virtual row renumbering ----
  // Virtual rows are renumbered in the order of their atlas key at the
  // middle channel (stable), as the single gridder renumbers its rows: the
  // atlas keys of different facets are disjoint and order each facet's rows
  // exactly as that facet's own keys do, so the sorted order of every facet
  // is the one its single gridder produces.
  const size_t n_vrows = n_virtual / n_chan;
  std::vector<unsigned long long> vrbase(F + 1);
  for (int f = 0; f <= F; ++f) vrbase[f] = vbase_[f] / n_chan;
  unsigned long long* dvrb =
      static_cast<unsigned long long*>(A.Get("fb.vrbase", size_t(F + 1) * sizeof(uint64_t)));
  WG_HIP(hipMemcpyAsync(dvrb, vrbase.data(), vrbase.size() * sizeof(uint64_t),
                        hipMemcpyHostToDevice, s));
  uint32_t* rperm = static_cast<uint32_t*>(A.Get("rows.perm", n_vrows * 4));
  double* uvw4 = static_cast<double*>(A.Get("uvw4", n_vrows * 4 * sizeof(double)));
  uint32_t* scratch = static_cast<uint32_t*>(
      A.Get("sort.scratch", RadixSortScratchWords(std::max(n_virtual, n_vrows)) * 4));
  if (n_vrows) {
    uint32_t* rk = static_cast<uint32_t*>(A.Get("rows.keys", n_vrows * 4));
    uint32_t* rk2 = static_cast<uint32_t*>(A.Get("rows.keys2", n_vrows * 4));
    uint32_t* rp2 = static_cast<uint32_t*>(A.Get("rows.perm2", n_vrows * 4));
    uint32_t* rphys = static_cast<uint32_t*>(A.Get("fb.rows.phys", n_vrows * 4));
    uint32_t* rfac = static_cast<uint32_t*>(A.Get("rows.facet", n_vrows * 4));
    LaunchFacetKeys(dfac, dvrb, F, uvw_dev, scale_ + n_chan / 2, 1, n_vrows, pa.ntu, pa.P0,
                    pa.key_shift, pa.key_T, rk, nullptr, s);
    if (DeviceRadixSort(rk, rperm, rk2, rp2, n_vrows, pa.key_bits, scratch, s, true))
      rperm = rp2;
    LaunchFacetRowMap(dfac, dvrb, F, rperm, n_vrows, rphys, rfac, s);
    // the padded rows carry their virtual row and their facet
    LaunchPadUvw(uvw_dev, rphys, rperm, n_vrows, uvw4, s, rfac);
    rphys_ = rphys;
  }
  n_vrows_ = n_vrows;

  // ---- keys, one stable sort, bucket offsets ----
  atlas_.n_vis = n_virtual;
  uint32_t* keys = static_cast<uint32_t*>(A.Get("atlas.keys", n_virtual * 4));
  uint32_t* vals = static_cast<uint32_t*>(A.Get("atlas.perm", n_virtual * 4));
  if (n_virtual) {
    // the sort's first pass generates the index values itself
    LaunchFacetKeysRenumbered(dfac, dvrb, F, uvw4, scale_, n_vrows, n_chan, pa.ntu, pa.P0,
                              pa.key_shift, pa.key_T, keys, s);
    uint32_t* keys2 = static_cast<uint32_t*>(A.Get("atlas.keys2", n_virtual * 4));
    uint32_t* vals2 = static_cast<uint32_t*>(A.Get("atlas.perm2", n_virtual * 4));
    if (DeviceRadixSort(keys, vals, keys2, vals2, n_virtual, pa.key_bits, scratch, s, true)) {
      keys = keys2;
      vals = vals2;
    }
  }
  atlas_.perm = vals;
  atlas_.off = static_cast<uint32_t*>(A.Get("atlas.off", (size_t(pa.nbuckets) + 1) * 4));
  DeviceBucketOffsets(keys, n_virtual, pa.nbuckets, uint32_t(pa.key_T) * uint32_t(pa.key_T),
                      atlas_.off, s);
  uint32_t nvalid = 0;
  WG_HIP(hipMemcpyAsync(&nvalid, atlas_.off + pa.nbuckets, 4, hipMemcpyDeviceToHost, s));
  WG_HIP(hipStreamSynchronize(s));
  WG_CHECK(nvalid == n_virtual, "facet batch: visibilities outside the atlas key space");
  atlas_.n_valid = nvalid;
  atlas_.coord = static_cast<float4*>(A.Get("atlas.coord", nvalid * sizeof(float4)));
  atlas_.phasor = static_cast<float2*>(A.Get("atlas.phasor", nvalid * sizeof(float2)));
  perm_phys_ = static_cast<uint32_t*>(A.Get("atlas.perm_phys", nvalid * 4));
  // a batch that the matrix-core kernel grids in one sub-slab keeps each
  // sorted visibility's renumbered virtual row for the per-item bounds
  const bool fused = pa.W <= 8 && pa.P <= pa.Sg;
  atlas_.vrow = fused ? static_cast<uint32_t*>(A.Get("atlas.vrow", nvalid * 4)) : nullptr;
  LaunchFacetCoords(dfac, dvrb, F, uvw4, scale_, n_chan, atlas_.perm, perm_phys_, nvalid,
                    atlas_.coord, atlas_.phasor, s, atlas_.vrow);
  eng_.BuildWorkItems(atlas_);
  for (int f = 0; f < F; ++f) eng_.BuildCorrection(fvs_[f]);
  prepared_ = true;
}

float2* FacetBatch::Region(float2* box, int f) const {
  const DevPlan& a = atlas_.dp;
  return box + size_t(tv0_[f]) * a.T * size_t(a.BU) + size_t(tu0_[f]) * a.T;
}

void FacetBatch::Invert(const float2* vis_dev, double* const* img_dev) {
  WG_CHECK(prepared_, "facet batch not prepared");
  const int F = NumFacets();
  const DevPlan& da = atlas_.dp;
  const Plan& pa = atlas_.plan;
  hipStream_t s = eng_.Stream();
  DeviceArena& A = eng_.Arena();
  if (atlas_.n_valid == 0) return;
  uint32_t* words = static_cast<uint32_t*>(A.Get("grid.words", 4 * sizeof(uint32_t)));
  // a single sub-slab: the matrix-core gridder gathers through perm_phys
  const bool fused = da.W <= 8 && atlas_.nsub == 1;
  float2* vsorted = nullptr;
  // the matrix-core gridders' per-item scales
  float2* bounds = nullptr;
  {
    EventTimer tm(eng_, &eng_.Times().sort);
    if (da.W <= 8)
      bounds = static_cast<float2*>(A.Get("grid.bounds", atlas_.n_items * sizeof(float2)));
    if (fused) {
      // row maxima of the renumbered virtual rows, read through their
      // physical rows
      float* rowmax = static_cast<float*>(A.Get("grid.rowmax", n_vrows_ * sizeof(float)));
      LaunchRowMax(vis_dev, rphys_, n_vrows_, n_chan_, reinterpret_cast<uint32_t*>(rowmax), s);
      LaunchItemBounds(atlas_.items, atlas_.n_items, nullptr, atlas_.vrow, rowmax, bounds, s);
    } else {
      vsorted = static_cast<float2*>(A.Get("vis_sorted", atlas_.n_valid * sizeof(float2)));
      LaunchGatherVis(perm_phys_, atlas_.phasor, atlas_.n_valid, vis_dev, vsorted, s);
      if (bounds)
        LaunchItemBounds(atlas_.items, atlas_.n_items, vsorted, nullptr, nullptr, bounds, s);
    }
  }
  StageTimes& st = eng_.Times();
  std::vector<float*> dirty(F);
  size_t lines = 0;
  for (int f = 0; f < F; ++f) {
    const size_t npix = size_t(geoms_[f].nx) * geoms_[f].ny;
    dirty[f] = static_cast<float*>(A.Get(fvs_[f].tag + ".dirty", npix * 4));
    LaunchFillZero(dirty[f], npix * 4, s);
    lines = std::max(lines, size_t(fvs_[f].plan.BV) * fvs_[f].plan.img.nx);
  }
  float2* box =
      static_cast<float2*>(A.Get("box", size_t(pa.Sf) * da.box_plane * sizeof(float2)));
  float2* rbuf[2], *cbuf[2];
  rbuf[0] = static_cast<float2*>(A.Get("rbuf", size_t(pa.Sf) * lines * sizeof(float2)));
  cbuf[0] = static_cast<float2*>(A.Get("cbuf", size_t(pa.Sf) * lines * sizeof(float2)));
  rbuf[1] = F > 1 ? static_cast<float2*>(A.Get("rbuf2", size_t(pa.Sf) * lines * sizeof(float2)))
                  : rbuf[0];
  cbuf[1] = F > 1 ? static_cast<float2*>(A.Get("cbuf2", size_t(pa.Sf) * lines * sizeof(float2)))
                  : cbuf[0];
  float2* scratch = static_cast<float2*>(
      A.Get("grid.scratch", GridScratchBytes(da, pa.Sg, atlas_.max_items_per_sub)));
  for (int f0 = 0; f0 < pa.P; f0 += pa.Sf) {
    const int np = std::min(pa.Sf, pa.P - f0);
    {
      EventTimer tm(eng_, &st.grid);
      const size_t E = size_t(da.T + da.W - 1);
      for (int g = f0 / pa.Sg; g * pa.Sg < f0 + np; ++g) {
        const int sub_lo = g * pa.Sg;
        const int sub_n = std::min(pa.Sg, pa.P - sub_lo);
        {
          // one persistent launch over the work items of every facet
          EventTimer tk(eng_, &st.grid_kernel_ms);
          LaunchGrid(da, atlas_, vsorted, scratch, words, bounds, g, sub_lo, sub_n, s,
                     fused ? vis_dev : nullptr, fused ? perm_phys_ : nullptr);
        }
        if (eng_.timing) {
          st.grid_launches += 1;
          st.grid_passes += atlas_.sub_vis[g];
          st.grid_bytes += (fused ? 36.0 : 24.0) * atlas_.sub_vis[g] +
                           8.0 * double(sub_n) * E * E * double(atlas_.sub_items[g]);
        }
        LaunchGatherBox(da, atlas_, box, scratch, g, sub_lo, sub_n, f0, s);
      }
    }
    EventTimer tf(eng_, &st.fft);
    WG_HIP(hipEventRecord(fork_, s));
    WG_HIP(hipStreamWaitEvent(aux_, fork_, 0));
    int k = 0;
    for (int f = 0; f < F; ++f) {
      const VisSet& v = fvs_[f];
      const int npf = std::min(np, v.plan.P - f0);
      if (npf <= 0) continue;
      const FFTDesc fu = MakeFFTDesc(v.dp.nu, eng_.Twiddles(v.dp.nu));
      const FFTDesc fv = MakeFFTDesc(v.dp.nv, eng_.Twiddles(v.dp.nv));
      const int b = k++ & 1;
      hipStream_t sf = b ? aux_ : s;
      LaunchRowFFTInvert(v.dp, fu, Region(box, f), rbuf[b], npf, sf);
      LaunchTranspose(rbuf[b], cbuf[b], v.dp.BV, v.dp.nx, npf, sf);
      LaunchColFFTInvert(v.dp, fv, cbuf[b], dirty[f], f0, npf, sf);
    }
    WG_HIP(hipEventRecord(join_, aux_));
    WG_HIP(hipStreamWaitEvent(s, join_, 0));
  }
  for (int f = 0; f < F; ++f)
    LaunchAccumulateImage(dirty[f], fvs_[f].corr, img_dev[f],
                          size_t(geoms_[f].nx) * geoms_[f].ny, s);
}

void FacetBatch::Predict(const double* const* img_dev, float2* vis_out_dev) {
  WG_CHECK(prepared_, "facet batch not prepared");
  const int F = NumFacets();
  const DevPlan& da = atlas_.dp;
  const Plan& pa = atlas_.plan;
  hipStream_t s = eng_.Stream();
  DeviceArena& A = eng_.Arena();
  if (atlas_.n_valid == 0) return;
  // a single sub-slab: the degridder writes the output through perm
  const bool fused = da.W <= 8 && atlas_.nsub == 1;
  float2* acc = vis_out_dev;
  if (!fused) {
    acc = static_cast<float2*>(A.Get("vis_acc", atlas_.n_valid * sizeof(float2)));
    LaunchFillZero(acc, atlas_.n_valid * sizeof(float2), s);
  }
  StageTimes& st = eng_.Times();
  size_t lines = 0;
  for (int f = 0; f < F; ++f)
    lines = std::max(lines, size_t(fvs_[f].plan.BV) * fvs_[f].plan.img.nx);
  float2* box =
      static_cast<float2*>(A.Get("box", size_t(pa.Sf) * da.box_plane * sizeof(float2)));
  float2* rbuf[2], *cbuf[2];
  rbuf[0] = static_cast<float2*>(A.Get("rbuf", size_t(pa.Sf) * lines * sizeof(float2)));
  cbuf[0] = static_cast<float2*>(A.Get("cbuf", size_t(pa.Sf) * lines * sizeof(float2)));
  rbuf[1] = F > 1 ? static_cast<float2*>(A.Get("rbuf2", size_t(pa.Sf) * lines * sizeof(float2)))
                  : rbuf[0];
  cbuf[1] = F > 1 ? static_cast<float2*>(A.Get("cbuf2", size_t(pa.Sf) * lines * sizeof(float2)))
                  : cbuf[0];
  uint32_t* queue = static_cast<uint32_t*>(A.Get("grid.words", 4 * sizeof(uint32_t)));
  for (int f0 = 0; f0 < pa.P; f0 += pa.Sf) {
    const int np = std::min(pa.Sf, pa.P - f0);
    // planes of a facet beyond its own P are never read: its work items
    // only reach planes below P
    {
      EventTimer tf(eng_, &st.fft);
      WG_HIP(hipEventRecord(fork_, s));
      WG_HIP(hipStreamWaitEvent(aux_, fork_, 0));
      int k = 0;
      for (int f = 0; f < F; ++f) {
        const VisSet& v = fvs_[f];
        const int npf = std::min(np, v.plan.P - f0);
        if (npf <= 0) continue;
        const FFTDesc fu = MakeFFTDesc(v.dp.nu, eng_.Twiddles(v.dp.nu));
        const FFTDesc fv = MakeFFTDesc(v.dp.nv, eng_.Twiddles(v.dp.nv));
        const int b = k++ & 1;
        hipStream_t sf = b ? aux_ : s;
        LaunchColFFTPredict(v.dp, fv, img_dev[f], v.corr, cbuf[b], f0, npf, sf);
        LaunchTranspose(cbuf[b], rbuf[b], v.dp.nx, v.dp.BV, npf, sf);
        LaunchRowFFTPredict(v.dp, fu, rbuf[b], Region(box, f), npf, sf);
      }
      WG_HIP(hipEventRecord(join_, aux_));
      WG_HIP(hipStreamWaitEvent(s, join_, 0));
    }
    EventTimer td(eng_, &st.degrid);
    const size_t E = size_t(da.T + da.W - 1);
    for (int g = f0 / pa.Sg; g * pa.Sg < f0 + np; ++g) {
      const int sub_lo = g * pa.Sg;
      const int sub_n = std::min(pa.Sg, pa.P - sub_lo);
      {
        EventTimer tk(eng_, &st.degrid_kernel_ms);
        LaunchDegrid(da, atlas_, box, acc, queue, g, sub_lo, sub_n, f0, s,
                     fused ? atlas_.perm : nullptr);
      }
      if (eng_.timing) {
        st.degrid_launches += 1;
        st.degrid_passes += atlas_.sub_vis[g];
        st.degrid_bytes += (fused ? 36.0 : 32.0) * atlas_.sub_vis[g] +
                           8.0 * double(sub_n) * E * E * double(atlas_.sub_items[g]);
      }
    }
  }
  if (!fused) {
    EventTimer tsc(eng_, &st.sort);
    LaunchScatterVis(atlas_.perm, atlas_.phasor, atlas_.n_valid, acc, vis_out_dev, s);
  }
}

std::string FacetBatch::Describe() const {
  std::ostringstream o;
  o.precision(17);
  o << "{\"atlas\":" << atlas_.plan.describe() << ",\"n_virtual\":" << atlas_.n_vis
    << ",\"n_items\":" << atlas_.n_items << ",\"facets\":[";
  for (int f = 0; f < NumFacets(); ++f) {
    if (f) o << ",";
    o << "{\"tu0\":" << tu0_[f] << ",\"tv0\":" << tv0_[f] << ",\"plan\":" << fvs_[f].plan.describe()
      << "}";
  }
  o << "]}";
  return o.str();
}

}  // namespace wg
