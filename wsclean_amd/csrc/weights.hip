// Imaging weights and the caller-side row preparation on the device.
//
// ImageWeights (structures/imageweights.h:22-132, imageweights.cpp:33-390)
// keeps a half-plane grid of FP64 weight sums, [height / 2][width], indexed
// by the uv cell of every (row, channel, polarization) weight; the grid is
// then turned into natural / uniform / Briggs imaging weights, optionally
// limited and tapered in uv, and sampled per visibility.  The row
// preparation restates MsGridderData::CalculateWeights + ApplyWeights
// (gridding/msgridderdata.h:761-789, :815-886) with RotateVisibilities
// (:929-941) and the polarization collapse (:37-52), and reduces the
// weight sums the caller divides the image by.
//
// The uv -> cell mapping runs in FP64 without contraction, as the
// reference's host code, so every visibility lands in the same cell.  Cell
// sums use FP64 atomics; f32 weights add exactly in FP64 unless a cell sum
// needs more than 53 significant bits, so the grid does not depend on the
// order of the adds.  The weight statistics are reduced in a fixed order
// (block partials, then one block).
#include <hip/hip_runtime.h>

#include <cmath>

#include "device.hpp"
#include "engine.hpp"

namespace wg {

namespace {

constexpr int kBlock = 256;
constexpr int kStatBlocks = 1024;

unsigned Blocks(size_t n) {
  size_t b = (n + kBlock - 1) / kBlock;
  return unsigned(b < 65536 ? (b ? b : 1) : 65536);
}

// imageweights.h:74-95: cell of (u, v) (flipped to v >= 0); -1 outside
__device__ __forceinline__ long long WeightCell(double u, double v, const WeightGrid& g) {
#pragma clang fp contract(off)
  if (v < 0.0) {
    u = -u;
    v = -v;
  }
  const double fx = floor(u * double(g.width) * g.psx + double(g.width / 2));
  const double fy = floor(v * double(g.height) * g.psy);
  // the range test on the doubles equals the reference's on the int casts
  // for every in-range value and avoids casting out-of-range ones
  if (fx >= 0.0 && fx < double(g.width) && fy < double(g.height / 2))
    return (long long)fx + (long long)fy * g.width;
  return -1;
}

__global__ void __launch_bounds__(kBlock) GridWeightsKernel(
    WeightGrid g, const double* __restrict__ uvw, const double* __restrict__ lambda,
    size_t n_rows, size_t n_chan, size_t n_pol, const float* __restrict__ weights,
    int as_taper, double* __restrict__ total_partial) {
#pragma clang fp contract(off)
  double local = 0.0;
  const size_t n = n_rows * n_chan;
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock) {
    const size_t row = i / n_chan, ch = i - row * n_chan;
    double um = uvw[3 * row], vm = uvw[3 * row + 1];
    if (vm < 0.0) {
      um = -um;
      vm = -vm;
    }
    const long long cell = WeightCell(um / lambda[ch], vm / lambda[ch], g);
    if (cell < 0) continue;
    for (size_t p = 0; p < n_pol; ++p) {
      float w = weights[i * n_pol + p];
      if (as_taper && w != 0.f) w = 1.f;
      atomicAdd(&g.grid[cell], double(w));
      local += double(w);
    }
  }
  // total of the gridded weights: per-block partials (summed in block order)
  __shared__ double sh[kBlock / 64];
  for (int m = 32; m >= 1; m >>= 1) local += __shfl_xor(local, m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) s += sh[k];
    total_partial[blockIdx.x] += s;
  }
}

// Sum of a partial array in index order (one block; deterministic).
__global__ void SumPartialsKernel(const double* __restrict__ part, int n, double* __restrict__ out,
                                  int accumulate) {
  __shared__ double sh[kBlock];
  double s = 0.0;
  // each thread a contiguous range, then the ranges in order
  const int per = (n + kBlock - 1) / kBlock;
  for (int k = threadIdx.x * per; k < min(n, (threadIdx.x + 1) * per); ++k) s += part[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kBlock; ++k) t += sh[k];
    *out = accumulate ? *out + t : t;
  }
}

// sum of squares of the grid (Briggs' average weight), block partials
__global__ void __launch_bounds__(kBlock) SquareSumKernel(const double* __restrict__ grid,
                                                          size_t n, double* __restrict__ part) {
  double local = 0.0;
  const size_t per = (n + size_t(gridDim.x) * kBlock - 1) / (size_t(gridDim.x) * kBlock);
  const size_t t = size_t(blockIdx.x) * kBlock + threadIdx.x;
  for (size_t i = t * per; i < min(n, (t + 1) * per); ++i) local += grid[i] * grid[i];
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = local;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < kBlock; ++k) s += sh[k];
    part[blockIdx.x] = s;
  }
}

// imageweights.cpp:116-150 (ssq: Briggs' scale, from the device sums)
__global__ void __launch_bounds__(kBlock) FinishKernel(double* __restrict__ grid, size_t n,
                                                       int mode, const double* __restrict__ sums,
                                                       double robust) {
  double ssq = 0.0;
  if (mode == 2) {
    const double avg = sums[1] / sums[0];  // sum of squares / total
    const double num = 5.0 * exp(-M_LN10 * robust);
    ssq = num * num / avg;
  }
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock) {
    const double v = grid[i];
    if (v == 0.0) continue;
    grid[i] = mode == 2 ? 1.0 / (1.0 + v * ssq) : (mode == 1 ? 1.0 / v : 1.0);
  }
}

__device__ __forceinline__ double Tukey(double x, double n) {
  return 0.5 * (1.0 + cos((M_PI / n) * (x - n)));
}

// uv-range limits and tapers over the half grid (imageweights.cpp:152-268,
// :333-356); kind: 0 min uv, 1 max uv, 2 Tukey, 3 inner Tukey, 4 edge,
// 5 edge Tukey, 6 Gaussian
__global__ void __launch_bounds__(kBlock) TaperKernel(WeightGrid g, int kind, double a, double b) {
#pragma clang fp contract(off)
  const size_t n = size_t(g.height / 2) * g.width;
  const double max_u = double(g.width - g.width / 2) / (double(g.width) * g.psx);
  const double max_v = double(g.height / 2) / (double(g.height) * g.psy);
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock) {
    const int y = int(i / g.width), x = int(i - size_t(y) * g.width);
    const double u = double(x - g.width / 2) / (double(g.width) * g.psx);
    const double v = double(y) / (double(g.height) * g.psy);
    const double d2 = u * u + v * v;
    double& w = g.grid[i];
    switch (kind) {
      case 0:
        if (d2 < a * a) w = 0.0;
        break;
      case 1:
        if (d2 > a * a) w = 0.0;
        break;
      case 2: {  // a: transition, b: max uv
        if (d2 > b * b)
          w = 0.0;
        else if (d2 > (b - a) * (b - a))
          w *= Tukey(b - sqrt(d2), a);
      } break;
      case 3: {  // a: transition, b: min uv
        if (d2 < b * b)
          w = 0.0;
        else if (d2 < (b + a) * (b + a))
          w *= Tukey(sqrt(d2) - b, a);
      } break;
      case 4:
        if (max_u - fabs(u) < a || max_v - fabs(v) < a) w = 0.0;
        break;
      case 5: {  // a: transition, b: edge size
        const double ud = max_u - fabs(u), vd = max_v - fabs(v), total = a + b;
        if (ud < b || vd < b) {
          w = 0.0;
        } else if (ud < total || vd < total) {
          double ru = ud - b, rv = vd - b;
          if (ru > a) ru = a;
          if (rv > a) rv = a;
          w *= Tukey(ru, a) * Tukey(rv, a);
        }
      } break;
      case 6:  // a: -2 (half-power uv * sigma-to-HP)^2
        if (w != 0.0) w = w * exp(d2 / a);
        break;
    }
  }
}

// imageweights.cpp:302-317, :358-388: cells above rank_limit x the mean of
// the nonzero cells of the window [x - d, x + d) x [y - d, y + d) are
// clipped to that value (summed in the reference's loop order)
__global__ void __launch_bounds__(kBlock) RankFilterKernel(WeightGrid g, const double* __restrict__ src,
                                                           double rank_limit, int win) {
  const int h2 = g.height / 2;
  const size_t n = size_t(h2) * g.width;
  const int d = win / 2;
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock) {
    const int y = int(i / g.width), x = int(i - size_t(y) * g.width);
    const double w = src[i];
    double out = w;
    if (w != 0.0) {
      const int x1 = x <= d ? 0 : x - d, y1 = y <= d ? 0 : y - d;
      const int x2 = x + d >= g.width ? g.width : x + d;
      const int y2 = y + d >= h2 ? h2 : y + d;
      size_t cnt = 0;
      double s = 0.0;
      for (int yi = y1; yi < y2; ++yi)
        for (int xi = x1; xi < x2; ++xi) {
          const double v = src[size_t(yi) * g.width + xi];
          if (v != 0.0) {
            ++cnt;
            s += v;
          }
        }
      const double mean = s / double(cnt);
      if (w > mean * rank_limit) out = mean * rank_limit;
    }
    g.grid[i] = out;
  }
}

// imageweights.cpp:270-287: the full [height][width] image of the half grid
__global__ void __launch_bounds__(kBlock) GetGridKernel(WeightGrid g, double* __restrict__ img) {
  const int h2 = g.height / 2;
  const size_t n = size_t(h2) * g.width;
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock) {
    const int y = int(i / g.width), x = int(i - size_t(y) * g.width);
    const double v = g.grid[i];
    img[size_t(h2 - 1 - y) * g.width + (g.width - x - 1)] = v;
    img[size_t(h2 + y) * g.width + x] = v;
  }
}

__global__ void __launch_bounds__(kBlock) SampleWeightsKernel(WeightGrid g, const double* __restrict__ u,
                                                              const double* __restrict__ v, size_t n,
                                                              double* __restrict__ out) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock) {
    const long long c = WeightCell(u[i], v[i], g);
    out[i] = c >= 0 ? g.grid[c] : 0.0;
  }
}

// Row preparation: one thread per (row, channel).  Statistics partials per
// block: [0] total weight, [1] visibility weight sum, [2] max gridded
// weight, [3] gridded count.
__global__ void __launch_bounds__(kBlock) PrepareRowsKernel(
    PrepArgs a, const double* __restrict__ uvw, const double* __restrict__ lambda,
    const float2* __restrict__ vis, const float* __restrict__ weights,
    const float2* __restrict__ model, const uint8_t* __restrict__ selected, WeightGrid g,
    int have_grid, float2* __restrict__ out, double* __restrict__ part) {
#pragma clang fp contract(off)
  double tot = 0.0, vws = 0.0, mx = 0.0, cnt = 0.0;
  const size_t n = a.n_rows * a.n_chan;
  const int npol = a.n_pol;
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock) {
    const size_t row = i / a.n_chan, ch = i - row * a.n_chan;
    const double lam = lambda[ch];
    float2 v[4];
    if (a.psf) {
      float2 ph = make_float2(1.f, 0.f);
      if (a.psf_dl != 0.0 || a.psf_dm != 0.0) {
        // RotateVisibilities: exp(i 2 pi (u dl + v dm + w dn) / lambda)
        const double shift = 2.0 * M_PI *
                             (uvw[3 * row] * a.psf_dl + uvw[3 * row + 1] * a.psf_dm +
                              uvw[3 * row + 2] * a.psf_dn);
        const double rad = shift / lam;
        ph = make_float2(float(cos(rad)), float(sin(rad)));
      }
      for (int p = 0; p < npol; ++p) v[p] = ph;
    } else {
      for (int p = 0; p < npol; ++p) v[p] = vis[i * npol + p];
    }
    if (model)
      for (int p = 0; p < npol; ++p) {
        const float2 m = model[i * npol + p];
        v[p] = make_float2(v[p].x - m.x, v[p].y - m.y);
      }
    float iw = 1.f;
    if (have_grid) {
      const long long c = WeightCell(uvw[3 * row] / lam, uvw[3 * row + 1] / lam, g);
      iw = c >= 0 ? float(g.grid[c]) : 0.f;
    }
    const bool sel = selected ? selected[ch] != 0 : true;
    float2 acc = make_float2(0.f, 0.f);
    for (int p = 0; p < npol; ++p) {
      float w = sel ? weights[i * npol + p] : 0.f;
      if (a.weighting == 1) w = w * w;
      else if (a.weighting == 2 && w != 0.f) w = 1.f;
      const float cum = w * iw;
      const bool has = cum != 0.f;
      if (p == 0) {
        vws += double(w * float(has));
        mx = fmax(mx, double(cum));
        cnt += has ? 1.0 : 0.0;
      }
      tot += double(cum);
      v[p] = make_float2(v[p].x * cum, v[p].y * cum);
    }
    // polarization collapse (msgridderdata.h:37-52): 2 -> sum, 4 -> the
    // gridder's polarization from the linear correlations
    if (npol == 1) {
      acc = v[0];
    } else if (npol == 2) {
      acc = make_float2(v[0].x + v[1].x, v[0].y + v[1].y);
    } else {
      acc = FromLinear(v, a.polarization);
    }
    out[i] = acc;
  }
  __shared__ double sh[4][kBlock];
  sh[0][threadIdx.x] = tot;
  sh[1][threadIdx.x] = vws;
  sh[2][threadIdx.x] = mx;
  sh[3][threadIdx.x] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < kBlock; ++k) {
      r[0] += sh[0][k];
      r[1] += sh[1][k];
      r[2] = fmax(r[2], sh[2][k]);
      r[3] += sh[3][k];
    }
    for (int q = 0; q < 4; ++q) part[size_t(q) * kStatBlocks + blockIdx.x] = r[q];
  }
}

__global__ void PrepStatsKernel(const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x == 0) {
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < kStatBlocks; ++b) {
      r[0] += part[b];
      r[1] += part[kStatBlocks + b];
      r[2] = fmax(r[2], part[2 * kStatBlocks + b]);
      r[3] += part[3 * kStatBlocks + b];
    }
    for (int q = 0; q < 4; ++q) out[q] = r[q];
  }
}

}  // namespace

void LaunchGridWeights(const WeightGrid& g, const double* uvw, const double* lambda,
                       size_t n_rows, size_t n_chan, size_t n_pol, const float* weights,
                       bool as_taper, double* total_partial, hipStream_t s) {
  if (!n_rows || !n_chan) return;
  GridWeightsKernel<<<kStatBlocks, kBlock, 0, s>>>(g, uvw, lambda, n_rows, n_chan, n_pol,
                                                   weights, as_taper ? 1 : 0, total_partial);
  WG_HIP(hipGetLastError());
}

void LaunchFinishWeights(const WeightGrid& g, int mode, double robust, const double* total_partial,
                         double* sums, double* scratch, hipStream_t s) {
  const size_t n = size_t(g.height / 2) * g.width;
  SumPartialsKernel<<<1, kBlock, 0, s>>>(total_partial, kStatBlocks, sums, 0);
  if (mode == 2) {
    SquareSumKernel<<<kStatBlocks, kBlock, 0, s>>>(g.grid, n, scratch);
    SumPartialsKernel<<<1, kBlock, 0, s>>>(scratch, kStatBlocks, sums + 1, 0);
  }
  FinishKernel<<<Blocks(n), kBlock, 0, s>>>(g.grid, n, mode, sums, robust);
  WG_HIP(hipGetLastError());
}

void LaunchWeightTaper(const WeightGrid& g, int kind, double a, double b, hipStream_t s) {
  const size_t n = size_t(g.height / 2) * g.width;
  TaperKernel<<<Blocks(n), kBlock, 0, s>>>(g, kind, a, b);
  WG_HIP(hipGetLastError());
}

void LaunchRankFilter(const WeightGrid& g, const double* src, double rank_limit, int win,
                      hipStream_t s) {
  const size_t n = size_t(g.height / 2) * g.width;
  RankFilterKernel<<<Blocks(n), kBlock, 0, s>>>(g, src, rank_limit, win);
  WG_HIP(hipGetLastError());
}

void LaunchGetWeightGrid(const WeightGrid& g, double* img, hipStream_t s) {
  const size_t n = size_t(g.height / 2) * g.width;
  GetGridKernel<<<Blocks(n), kBlock, 0, s>>>(g, img);
  WG_HIP(hipGetLastError());
}

void LaunchSampleWeights(const WeightGrid& g, const double* u, const double* v, size_t n,
                         double* out, hipStream_t s) {
  if (!n) return;
  SampleWeightsKernel<<<Blocks(n), kBlock, 0, s>>>(g, u, v, n, out);
  WG_HIP(hipGetLastError());
}

void LaunchPrepareRows(const PrepArgs& a, const double* uvw, const double* lambda,
                       const float2* vis, const float* weights, const float2* model,
                       const uint8_t* selected, const WeightGrid* g, float2* out, double* part,
                       double* stats, hipStream_t s) {
  WG_CHECK(a.n_pol == 1 || a.n_pol == 2 || a.n_pol == 4, "1, 2 or 4 polarizations");
  WeightGrid none{};
  PrepareRowsKernel<<<kStatBlocks, kBlock, 0, s>>>(a, uvw, lambda, vis, weights, model, selected,
                                                   g ? *g : none, g ? 1 : 0, out, part);
  PrepStatsKernel<<<1, 64, 0, s>>>(part, stats);
  WG_HIP(hipGetLastError());
}

size_t WeightStatPartials() { return size_t(4) * kStatBlocks; }

}  // namespace wg
