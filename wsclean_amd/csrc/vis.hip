// Per-visibility kernels: coordinate extents, bucket keys, sorted
// coordinates, the visibility gather (invert) / scatter (predict) with the
// phase-centre shift, and the per-pixel grid correction.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "device.hpp"
#include "engine.hpp"

namespace wg {

namespace {

constexpr int kBlock = 256;

// min / max that keep a NaN operand (fmin / fmax return the other one), so
// a NaN coordinate reaches the extents and the host's finiteness check
__device__ __forceinline__ double MinKeepNaNd(double a, double b) {
  return (a < b || a != a) ? a : b;
}
__device__ __forceinline__ double MaxKeepNaNd(double a, double b) {
  return (a > b || a != a) ? a : b;
}
__device__ __forceinline__ double WaveMin(double v) {
  for (int o = 32; o > 0; o >>= 1) v = MinKeepNaNd(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double WaveMax(double v) {
  for (int o = 32; o > 0; o >>= 1) v = MaxKeepNaNd(v, __shfl_xor(v, o, 64));
  return v;
}

// Per block: min/max of the flipped u, v, w (wavelengths) over its rows.
__global__ void __launch_bounds__(kBlock) ExtentsKernel(const double* __restrict__ uvw,
                                                        size_t n_rows, double smin, double smax,
                                                        double* __restrict__ partial) {
  double r[6] = {DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX};
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n_rows;
       i += size_t(gridDim.x) * kBlock) {
#pragma clang fp contract(off)
    double u = uvw[3 * i], v = uvw[3 * i + 1], w = uvw[3 * i + 2];
    if (w < 0.0) {
      u = -u;
      v = -v;
      w = -w;
    }
    const double u1 = u * smin, u2 = u * smax;
    const double v1 = v * smin, v2 = v * smax;
    const double w1 = w * smin, w2 = w * smax;
    r[0] = MinKeepNaNd(r[0], MinKeepNaNd(u1, u2));
    r[1] = MaxKeepNaNd(r[1], MaxKeepNaNd(u1, u2));
    r[2] = MinKeepNaNd(r[2], MinKeepNaNd(v1, v2));
    r[3] = MaxKeepNaNd(r[3], MaxKeepNaNd(v1, v2));
    r[4] = MinKeepNaNd(r[4], MinKeepNaNd(w1, w2));
    r[5] = MaxKeepNaNd(r[5], MaxKeepNaNd(w1, w2));
  }
  __shared__ double sh[kBlock / 64][6];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 0; k < 6; ++k) {
    const double v = (k & 1) ? WaveMax(r[k]) : WaveMin(r[k]);
    if (lane == 0) sh[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    double v = sh[0][k];
    for (int w = 1; w < kBlock / 64; ++w)
      v = (k & 1) ? MaxKeepNaNd(v, sh[w][k]) : MinKeepNaNd(v, sh[w][k]);
    partial[size_t(blockIdx.x) * 6 + k] = v;
  }
}

// The blocks' partial extents reduced by one block: each thread takes every
// kBlock-th partial, then the waves and the block combine (min / max are
// exact, so the order does not matter).
__global__ void __launch_bounds__(kBlock) ExtentsFinalKernel(const double* __restrict__ partial,
                                                             int nblocks,
                                                             double* __restrict__ out) {
  double r[6] = {DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX, DBL_MAX, -DBL_MAX};
  for (int b = threadIdx.x; b < nblocks; b += kBlock) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double v = partial[size_t(b) * 6 + k];
      r[k] = (k & 1) ? MaxKeepNaNd(r[k], v) : MinKeepNaNd(r[k], v);
    }
  }
  __shared__ double sh[kBlock / 64][6];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double v = (k & 1) ? WaveMax(r[k]) : WaveMin(r[k]);
    if (lane == 0) sh[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    double v = sh[0][k];
    for (int w = 1; w < kBlock / 64; ++w)
      v = (k & 1) ? MaxKeepNaNd(v, sh[w][k]) : MinKeepNaNd(v, sh[w][k]);
    out[k] = v;
  }
}

// Sort key of one visibility: ((tile * P0 + p0) * key_T + cell row) * key_T +
// cell column (SortKey).

__device__ __forceinline__ uint32_t VisKey(const VisLoc& L, const DevPlan& dp) {
  const uint32_t bucket = uint32_t(L.tile) * uint32_t(dp.P0) + uint32_t(L.p0);
  return SortKey(bucket, L.du, L.dv, dp.key_shift, dp.key_T);
}

__global__ void __launch_bounds__(kBlock) KeysKernel(DevPlan dp, const double* __restrict__ uvw,
                                                     const double* __restrict__ scale,
                                                     size_t n_rows, size_t n_chan,
                                                     const float2* __restrict__ vis,
                                                     uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals) {
  const size_t n = n_rows * n_chan;
  const double inv_chan = 1.0 / double(n_chan);
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock) {
    const size_t row = DivSmall(uint32_t(i), uint32_t(n_chan), inv_chan), ch = i - row * n_chan;
    uint32_t key = dp.nkeys;
    bool valid = true;
    if (vis) {
      const float2 V = vis[i];
      valid = (V.x != 0.f) || (V.y != 0.f);
    }
    if (valid)
      key = VisKey(LocateVis(uvw[3 * row], uvw[3 * row + 1], uvw[3 * row + 2], scale[ch], dp), dp);
    keys[i] = key;
    if (vals) vals[i] = uint32_t(i);
  }
}

// Row maxima of |V| from one wave's 64 consecutive elements: lane L holds
// the squared modulus m2 of an element of row `row` (~0u past the end), and
// the lanes of one row are consecutive.  Each row segment's max of
// sqrt(m2) goes to rowmax_bits[row] with atomicMax on the bits of a
// non-negative float (the words start at zero), which also joins a row
// split between waves.  sqrt is monotonic and correctly rounded, so this is
// max |V| of the row exactly.  Every lane of the wave takes part.
__device__ __forceinline__ void RowSegmentMax(float m2, uint32_t row,
                                              uint32_t* __restrict__ rowmax_bits) {
  const int lane = threadIdx.x & 63;
  const uint32_t r0 = __builtin_amdgcn_readfirstlane(row);
  if (__all(row == r0)) {
    for (int m = 32; m >= 1; m >>= 1) m2 = MaxKeepNaN(m2, __shfl_xor(m2, m));
    if (lane == 0 && r0 != ~0u) atomicMax(rowmax_bits + r0, __float_as_uint(sqrtf(m2)));
    return;
  }
  // lane L ends with the max over [L, end of its segment): windows that
  // double in length, taken only from lanes of the same row
  for (int off = 1; off < 64; off <<= 1) {
    const float o = __shfl_down(m2, off);
    const uint32_t ro = __shfl_down(row, off);
    if (lane + off < 64 && ro == row) m2 = MaxKeepNaN(m2, o);
  }
  const uint32_t rp = __shfl_up(row, 1);
  if ((lane == 0 || rp != row) && row != ~0u)
    atomicMax(rowmax_bits + row, __float_as_uint(sqrtf(m2)));
}

// Keys in renumbered row order: element e = r * n_chan + ch is channel ch of
// renumbered row r, i.e. of caller row RowTag(uvw4, r); uvw4 holds the
// renumbered rows padded to 32 bytes.  Keys are the same as KeysKernel's for
// the same visibility, only their order differs.  ROWMAX (with vis): also
// max |V| of every renumbered row (the values RowMaxKernel gives with the
// rows' caller map) from the same reads; its loop steps whole waves, so
// every lane takes part in the row maxima.
template <bool ROWMAX>
__global__ void __launch_bounds__(kBlock) KeysRenumberedKernel(
    DevPlan dp, const double* __restrict__ uvw4,
    const double* __restrict__ scale, size_t n_rows, size_t n_chan,
    const float2* __restrict__ vis, uint32_t* __restrict__ keys,
    uint32_t* __restrict__ rowmax_bits) {
  const size_t n = n_rows * n_chan;
  const double inv_chan = 1.0 / double(n_chan);
  const size_t lane = ROWMAX ? threadIdx.x & 63 : threadIdx.x;
  const size_t first = size_t(blockIdx.x) * kBlock + (ROWMAX ? (threadIdx.x & ~63u) : 0u);
  for (size_t e0 = first; ROWMAX ? e0 < n : e0 + lane < n; e0 += size_t(gridDim.x) * kBlock) {
    const size_t e = e0 + lane;
    const bool in = !ROWMAX || e < n;
    float m2 = 0.f;
    uint32_t rrow = ~0u;
    if (in) {
      const size_t r = DivSmall(uint32_t(e), uint32_t(n_chan), inv_chan), ch = e - r * n_chan;
      uint32_t key = dp.nkeys;
      bool valid = true;
      const double2 uv = reinterpret_cast<const double2*>(uvw4)[2 * r];
      const double2 wt = reinterpret_cast<const double2*>(uvw4)[2 * r + 1];
      if (vis) {
        const float2 V = vis[size_t(RowTag(wt)) * n_chan + ch];
        valid = (V.x != 0.f) || (V.y != 0.f);
        m2 = SquaredModulus(V);
        rrow = uint32_t(r);
      }
      if (valid) key = VisKey(LocateVis(uv.x, uv.y, wt.x, scale[ch], dp), dp);
      keys[e] = key;
    }
    if constexpr (ROWMAX) RowSegmentMax(m2, rrow, rowmax_bits);
  }
}

// rowmax_bits[r] = bits of max |V| over the n_chan visibilities of row
// row_map[r] (row r without a map) of vis; the words start at zero.  Each
// wave step reads four 64-element chunks before reducing them.
__global__ void __launch_bounds__(kBlock) RowMaxKernel(const float2* __restrict__ vis,
                                                      const uint32_t* __restrict__ row_map,
                                                      size_t n_rows, size_t n_chan,
                                                      uint32_t* __restrict__ rowmax_bits) {
  const size_t n = n_rows * n_chan;
  const double inv_chan = 1.0 / double(n_chan);
  const size_t lane = threadIdx.x & 63;
  const size_t stride = size_t(gridDim.x) * kBlock;
  for (size_t e0 = size_t(blockIdx.x) * kBlock + (threadIdx.x & ~63u); e0 < n;
       e0 += 4 * stride) {
    float2 v[4];
    uint32_t row[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const size_t e = e0 + k * stride + lane;
      row[k] = ~0u;
      v[k] = make_float2(0.f, 0.f);
      if (e < n) {
        const uint32_t r = DivSmall(uint32_t(e), uint32_t(n_chan), inv_chan);
        const size_t src = row_map ? row_map[r] : r;
        v[k] = vis[src * n_chan + (e - size_t(r) * n_chan)];
        row[k] = r;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (e0 + k * stride >= n) break;  // the same for every lane
      RowSegmentMax(SquaredModulus(v[k]), row[k], rowmax_bits);
    }
  }
}

// As RowMaxKernel without a row map, for an even n_chan: each lane reads
// two visibilities of one row with a 16-byte load (no pair spans two rows),
// four 128-element chunks per wave step.
__global__ void __launch_bounds__(kBlock) RowMaxPairsKernel(const float4* __restrict__ vis2,
                                                           size_t n_rows, size_t n_chan,
                                                           uint32_t* __restrict__ rowmax_bits) {
  const size_t npair = n_rows * n_chan / 2;
  const double inv_chan = 1.0 / double(n_chan);
  const size_t lane = threadIdx.x & 63;
  const size_t stride = size_t(gridDim.x) * kBlock;
  for (size_t p0 = size_t(blockIdx.x) * kBlock + (threadIdx.x & ~63u); p0 < npair;
       p0 += 4 * stride) {
    float4 v[4];
    uint32_t row[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const size_t p = p0 + k * stride + lane;
      row[k] = ~0u;
      v[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (p < npair) {
        v[k] = vis2[p];
        row[k] = DivSmall(uint32_t(2 * p), uint32_t(n_chan), inv_chan);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (p0 + k * stride >= npair) break;  // the same for every lane
      RowSegmentMax(MaxKeepNaN(SquaredModulus(make_float2(v[k].x, v[k].y)),
                          SquaredModulus(make_float2(v[k].z, v[k].w))),
                    row[k], rowmax_bits);
    }
  }
}

// out[r] = in[rows[r]]: per-row values in another row order.
__global__ void __launch_bounds__(kBlock) GatherRowsKernel(const float* __restrict__ in,
                                                          const uint32_t* __restrict__ rows,
                                                          size_t n, float* __restrict__ out) {
  for (size_t r = size_t(blockIdx.x) * kBlock + threadIdx.x; r < n;
       r += size_t(gridDim.x) * kBlock)
    out[r] = in[rows[r]];
}

// bounds[it] = (max, sum) over work item it's visibilities of a bound b_i
// on |Re| and |Im| of each rotated visibility: SORTED, the larger component
// of vis_sorted[i] itself; else max |V| of the visibility's renumbered row
// (rowmax[vrow[i]]), which bounds both components of V times a unit
// phasor.  Rows are renumbered in uv order, so an item reads the maxima of
// a few short row ranges.  One block per item; the sums are formed in a
// fixed order, so the bounds (and the gridding scales taken from them)
// repeat bitwise.
template <bool SORTED>
__global__ void __launch_bounds__(256) ItemBoundsKernel(const WorkItem* __restrict__ items,
                                                       const float2* __restrict__ vis_sorted,
                                                       const uint32_t* __restrict__ vrow,
                                                       const float* __restrict__ rowmax,
                                                       float2* __restrict__ bounds) {
  __shared__ float part[2][4];
  const WorkItem wi = items[blockIdx.x];
  auto bound = [&](uint32_t i) {
    if constexpr (SORTED) {
      const float2 v = vis_sorted[i];
      return MaxKeepNaN(fabsf(v.x), fabsf(v.y));
    } else {
      return rowmax[vrow[i]];
    }
  };
  float mx = 0.f, sm = 0.f;
  uint32_t i = wi.begin + threadIdx.x;
  for (; i + 3u * 256u < wi.end; i += 4u * 256u) {
    float b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) b[k] = bound(i + 256u * k);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      mx = MaxKeepNaN(mx, b[k]);
      sm += b[k];
    }
  }
  for (; i < wi.end; i += 256u) {
    const float b = bound(i);
    mx = MaxKeepNaN(mx, b);
    sm += b;
  }
  for (int m = 32; m >= 1; m >>= 1) {
    mx = MaxKeepNaN(mx, __shfl_xor(mx, m));
    sm += __shfl_xor(sm, m);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[0][wave] = mx;
    part[1][wave] = sm;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    bounds[blockIdx.x] =
        make_float2(MaxKeepNaN(MaxKeepNaN(part[0][0], part[0][1]),
                               MaxKeepNaN(part[0][2], part[0][3])),
                    (part[1][0] + part[1][1]) + (part[1][2] + part[1][3]));
}

// uvw4[r] = (uvw[rfetch[r]], tag): the renumbered rows padded to 32 bytes,
// the fourth word holding rtag[r] in its low half (read back by RowTag) and
// rhi[r] (or 0) in its high half (RowHi), so that kernels reading a
// renumbered row find its caller row -- and a facet batch its facet -- in
// the same cache line.
__global__ void __launch_bounds__(kBlock) PadUvwKernel(const double* __restrict__ uvw,
                                                       const uint32_t* __restrict__ rfetch,
                                                       const uint32_t* __restrict__ rtag,
                                                       const uint32_t* __restrict__ rhi,
                                                       size_t n_rows,
                                                       double* __restrict__ uvw4) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < 4 * n_rows;
       i += size_t(gridDim.x) * kBlock) {
    const size_t r = i >> 2;
    const int k = int(i & 3);
    const uint64_t tag = uint64_t(rtag[r]) | (rhi ? uint64_t(rhi[r]) << 32 : 0ull);
    uvw4[i] = k < 3 ? uvw[3 * size_t(rfetch[r]) + k] : __longlong_as_double((long long)tag);
  }
}

__global__ void __launch_bounds__(kBlock) MaskKernel(const uint32_t* __restrict__ in, size_t n,
                                                     uint32_t mask, uint32_t* __restrict__ out) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock)
    out[i] = in[i] & mask;
}

__global__ void __launch_bounds__(kBlock) GatherVisKernel(
    const uint32_t* __restrict__ perm, const float2* __restrict__ phasor, size_t n_valid,
    const float2* __restrict__ vis, float2* __restrict__ out) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n_valid;
       i += size_t(gridDim.x) * kBlock) {
    const uint32_t pw = perm[i];
    float2 V = vis[pw & ~kPermFlip];
    if (pw & kPermFlip) V.y = -V.y;
    const float2 ph = phasor[i];
    out[i] = make_float2(V.x * ph.x - V.y * ph.y, V.x * ph.y + V.y * ph.x);
  }
}

__global__ void __launch_bounds__(kBlock) ScatterVisKernel(
    const uint32_t* __restrict__ perm, const float2* __restrict__ phasor, size_t n_valid,
    const float2* __restrict__ acc, float2* __restrict__ out) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n_valid;
       i += size_t(gridDim.x) * kBlock) {
    const uint32_t pw = perm[i];
    const float2 a = acc[i];
    const float2 ph = phasor[i];  // conjugated below: the predict direction
    float2 V = make_float2(a.x * ph.x + a.y * ph.y, a.y * ph.x - a.x * ph.y);
    if (pw & kPermFlip) V.y = -V.y;
    out[pw & ~kPermFlip] = V;
  }
}

// corr[x][y] = (-1)^(q_l + q_m) / (Phi_u(q_l/nu) Phi_v(q_m/nv) Phi_w(dw (n-1-zc)))
// Phi_w is integrated per pixel in FP64: sum_i a_i cos(b_i t).
__global__ void __launch_bounds__(kBlock) CorrectionKernel(
    DevPlan dp, const double* __restrict__ phi_u, const double* __restrict__ phi_v,
    const double* __restrict__ qa, const double* __restrict__ qb, int nq,
    float* __restrict__ corr) {
  const size_t n = size_t(dp.nx) * dp.ny;
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock) {
    const int x = int(i / dp.ny), y = int(i % dp.ny);
    const double l = double(dp.nx / 2 - x) * dp.dx_rad + dp.l_shift;
    const double m = double(y - dp.ny / 2) * dp.dy_rad + dp.m_shift;
    const double r2 = l * l + m * m;
    const double z = (r2 < 1.0 ? -r2 / (sqrt(1.0 - r2) + 1.0) : 0.0) - dp.zc;
    const double t = dp.dw * z;
    double pw = 0.0;
    for (int q = 0; q < nq; ++q) pw += qa[q] * cos(qb[q] * t);
    const int ql = dp.nx / 2 - x, qm = y - dp.ny / 2;
    const double sgn = ((ql + qm) & 1) ? -1.0 : 1.0;
    corr[i] = float(sgn / (phi_u[x] * phi_v[y] * pw));
  }
}

// phi_u[x] = Phi((nx/2 - x) / nu), phi_v[y] = Phi((y - ny/2) / nv), FP64.
__global__ void PhiTablesKernel(DevPlan dp, const double* __restrict__ qa,
                                const double* __restrict__ qb, int nq, double* __restrict__ pu,
                                double* __restrict__ pv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double t;
  double* out;
  if (i < dp.nx) {
    t = double(dp.nx / 2 - i) / double(dp.nu);
    out = pu + i;
  } else if (i < dp.nx + dp.ny) {
    t = double(i - dp.nx - dp.ny / 2) / double(dp.nv);
    out = pv + (i - dp.nx);
  } else {
    return;
  }
  double s = 0.0;
  for (int q = 0; q < nq; ++q) s += qa[q] * cos(qb[q] * t);
  *out = s;
}

__global__ void AccumulateImageKernel(const float* __restrict__ dirty,
                                      const float* __restrict__ corr, double* __restrict__ img,
                                      size_t n) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock)
    img[i] += double(dirty[i]) * double(corr[i]);
}

__global__ void StridedGatherKernel(const uint32_t* __restrict__ in, size_t stride,
                                    size_t count, uint32_t* __restrict__ out) {
  const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < count) out[k] = in[k * stride];
}

// ---- per-antenna gain corrections (correction callback) ----
// 2x2 complex matrices with elements [0] xx, [1] xy, [2] yx, [3] yy.
struct C2x2 {
  float2 e[4];
};
__device__ __forceinline__ float2 cm(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 ca(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
// A B (AH: A^H B, BH: A B^H)
template <bool AH, bool BH>
__device__ __forceinline__ C2x2 MatMul(const C2x2& A, const C2x2& B) {
  C2x2 R;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      float2 s = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const float2 a = AH ? cj(A.e[k * 2 + i]) : A.e[i * 2 + k];
        const float2 b = BH ? cj(B.e[j * 2 + k]) : B.e[k * 2 + j];
        s = ca(s, cm(a, b));
      }
      R.e[i * 2 + j] = s;
    }
  return R;
}

__device__ __forceinline__ C2x2 LoadJones(const float2* __restrict__ p, int nparms) {
  C2x2 g;
  if (nparms == 2) {
    g.e[0] = p[0];
    g.e[1] = g.e[2] = make_float2(0.f, 0.f);
    g.e[3] = p[1];
  } else {
    for (int i = 0; i < 4; ++i) g.e[i] = p[i];
  }
  return g;
}

// ApplyGain (conj = false) / ApplyConjugatedGain (conj = true) of
// gridding/visibilitymodifier.h:346-401 on the n_pol values of v.
__device__ __forceinline__ void ApplyJones(int mode, bool conj, float2* v, const C2x2& a,
                                           const C2x2& b) {
  switch (mode) {
    case 0:  // xx
      v[0] = conj ? cm(cm(cj(a.e[0]), v[0]), b.e[0]) : cm(cm(a.e[0], v[0]), cj(b.e[0]));
      break;
    case 1:  // yy
      v[0] = conj ? cm(cm(cj(a.e[3]), v[0]), b.e[3]) : cm(cm(a.e[3], v[0]), cj(b.e[3]));
      break;
    case 2: {  // trace: 0.5 * sum_i x_i conj(y_i)
      float2 dd = make_float2(0.f, 0.f);
      for (int i = 0; i < 4; ++i)
        dd = ca(dd, conj ? cm(b.e[i], cj(a.e[i])) : cm(a.e[i], cj(b.e[i])));
      v[0] = cm(v[0], make_float2(0.5f * dd.x, 0.5f * dd.y));
      break;
    }
    case 3:  // two diagonal visibilities; the second uses the updated first
      if (conj) {
        v[0] = ca(cm(cm(cj(a.e[0]), v[0]), b.e[0]), cm(cm(cj(a.e[2]), v[1]), b.e[2]));
        v[1] = ca(cm(cm(cj(a.e[3]), v[1]), b.e[3]), cm(cm(cj(a.e[1]), v[0]), b.e[1]));
      } else {
        v[0] = ca(cm(cm(a.e[0], v[0]), cj(b.e[0])), cm(cm(a.e[1], v[1]), cj(b.e[1])));
        v[1] = ca(cm(cm(a.e[3], v[1]), cj(b.e[3])), cm(cm(a.e[2], v[0]), cj(b.e[2])));
      }
      break;
    default: {  // full 2x2: a^H V b, or a V b^H
      C2x2 V;
      for (int i = 0; i < 4; ++i) V.e[i] = v[i];
      const C2x2 R = conj ? MatMul<false, false>(MatMul<true, false>(a, V), b)
                          : MatMul<false, true>(MatMul<false, false>(a, V), b);
      for (int i = 0; i < 4; ++i) v[i] = R.e[i];
    }
  }
}

// out[row][ch] = collapse(conj-gain(forward-gain(vis[row][ch][0..npol))))
__global__ void __launch_bounds__(kBlock) CorrectCollapseKernel(
    const float2* __restrict__ vis, const uint32_t* __restrict__ antennas,
    const uint64_t* __restrict__ time_offsets, const float2* __restrict__ gains,
    size_t n_rows, size_t n_chan, size_t n_ant, int nparms, int mode, int npol, int forward,
    int pol, float2* __restrict__ out) {
  const size_t n = n_rows * n_chan;
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock) {
    const size_t row = i / n_chan, ch = i - row * n_chan;
    const size_t base = (size_t(time_offsets[row]) * n_chan + ch) * n_ant;
    const C2x2 a = LoadJones(gains + (base + antennas[2 * row]) * nparms, nparms);
    const C2x2 b = LoadJones(gains + (base + antennas[2 * row + 1]) * nparms, nparms);
    float2 v[4];
    for (int p = 0; p < npol; ++p) v[p] = vis[i * npol + p];
    if (forward) ApplyJones(mode, false, v, a, b);
    ApplyJones(mode, true, v, a, b);
    // collapse (gridding/msgridderdata.h:37-52): 2 -> sum, 4 -> the
    // gridder's polarization from the linear correlations
    float2 r = v[0];
    if (npol == 2) r = ca(v[0], v[1]);
    if (npol == 4) r = FromLinear(v, pol);
    out[i] = r;
  }
}

__global__ void StrideRowsKernel(const uint64_t* __restrict__ in, size_t stride, size_t n_out,
                                 size_t row_words, uint64_t* __restrict__ out) {
  const size_t n = n_out * row_words;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const size_t r = i / row_words, k = i - r * row_words;
    out[i] = in[r * stride * row_words + k];
  }
}

__global__ void FillZeroKernel(uint32_t* __restrict__ p, size_t nwords) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < nwords;
       i += size_t(gridDim.x) * kBlock)
    p[i] = 0u;
}

unsigned Blocks(size_t n) {
  size_t b = (n + kBlock - 1) / kBlock;
  return unsigned(b < 65536 ? (b ? b : 1) : 65536);
}

// splitmix64 finaliser: counter-based, so the noise of visibility i does not
// depend on the launch shape.
__device__ __forceinline__ uint64_t Mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// V(row, chan) = sum_s S_s g_s(u, v) exp(+2 pi i (u l_s + v m_s + w (n_s - 1)))
// + complex Gaussian noise; g_s = exp(-2 pi^2 sigma_s^2 (u^2 + v^2)).
// Sources are [l, m, n - 1, S, sigma] in FP64; phases are reduced in FP64.
constexpr int kMaxSrcLds = 1024;
__global__ void __launch_bounds__(kBlock) SynthKernel(
    const double* __restrict__ uvw, size_t n_rows, size_t n_chan,
    const double* __restrict__ scale, const double* __restrict__ src, int n_src,
    double noise, uint64_t seed, float2* __restrict__ vis) {
  __shared__ double s_src[kMaxSrcLds * 5];
  for (int i = threadIdx.x; i < n_src * 5; i += kBlock) s_src[i] = src[i];
  __syncthreads();
  const size_t n = n_rows * n_chan;
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += size_t(gridDim.x) * kBlock) {
    const size_t row = i / n_chan, ch = i - row * n_chan;
    const double s = scale[ch];
    const double u = uvw[3 * row] * s, v = uvw[3 * row + 1] * s, w = uvw[3 * row + 2] * s;
    const double uv2 = u * u + v * v;
    double re = 0.0, im = 0.0;
    for (int k = 0; k < n_src; ++k) {
      const double* q = s_src + 5 * k;
      const double cyc = u * q[0] + v * q[1] + w * q[2];
      const double f = cyc - rint(cyc);
      double amp = q[3];
      if (q[4] > 0.0) amp *= exp(-2.0 * M_PI * M_PI * q[4] * q[4] * uv2);
      float sn, cs;
      sincospif(2.0f * float(f), &sn, &cs);
      re += amp * cs;
      im += amp * sn;
    }
    if (noise > 0.0) {
      const uint64_t a = Mix64(seed ^ (uint64_t(i) * 2ull));
      const uint64_t b = Mix64(seed ^ (uint64_t(i) * 2ull + 1ull));
      const double u1 = (double(a >> 11) + 1.0) * (1.0 / 9007199254740993.0);
      const double u2 = double(b >> 11) * (1.0 / 9007199254740992.0);
      const double r = sqrt(-2.0 * log(u1)) * noise;
      re += r * cos(2.0 * M_PI * u2);
      im += r * sin(2.0 * M_PI * u2);
    }
    vis[i] = make_float2(float(re), float(im));
  }
}


// ------------------------------------------------------------ facet batches
// Facet f owns the virtual visibility indices [vbase[f], vbase[f + 1]); its
// visibility (row r, channel c) is the physical entry
// (row_off_f + r) * n_chan + c of the shared uvw / visibility buffers.
__device__ __forceinline__ int FacetOf(const unsigned long long* __restrict__ vbase, int F,
                                       unsigned long long v) {
  int lo = 0, hi = F - 1;  // the last facet with vbase <= v
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (vbase[mid] <= v) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Keys of a facet batch: every facet maps its visibilities with its own
// plan, and its tiles are placed at (tu0, tv0) of the atlas tile grid, so
// one sort and one set of work items covers all facets.
__global__ void __launch_bounds__(kBlock) FacetKeysKernel(
    const FacetDev* __restrict__ fac, const unsigned long long* __restrict__ vbase, int F,
    const double* __restrict__ uvw, const double* __restrict__ scale, size_t n_chan,
    size_t n_virtual, int atlas_ntu, int atlas_P0, int key_shift, int key_T,
    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n_virtual;
       i += size_t(gridDim.x) * kBlock) {
    const int f = FacetOf(vbase, F, i);
    const size_t local = i - vbase[f];
    const size_t row = fac[f].row_off + local / n_chan, ch = local % n_chan;
    const DevPlan& dp = fac[f].dp;
    const VisLoc L = LocateVis(uvw[3 * row], uvw[3 * row + 1], uvw[3 * row + 2], scale[ch], dp);
    const int tu = L.tile % dp.ntu, tv = L.tile / dp.ntu;
    const uint32_t tile = uint32_t(fac[f].tv0 + tv) * uint32_t(atlas_ntu) + uint32_t(fac[f].tu0 + tu);
    const uint32_t bucket = tile * uint32_t(atlas_P0) + uint32_t(L.p0);
    keys[i] = SortKey(bucket, L.du, L.dv, key_shift, key_T);
    if (vals) vals[i] = uint32_t(i);
  }
}

// Keys of a facet batch in renumbered row order: element e = r * n_chan +
// ch is channel ch of renumbered virtual row r, i.e. of virtual row
// RowTag(uvw4, r) (virtual rows of facet f: [vrbase[f], vrbase[f + 1]),
// physical row row_off_f + local); uvw4 holds the renumbered rows padded to
// 32 bytes.
__global__ void __launch_bounds__(kBlock) FacetKeysRenumberedKernel(
    const FacetDev* __restrict__ fac, const unsigned long long* __restrict__ vrbase, int F,
    const double* __restrict__ uvw4,
    const double* __restrict__ scale, size_t n_vrows, size_t n_chan, int atlas_ntu,
    int atlas_P0, int key_shift, int key_T, uint32_t* __restrict__ keys) {
  const size_t n = n_vrows * n_chan;
  const double inv_chan = 1.0 / double(n_chan);
  for (size_t e = size_t(blockIdx.x) * kBlock + threadIdx.x; e < n;
       e += size_t(gridDim.x) * kBlock) {
    const size_t r = DivSmall(uint32_t(e), uint32_t(n_chan), inv_chan), ch = e - r * n_chan;
    const double2 uv = reinterpret_cast<const double2*>(uvw4)[2 * r];
    const double2 wt = reinterpret_cast<const double2*>(uvw4)[2 * r + 1];
    const int f = int(RowHi(wt));  // the row's facet, from its padded entry
    auto key = [&](const FacetDev& fd) {
      const DevPlan& dp = fd.dp;
      const VisLoc L = LocateVis(uv.x, uv.y, wt.x, scale[ch], dp);
      const int tu = L.tile % dp.ntu, tv = L.tile / dp.ntu;
      const uint32_t tile = uint32_t(fd.tv0 + tv) * uint32_t(atlas_ntu) + uint32_t(fd.tu0 + tu);
      const uint32_t bucket = tile * uint32_t(atlas_P0) + uint32_t(L.p0);
      return SortKey(bucket, L.du, L.dv, key_shift, key_T);
    };
    // a wave's elements mostly share one facet: its plan is then read with
    // scalar loads
    const int f0 = __builtin_amdgcn_readfirstlane(f);
    keys[e] = __all(f == f0) ? key(fac[f0]) : key(fac[f]);
  }
}

// Physical row and facet of every renumbered virtual row: rphys[r] =
// row_off_f + (rperm[r] - vrbase[f]), rfac[r] = f.
__global__ void __launch_bounds__(kBlock) FacetRowMapKernel(
    const FacetDev* __restrict__ fac, const unsigned long long* __restrict__ vrbase, int F,
    const uint32_t* __restrict__ rperm, size_t n_vrows, uint32_t* __restrict__ rphys,
    uint32_t* __restrict__ rfac) {
  for (size_t r = size_t(blockIdx.x) * kBlock + threadIdx.x; r < n_vrows;
       r += size_t(gridDim.x) * kBlock) {
    const uint32_t vr = rperm[r];
    const int f = FacetOf(vrbase, F, vr);
    rphys[r] = uint32_t(fac[f].row_off + (vr - vrbase[f]));
    rfac[r] = uint32_t(f);
  }
}

// Sorted coordinates and phasors of a facet batch (as the visibility sort's
// coordinates epilogue, SortedCoords in sort.hip, with
// the facet's own plan, phase centre shift and w centre).  perm enters with
// sorted renumbered elements; it leaves with the virtual index (the predict
// output layout) and perm_phys with the physical index into the shared
// visibility buffer; both carry the flip bit.  With vrow, each sorted
// element's renumbered virtual row is kept.
__global__ void __launch_bounds__(kBlock) FacetCoordsKernel(
    const FacetDev* __restrict__ fac, const unsigned long long* __restrict__ vrbase, int F,
    const double* __restrict__ uvw4,
    const double* __restrict__ scale, size_t n_chan, uint32_t* __restrict__ perm,
    uint32_t* __restrict__ perm_phys, size_t n_valid, float4* __restrict__ coord,
    float2* __restrict__ phasor, uint32_t* __restrict__ vrow) {
  const double inv_chan = 1.0 / double(n_chan);
  for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n_valid;
       i += size_t(gridDim.x) * kBlock) {
    const uint32_t e = perm[i];
    const size_t r = DivSmall(e, uint32_t(n_chan), inv_chan), ch = e - r * n_chan;
    const double2 uv = reinterpret_cast<const double2*>(uvw4)[2 * r];
    const double2 wt = reinterpret_cast<const double2*>(uvw4)[2 * r + 1];
    const uint32_t vr = RowTag(wt);
    const int f = int(RowHi(wt));  // the row's facet, from its padded entry
    if (vrow) vrow[i] = uint32_t(r);
    auto write = [&](const FacetDev& fd, unsigned long long vrb) {
      const DevPlan& dp = fd.dp;
      const VisLoc L = LocateVis(uv.x, uv.y, wt.x, scale[ch], dp);
      coord[i] = make_float4(L.fu, L.fv, L.fw, __uint_as_float(PackLoc(L)));
      phasor[i] = ShiftPhasor(dp, L, -1.f);
      const uint32_t flip = L.flip ? kPermFlip : 0u;
      const size_t row = fd.row_off + (vr - vrb);
      perm[i] = uint32_t(size_t(vr) * n_chan + ch) | flip;
      perm_phys[i] = uint32_t(row * n_chan + ch) | flip;
    };
    // neighbouring sorted visibilities mostly share one facet: its plan is
    // then read with scalar loads
    const int f0 = __builtin_amdgcn_readfirstlane(f);
    if (__all(f == f0))
      write(fac[f0], vrbase[f0]);
    else
      write(fac[f], vrbase[f]);
  }
}

}  // namespace

void LaunchSynth(const double* uvw, size_t n_rows, size_t n_chan, const double* scale,
                 const double* src, int n_src, double noise, uint64_t seed, float2* vis,
                 hipStream_t s) {
  WG_CHECK(n_src >= 0 && n_src <= kMaxSrcLds, "at most 1024 synthetic sources");
  SynthKernel<<<Blocks(n_rows * n_chan), kBlock, 0, s>>>(uvw, n_rows, n_chan, scale, src, n_src,
                                                         noise, seed, vis);
  WG_HIP(hipGetLastError());
}

void LaunchExtents(const double* uvw, size_t n_rows, double smin, double smax,
                   double* partial, double* out, hipStream_t s) {
  const unsigned nb = std::min<unsigned>(Blocks(n_rows), 1024);
  ExtentsKernel<<<nb, kBlock, 0, s>>>(uvw, n_rows, smin, smax, partial);
  ExtentsFinalKernel<<<1, kBlock, 0, s>>>(partial, int(nb), out);
  WG_HIP(hipGetLastError());
}

void LaunchKeys(const DevPlan& dp, const double* uvw, const double* scale, size_t n_rows,
                size_t n_chan, const float2* vis, uint32_t* keys, uint32_t* vals,
                hipStream_t s) {
  KeysKernel<<<Blocks(n_rows * n_chan), kBlock, 0, s>>>(dp, uvw, scale, n_rows, n_chan, vis,
                                                       keys, vals);
  WG_HIP(hipGetLastError());
}

void LaunchFacetKeys(const FacetDev* fac, const unsigned long long* vbase, int F,
                     const double* uvw, const double* scale, size_t n_chan, size_t n_virtual,
                     int atlas_ntu, int atlas_P0, int key_shift, int key_T, uint32_t* keys,
                     uint32_t* vals, hipStream_t s) {
  if (!n_virtual) return;
  FacetKeysKernel<<<Blocks(n_virtual), kBlock, 0, s>>>(fac, vbase, F, uvw, scale, n_chan,
                                                        n_virtual, atlas_ntu, atlas_P0,
                                                        key_shift, key_T, keys, vals);
  WG_HIP(hipGetLastError());
}

void LaunchFacetKeysRenumbered(const FacetDev* fac, const unsigned long long* vrbase, int F,
                               const double* uvw4, const double* scale,
                               size_t n_vrows, size_t n_chan, int atlas_ntu, int atlas_P0,
                               int key_shift, int key_T, uint32_t* keys, hipStream_t s) {
  if (!n_vrows) return;
  WG_CHECK((reinterpret_cast<uintptr_t>(uvw4) & 31) == 0, "padded uvw must be 32-byte aligned");
  FacetKeysRenumberedKernel<<<Blocks(n_vrows * n_chan), kBlock, 0, s>>>(
      fac, vrbase, F, uvw4, scale, n_vrows, n_chan, atlas_ntu, atlas_P0, key_shift, key_T, keys);
  WG_HIP(hipGetLastError());
}

void LaunchFacetRowMap(const FacetDev* fac, const unsigned long long* vrbase, int F,
                       const uint32_t* rperm, size_t n_vrows, uint32_t* rphys, uint32_t* rfac,
                       hipStream_t s) {
  if (!n_vrows) return;
  FacetRowMapKernel<<<Blocks(n_vrows), kBlock, 0, s>>>(fac, vrbase, F, rperm, n_vrows, rphys,
                                                        rfac);
  WG_HIP(hipGetLastError());
}

void LaunchFacetCoords(const FacetDev* fac, const unsigned long long* vrbase, int F,
                       const double* uvw4, const double* scale,
                       size_t n_chan, uint32_t* perm, uint32_t* perm_phys, size_t n_valid,
                       float4* coord, float2* phasor, hipStream_t s, uint32_t* vrow) {
  if (!n_valid) return;
  WG_CHECK((reinterpret_cast<uintptr_t>(uvw4) & 31) == 0, "padded uvw must be 32-byte aligned");
  FacetCoordsKernel<<<Blocks(n_valid), kBlock, 0, s>>>(fac, vrbase, F, uvw4, scale, n_chan,
                                                        perm, perm_phys, n_valid, coord, phasor,
                                                        vrow);
  WG_HIP(hipGetLastError());
}

void LaunchPadUvw(const double* uvw, const uint32_t* rfetch, const uint32_t* rtag, size_t n_rows,
                  double* uvw4, hipStream_t s, const uint32_t* rhi) {
  if (!n_rows) return;
  PadUvwKernel<<<Blocks(4 * n_rows), kBlock, 0, s>>>(uvw, rfetch, rtag, rhi, n_rows, uvw4);
  WG_HIP(hipGetLastError());
}

void LaunchKeysRenumbered(const DevPlan& dp, const double* uvw4, const double* scale,
                          size_t n_rows, size_t n_chan, const float2* vis,
                          uint32_t* keys, hipStream_t s, uint32_t* rowmax_bits) {
  WG_CHECK(!rowmax_bits || vis, "the row maxima need the visibilities");
  if (!n_rows) return;
  if (rowmax_bits) LaunchFillZero(rowmax_bits, n_rows * sizeof(uint32_t), s);
  WG_CHECK((reinterpret_cast<uintptr_t>(uvw4) & 31) == 0, "padded uvw must be 32-byte aligned");
  if (rowmax_bits)
    KeysRenumberedKernel<true><<<Blocks(n_rows * n_chan), kBlock, 0, s>>>(
        dp, uvw4, scale, n_rows, n_chan, vis, keys, rowmax_bits);
  else
    KeysRenumberedKernel<false><<<Blocks(n_rows * n_chan), kBlock, 0, s>>>(
        dp, uvw4, scale, n_rows, n_chan, vis, keys, nullptr);
  WG_HIP(hipGetLastError());
}

void LaunchMaskBits(const uint32_t* in, size_t n, uint32_t mask, uint32_t* out, hipStream_t s) {
  if (!n) return;
  MaskKernel<<<Blocks(n), kBlock, 0, s>>>(in, n, mask, out);
  WG_HIP(hipGetLastError());
}

void LaunchGatherVis(const uint32_t* perm, const float2* phasor, size_t n_valid,
                     const float2* vis, float2* vis_sorted, hipStream_t s) {
  if (!n_valid) return;
  GatherVisKernel<<<Blocks(n_valid), kBlock, 0, s>>>(perm, phasor, n_valid, vis, vis_sorted);
  WG_HIP(hipGetLastError());
}

void LaunchRowMax(const float2* vis, const uint32_t* row_map, size_t n_rows, size_t n_chan,
                  uint32_t* rowmax_bits, hipStream_t s) {
  if (!n_rows || !n_chan) return;
  LaunchFillZero(rowmax_bits, n_rows * sizeof(uint32_t), s);
  const size_t n = n_rows * n_chan;
  if (!row_map && n_chan % 2 == 0 && (reinterpret_cast<uintptr_t>(vis) & 15) == 0) {
    // four 128-element chunks per wave step, about four steps per wave
    const unsigned nb =
        unsigned(std::min<size_t>((n / 2 + 16 * kBlock - 1) / (16 * kBlock), 16384));
    RowMaxPairsKernel<<<std::max(nb, 1u), kBlock, 0, s>>>(
        reinterpret_cast<const float4*>(vis), n_rows, n_chan, rowmax_bits);
  } else {
    // four 64-element chunks per wave step, about four steps per wave
    const unsigned nb = unsigned(std::min<size_t>((n + 16 * kBlock - 1) / (16 * kBlock), 16384));
    RowMaxKernel<<<std::max(nb, 1u), kBlock, 0, s>>>(vis, row_map, n_rows, n_chan, rowmax_bits);
  }
  WG_HIP(hipGetLastError());
}

void LaunchGatherRows(const float* in, const uint32_t* rows, size_t n, float* out,
                      hipStream_t s) {
  if (!n) return;
  GatherRowsKernel<<<Blocks(n), kBlock, 0, s>>>(in, rows, n, out);
  WG_HIP(hipGetLastError());
}

void LaunchItemBounds(const WorkItem* items, size_t n_items, const float2* vis_sorted,
                      const uint32_t* vrow, const float* rowmax, float2* bounds, hipStream_t s) {
  if (!n_items) return;
  WG_CHECK(n_items < (size_t(1) << 31), "too many work items");
  if (vis_sorted) {
    ItemBoundsKernel<true><<<unsigned(n_items), 256, 0, s>>>(items, vis_sorted, nullptr, nullptr,
                                                              bounds);
  } else {
    WG_CHECK(vrow && rowmax, "item bounds need the sorted rows and the row maxima");
    ItemBoundsKernel<false><<<unsigned(n_items), 256, 0, s>>>(items, nullptr, vrow, rowmax,
                                                               bounds);
  }
  WG_HIP(hipGetLastError());
}

void LaunchScatterVis(const uint32_t* perm, const float2* phasor, size_t n_valid,
                      const float2* acc, float2* vis_out, hipStream_t s) {
  if (!n_valid) return;
  ScatterVisKernel<<<Blocks(n_valid), kBlock, 0, s>>>(perm, phasor, n_valid, acc, vis_out);
  WG_HIP(hipGetLastError());
}

void LaunchCorrection(const DevPlan& dp, const double* phi_u, const double* phi_v,
                      const double* quad_a, const double* quad_b, int nquad, float* corr,
                      hipStream_t s) {
  CorrectionKernel<<<Blocks(size_t(dp.nx) * dp.ny), kBlock, 0, s>>>(dp, phi_u, phi_v, quad_a,
                                                                    quad_b, nquad, corr);
  WG_HIP(hipGetLastError());
}

void LaunchPhiTables(const DevPlan& dp, const double* quad_a, const double* quad_b, int nquad,
                     double* phi_u, double* phi_v, hipStream_t s) {
  const int n = dp.nx + dp.ny;
  PhiTablesKernel<<<(n + 255) / 256, 256, 0, s>>>(dp, quad_a, quad_b, nquad, phi_u, phi_v);
  WG_HIP(hipGetLastError());
}

void LaunchAccumulateImage(const float* dirty, const float* corr, double* img, size_t n,
                           hipStream_t s) {
  AccumulateImageKernel<<<Blocks(n), kBlock, 0, s>>>(dirty, corr, img, n);
  WG_HIP(hipGetLastError());
}

void LaunchStridedGather(const uint32_t* in, size_t stride, size_t count, uint32_t* out,
                         hipStream_t s) {
  if (!count) return;
  StridedGatherKernel<<<unsigned((count + 255) / 256), 256, 0, s>>>(in, stride, count, out);
  WG_HIP(hipGetLastError());
}

void LaunchStrideRows(const void* in, size_t stride, size_t n_out, size_t row_words, void* out,
                      hipStream_t s) {
  const size_t n = n_out * row_words;
  if (!n) return;
  StrideRowsKernel<<<Blocks(n), kBlock, 0, s>>>(static_cast<const uint64_t*>(in), stride, n_out,
                                                 row_words, static_cast<uint64_t*>(out));
  WG_HIP(hipGetLastError());
}

void LaunchFillZero(void* p, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  WG_CHECK(bytes % 4 == 0, "fill size must be a multiple of 4");
  FillZeroKernel<<<Blocks(bytes / 4), kBlock, 0, s>>>(static_cast<uint32_t*>(p), bytes / 4);
  WG_HIP(hipGetLastError());
}

void LaunchCorrectCollapse(const float2* vis, const uint32_t* antennas,
                           const uint64_t* time_offsets, const float2* gains, size_t n_rows,
                           size_t n_chan, size_t n_ant, int nparms, int mode, int npol,
                           bool forward, int pol, float2* out, hipStream_t s) {
  const size_t n = n_rows * n_chan;
  if (!n) return;
  CorrectCollapseKernel<<<Blocks(n), kBlock, 0, s>>>(vis, antennas, time_offsets, gains, n_rows,
                                                     n_chan, n_ant, nparms, mode, npol,
                                                     forward ? 1 : 0, pol, out);
  WG_HIP(hipGetLastError());
}

}  // namespace wg
