// Split-precision matrix-core gridding and degridding (W <= 8).
//
// The FP32 matrix instruction v_mfma_f32_16x16x4_f32 issues at the rate of
// the FP32 VALU and never overlaps it, so the FP32 kernels are bound by the
// sum of their MFMA and VALU issue.  v_mfma_f32_16x16x32_f16 does eight
// times the work per issue, in 20 instead of 32 cycles, and overlaps VALU
// work.  These kernels keep FP32 accuracy on it by carrying every operand
// value x as two halves, hi = RN16(x) and lo = RN16(x - hi), and every
// product of two operands as hi hi' + hi lo' + lo hi' (three MFMAs), which
// leaves out lo lo' and the rounding of lo: about 2^-21 of |x x'| while lo
// is a normal half.  Operands are scaled by powers of two to within 2^14
// of the half range, so that holds for every term above ~1e-5 of the
// largest, and smaller ones err by at most 2^-25 of the largest.
// Each hi / lo pair comes from two v_fma_mix{lo,hi}_f16 per product: the
// product and the residual are formed in one fused operation each.
//
// Gridding of a 32-visibility group of one run (same first cell and plane):
//   D_m[(c, l)][(k, t')] += sum_s A[(c, l)][s] B_m[s][(k, t')]
// with A = (Re | Im V_s) kv_s[l] and B_m = ku_s[k] kw_s[2m + t'], one
// 16x16x32 product per plane pair m and part.  Operand layout: lane L holds
// row (or column) L % 16 of A (B) and eight of the 32 slots; the slot order
// inside the instruction is the same for A and B, so lane group L / 16 is
// given slots 8 (L / 16) .. + 7.  D_m: lane L holds rows 4 (L / 16) + r,
// column L % 16, the layout of the FP32 kernel, so the flush is the same.
// Kernel values: lane L evaluates, for slot L % 32, the W / 2 taps k < W / 2
// of u, v and w at +g (L < 32) or at -g (L >= 32, giving taps W - 1 - k:
// phi is even), so the polynomial coefficients are wave-uniform; the values
// go through a per-wave LDS table of 26 rows (ku 0..7, kv 0..7, kw 0..7,
// Re V, Im V) x 32 slots, from which each lane reads its eight slots of the
// rows it needs with 16-byte loads.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.hpp"
#include "engine.hpp"
#include "mfma_common.hpp"

namespace wg {

namespace {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

// hi = RN16(a0 b0) | RN16(a1 b1) << 16, lo = RN16(a0 b0 - hi0) | ... << 16.
__device__ __forceinline__ void SplitProducts(float a0, float b0, float a1, float b1,
                                              uint32_t& hi, uint32_t& lo) {
  uint32_t h, q;
  // pure arithmetic: not volatile, so loads feeding it can be scheduled
  // ahead of it
  asm("v_fma_mixlo_f16 %0, %2, %3, 0\n\t"
      "v_fma_mixhi_f16 %0, %4, %5, 0\n\t"
      "v_fma_mixlo_f16 %1, %2, %3, -%0 op_sel_hi:[0,0,1]\n\t"
      "v_fma_mixhi_f16 %1, %4, %5, -%0 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
      : "=&v"(h), "=&v"(q)
      : "v"(a0), "v"(b0), "v"(a1), "v"(b1));
  hi = h;
  lo = q;
}

// Eight products a[e] b[e] as hi and lo halves (element e in half e % 2 of
// word e / 2).
__device__ __forceinline__ void SplitProducts8(const float4& a0, const float4& a1,
                                               const float4& b0, const float4& b1,
                                               half8_t& hi, half8_t& lo) {
  uint32_t h0, h1, h2, h3, q0, q1, q2, q3;
  SplitProducts(a0.x, b0.x, a0.y, b0.y, h0, q0);
  SplitProducts(a0.z, b0.z, a0.w, b0.w, h1, q1);
  SplitProducts(a1.x, b1.x, a1.y, b1.y, h2, q2);
  SplitProducts(a1.z, b1.z, a1.w, b1.w, h3, q3);
  hi = __builtin_bit_cast(half8_t, u32x4_t{h0, h1, h2, h3});
  lo = __builtin_bit_cast(half8_t, u32x4_t{q0, q1, q2, q3});
}

// The three split products of one group, summed from zero (the small
// cross terms first).  The matrix instructions do not round the exact sum
// of C and their products: each product is aligned to the result's
// exponent, so addends far below a large C lose their low bits (adding
// 32 x 2^-25 to C = 1 returns 1).  Runs are therefore summed across groups
// with FP32 VALU adds, never in the instruction's accumulator.
__device__ __forceinline__ mfma_f4 Mfma3(const half8_t& ah, const half8_t& al, const half8_t& bh,
                                         const half8_t& bl) {
  mfma_f4 d = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, mfma_f4{0.f, 0.f, 0.f, 0.f}, 0, 0,
                                                     0);
  d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, d, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, d, 0, 0, 0);
}

// Both halves of x broadcast: lo gets lane L % 32, hi lane 32 + L % 32, for
// every lane L (one v_permlane32_swap of x with itself).
__device__ __forceinline__ void HalfBroadcast(float x, float& lo, float& hi) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x), false,
                                                  false);
  lo = __uint_as_float(r[0]);
  hi = __uint_as_float(r[1]);
}

// Per-wave kernel-value table: rows x 32 slots, rows 36 floats apart (the
// rows of one 16-byte read start in different LDS banks).
constexpr int kRowKu = 0, kRowKv = 8, kRowKw = 16, kRowV = 24;
constexpr int kSplitRows = 26;
constexpr int kSplitPitch = 36;
constexpr int kSplitTable = kSplitRows * kSplitPitch;  // floats per wave
// scale of kw in the table (B operands up to 2^8); A operands (kv V) are
// scaled through V to at most 2^14
constexpr float kSplitB = 256.f;

// Lane mask of a slot range: the eight slots of lane group g (two per
// 32-bit word of a half8 operand) inside [a, b) keep their half.
__device__ __forceinline__ u32x4_t SlotMask(uint32_t a, uint32_t b, int g) {
  const uint32_t range = (b >= 32 ? ~0u : ((1u << b) - 1u)) & ~((1u << a) - 1u);
  const uint32_t mb = (range >> (8 * g)) & 0xffu;
  u32x4_t m;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const uint32_t two = (mb >> (2 * p)) & 3u;
    m[p] = ((two | (two << 15)) & 0x10001u) * 0xffffu;
  }
  return m;
}

__device__ __forceinline__ half8_t MaskHalves(const half8_t& x, const u32x4_t& m) {
  return __builtin_bit_cast(half8_t, __builtin_bit_cast(u32x4_t, x) & m);
}

template <int W, bool FUSED>
__global__ void __launch_bounds__(64 * kMfmaWaves)
    __attribute__((amdgpu_waves_per_eu(4, 4))) GridKernelSplit(
    DevPlan dp, const WorkItem* __restrict__ items, const uint32_t* __restrict__ item_pos,
    int pos_lo, int pos_hi, const float4* __restrict__ coord,
    const float2* __restrict__ vis, const uint32_t* __restrict__ perm,
    const float2* __restrict__ phasor, float2* __restrict__ scratch, int sub_lo, int sub_n,
    uint32_t* __restrict__ queue, const float2* __restrict__ bounds) {
  static_assert(W <= 8 && W % 2 == 0, "split-precision gridding covers W <= 8");
  constexpr int H = W / 2;  // taps per coordinate and lane; plane pairs
  constexpr int D = EsPolyDegree(W);
  __shared__ uint32_t next_item;
  // [sub_n][E][E] fixed-point re, im; then the per-wave tables (16-byte
  // aligned)
  extern __shared__ __attribute__((aligned(16))) int split_lds[];
  int* sub_re = split_lds;
  const int E = dp.T + W - 1;
  const int ncell = sub_n * E * E;
  int* sub_im = sub_re + ncell;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float* tbl = static_cast<float*>(__builtin_assume_aligned(
      reinterpret_cast<float*>(sub_re + ((2 * ncell + 3) & ~3)) + wave * kSplitTable, 16));
  // rows of taps >= W are never written and stay zero
  for (int i = lane; i < kSplitTable; i += 64) tbl[i] = 0.f;
  // evaluation lane: slot es of the group, +g (eh = 0) or -g (eh = 1)
  const int es = lane & 31, eh = lane >> 5;
  // operand lane: row / column o16, slots 8 og .. 8 og + 7
  const int o16 = lane & 15, og = lane >> 4;
  const int a_l = o16 & 7, a_c = o16 >> 3;  // A row: v tap, component
  const int b_k = o16 & 7, b_t = o16 >> 3;  // B column: u tap, w tap parity
  const float* a_kv = tbl + (kRowKv + a_l) * kSplitPitch + 8 * og;
  const float* a_v = tbl + (kRowV + a_c) * kSplitPitch + 8 * og;
  const float* b_ku = tbl + (kRowKu + b_k) * kSplitPitch + 8 * og;
  const float* b_kw = tbl + (kRowKw + b_t) * kSplitPitch + 8 * og;
  const int f_c = og >> 1;  // D rows 4 og + r: component
  const uint32_t it0 = item_pos[pos_lo], it1 = item_pos[pos_hi];
  for (;;) {
    if (threadIdx.x == 0) next_item = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t it = it0 + next_item;
    if (it >= it1) break;
    const WorkItem wi = items[it];
    for (int i = threadIdx.x; i < ncell; i += 64 * kMfmaWaves) {
      sub_re[i] = 0;
      sub_im[i] = 0;
    }
    __syncthreads();
    const uint32_t n = wi.end - wi.begin;
    // operands: V sA <= 2^14 in magnitude for the item's largest component
    // bound; fixed-point subgrid as in the FP32 kernel, |cell| <=
    // min(n max b_i, sum b_i) <= 2^e
    const float2 bd = bounds[it];
    int ev = 0;
    (void)frexpf(fmaxf(bd.x, 1e-30f), &ev);
    const float sA = ldexpf(1.0f, 14 - ev);
    const float bnd = fminf(float(n) * bd.x, bd.y);
    int e = 0;
    (void)frexpf(fmaxf(bnd, 1e-30f), &e);
    e = max(e, -90);
    const float fx_scale = (bnd < 3.0e38f) ? ldexpf(1.0f, 30 - e) : 0.f;
    // a non-finite bound (NaN or infinite visibilities) writes the subgrid as NaN
    const float fx_inv = (bnd < 3.0e38f) ? ldexpf(1.0f, e - 30) : __builtin_nanf("");
    // D holds sA kSplitB times the sums
    const float fxs = fx_scale / (sA * kSplitB);
    const uint32_t per = (n + kMfmaWaves - 1) / kMfmaWaves;
    const uint32_t w0 = wi.begin + min(n, wave * per);
    const uint32_t w1 = wi.begin + min(n, (wave + 1) * per);
    mfma_f4 acc[H];
#pragma unroll
    for (int m = 0; m < H; ++m) acc[m] = mfma_f4{0.f, 0.f, 0.f, 0.f};
    int cur = -1;

#define WG_FLUSH_SPLIT()                                                         \
  do {                                                                           \
    const int du_ = cur & 255, dv_ = (cur >> 8) & 255;                           \
    const int q0_ = ((cur >> 16) & 0x7fff) - sub_lo;                             \
    int* dst_ = f_c ? sub_im : sub_re;                                           \
    _Pragma("unroll") for (int m = 0; m < H; ++m) {                              \
      const int t_ = 2 * m + b_t;                                                \
      const int q_ = q0_ + t_;                                                   \
      _Pragma("unroll") for (int r = 0; r < 4; ++r) {                            \
        const int l_ = (4 * og + r) & 7;                                         \
        if (t_ < W && l_ < W && b_k < W && q_ >= 0 && q_ < sub_n)                \
          atomicAdd(&dst_[(q_ * E + dv_ + l_) * E + du_ + b_k],                  \
                    __float2int_rn(acc[m][r] * fxs));                            \
        acc[m][r] = 0.f;                                                         \
      }                                                                          \
    }                                                                            \
  } while (0)

    if (w0 < w1) {
      // batch b = [w0 + 64 b, w0 + 64 b + 64): lane L holds visibility L;
      // the next batch is in flight while this one is processed
      float4 pc;
      VisSource<FUSED> src;
      src.vis = vis;
      if constexpr (FUSED) {
        src.perm = perm;
        src.phasor = phasor;
      }
      auto load = [&](uint32_t first) {
        // lanes past the range: their key marks the end and their V is 0
        pc = coord[StageIndex(first, w1, lane)];
        src.Issue(first, w1, lane);
      };
      src.Prime(w0, w1, lane);
      load(w0);
      int prev_last = -1;
      for (uint32_t bb = w0; bb < w1; bb += kStage) {
        const float4 cc = pc;
        const float2 v = src.Value();
        const float vre = v.x * sA, vim = v.y * sA;
        const int key = StageKey(bb, w1, lane, cc);
        const uint64_t sm = RunStarts(key, prev_last, lane);
        prev_last = __builtin_amdgcn_readlane(key, 63);
        if (bb + kStage < w1) load(bb + kStage);
        // both 32-visibility groups of the batch, seen by every lane
        // (evaluation lanes eh = 0 / 1 store the real / imaginary part)
        float bu[2], bv[2], bx[2];
        HalfBroadcast(cc.x, bu[0], bu[1]);
        HalfBroadcast(cc.y, bv[0], bv[1]);
        // the w taps of the lane's own visibility, slot es of group eh:
        // each visibility's w polynomial pairs are evaluated once
        float tw[W];
        {
          const float gw = __builtin_fmaf(cc.z, 2.f, -1.f);
          const float yw = gw * gw;
#pragma unroll
          for (int q = 0; q < H; ++q) {
            float c[D + 1];
#pragma unroll
            for (int d = 0; d <= D; ++d) c[d] = dp.es_poly[q][d];
            float E, O;
            EsPolyEvenOdd<D>(c, yw, E, O);
            tw[q] = __builtin_fmaf(gw, O, E) * kSplitB;
            tw[W - 1 - q] = __builtin_fmaf(-gw, O, E) * kSplitB;
          }
        }
        {
          float r[2], i[2];
          HalfBroadcast(vre, r[0], r[1]);
          HalfBroadcast(vim, i[0], i[1]);
          bx[0] = eh ? i[0] : r[0];
          bx[1] = eh ? i[1] : r[1];
        }
#pragma unroll
        for (int half = 0; half < 2; ++half) {
          const uint32_t gb = bb + 32u * uint32_t(half);
          if (gb >= w1) break;
          const uint32_t glen = min(32u, w1 - gb);
          // kernel values of slot es into the table.  The even and odd
          // parts of tap q's polynomial at y = g^2 give tap q (+g) and tap
          // W - 1 - q (-g) together: lanes eh = 0 / 1 take every tap of u /
          // v; the w taps come from the lanes of this group's visibilities.
          {
            const float g1 = __builtin_fmaf(eh ? bv[half] : bu[half], 2.f, -1.f);
            const float y1 = g1 * g1;
            float* r1 = tbl + (eh ? kRowKv : kRowKu) * kSplitPitch + es;
            float* rw = tbl + kRowKw * kSplitPitch + es;
#pragma unroll
            for (int q = 0; q < H; ++q) {
              float c[D + 1];
#pragma unroll
              for (int d = 0; d <= D; ++d) c[d] = dp.es_poly[q][d];
              float E, O;
              EsPolyEvenOdd<D>(c, y1, E, O);
              r1[q * kSplitPitch] = __builtin_fmaf(g1, O, E);
              r1[(W - 1 - q) * kSplitPitch] = __builtin_fmaf(-g1, O, E);
            }
            if (eh == half) {
#pragma unroll
              for (int k = 0; k < W; ++k) rw[k * kSplitPitch] = tw[k];
            }
          }
          tbl[(kRowV + eh) * kSplitPitch + es] = bx[half];
          WaveLdsSync();
          // A: kv[a_l] (Re | Im V) of the lane's eight slots
          half8_t ah, al;
          SplitProducts8(*reinterpret_cast<const float4*>(a_kv),
                         *reinterpret_cast<const float4*>(a_kv + 4),
                         *reinterpret_cast<const float4*>(a_v),
                         *reinterpret_cast<const float4*>(a_v + 4), ah, al);
          // ku[b_k] of the lane's slots; B_m = ku kw[2 m + b_t] is formed
          // per plane pair next to its products
          const float4 u0 = *reinterpret_cast<const float4*>(b_ku);
          const float4 u1 = *reinterpret_cast<const float4*>(b_ku + 4);
          auto products = [&](const half8_t& xh, const half8_t& xl) {
#pragma unroll
            for (int m = 0; m < H; ++m) {
              const float* kw = b_kw + 2 * m * kSplitPitch;
              half8_t bh, bl;
              SplitProducts8(u0, u1, *reinterpret_cast<const float4*>(kw),
                             *reinterpret_cast<const float4*>(kw + 4), bh, bl);
              acc[m] += Mfma3(xh, xl, bh, bl);
            }
          };
          // runs starting in this group (slots past the range carry V = 0)
          const uint32_t gmask = glen >= 32 ? ~0u : ((1u << glen) - 1u);
          const uint32_t starts = uint32_t(sm >> (32 * half)) & gmask;
          if (starts == 0u) {
            products(ah, al);
          } else {
            uint32_t a = 0;
            while (a < 32u) {
              const uint32_t above = a < 31u ? (starts & (~0u << (a + 1u))) : 0u;
              const uint32_t b = above ? uint32_t(__builtin_ctz(above)) : 32u;
              if ((starts >> a) & 1u) {
                if (cur >= 0) WG_FLUSH_SPLIT();
                cur = __builtin_amdgcn_readlane(key, int(32u * uint32_t(half) + a));
              }
              const u32x4_t mk = SlotMask(a, b, og);
              products(MaskHalves(ah, mk), MaskHalves(al, mk));
              a = b;
            }
          }
          WaveLdsSync();
        }
      }
      if (cur >= 0) WG_FLUSH_SPLIT();
    }
#undef WG_FLUSH_SPLIT
    __syncthreads();
    float2* slot = scratch + size_t(it - it0) * size_t(ncell);
    for (int i = threadIdx.x; i < ncell; i += 64 * kMfmaWaves)
      slot[i] = make_float2(float(sub_re[i]) * fx_inv, float(sub_im[i]) * fx_inv);
    __syncthreads();
  }
}

// Degridding of a 16-visibility column group of one run:
//   D[(t, c)][j] = sum_{l, k} G[q0 + t][dv + l][du + k].c kv_j[l] ku_j[k]
// over the 64 pairs (l, k) as two 16x16x32 products: K-half h gives lane
// group L / 16 the pairs l = 4 h + L / 16, k = 0..7.  A = the run's box
// block (row L % 16 = (t, c), split once per run, scaled by a power of two
// to within 2^14 of the half range from the item's largest component),
// B = kv ku of the column's visibility (split per group).  Then V_j =
// sum_t kw_j[t] (D[(t, re)][j] + i D[(t, im)][j]): lane L holds rows
// 4 (L / 16) + r, i.e. planes 2 (L / 16) and 2 (L / 16) + 1, and the four
// lanes of column j add their parts.  A run boundary inside a group gives
// one product per run, each kept for its own columns.  Kernel values as in
// the gridder, through a per-wave table [32 slots][28] (ku 0..7, kv 8..15,
// kw 16..23 of each slot).
constexpr int kDgPitch = 28;
constexpr int kDgTable = 32 * kDgPitch;  // floats per wave

template <int W, bool FUSED>
__global__ void __launch_bounds__(64 * kMfmaWaves)
    __attribute__((amdgpu_waves_per_eu(4, 4))) DegridKernelSplit(
    DevPlan dp, const WorkItem* __restrict__ items, const uint32_t* __restrict__ item_pos,
    int pos_lo, int pos_hi, const float4* __restrict__ coord, const float2* __restrict__ box,
    float2* __restrict__ acc, const uint32_t* __restrict__ perm, const float2* __restrict__ phasor,
    int sub_lo, int sub_n, int slab_lo, uint32_t* __restrict__ queue) {
  static_assert(W <= 8 && W % 2 == 0, "split-precision degridding covers W <= 8");
  constexpr int H = W / 2;
  constexpr int D = EsPolyDegree(W);
  __shared__ uint32_t next_item;
  __shared__ uint32_t item_max;  // largest |component| of the item's block, as float bits
  extern __shared__ __attribute__((aligned(16))) float2 dsplit_lds[];
  float2* sub = dsplit_lds;  // [sub_n][E][E] box block of the item's tile
  const int E = dp.T + W - 1;
  const int ncell = sub_n * E * E;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float* tbl = static_cast<float*>(__builtin_assume_aligned(
      reinterpret_cast<float*>(dsplit_lds + ((ncell + 1) & ~1)) + wave * kDgTable, 16));
  for (int i = lane; i < kDgTable; i += 64) tbl[i] = 0.f;
  const int es = lane & 31, eh = lane >> 5;
  const int jc = lane & 15, og = lane >> 4;
  const int a_t = jc >> 1, a_c = jc & 1;  // A row (plane, component)
  // w taps: lane half eh evaluates the even / odd pairs j + eh HW (taps q
  // and W - 1 - q of its slot), so each pair is evaluated once per slot;
  // their coefficients are held per lane
  constexpr int HW = (H + 1) / 2;
  float cw[HW][D + 1];
#pragma unroll
  for (int j = 0; j < HW; ++j)
#pragma unroll
    for (int d = 0; d <= D; ++d)
      cw[j][d] = eh ? (j + HW < H ? dp.es_poly[j + HW < H ? j + HW : j][d] : 0.f)
                    : dp.es_poly[j][d];
  const uint32_t it0 = item_pos[pos_lo], it1 = item_pos[pos_hi];
  for (;;) {
    if (threadIdx.x == 0) {
      next_item = atomicAdd(queue, 1u);
      item_max = 0u;
    }
    __syncthreads();
    const uint32_t it = it0 + next_item;
    if (it >= it1) break;
    const WorkItem wi = items[it];
    const int tu = int(wi.tile % uint32_t(dp.ntu)), tv = int(wi.tile / uint32_t(dp.ntu));
    float bmax = 0.f;
    for (int i = threadIdx.x; i < ncell; i += 64 * kMfmaWaves) {
      const int p = i / (E * E), r = (i / E) % E, c = i % E;
      const float2 v = box[(size_t(sub_lo + p - slab_lo) * dp.BV + size_t(tv * dp.T + r)) * dp.BU +
                           size_t(tu * dp.T + c)];
      sub[i] = v;
      bmax = fmaxf(bmax, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
    for (int m = 32; m >= 1; m >>= 1) bmax = fmaxf(bmax, __shfl_xor(bmax, m));
    if (lane == 0) atomicMax(&item_max, __float_as_uint(bmax));
    __syncthreads();
    int eg = 0;
    (void)frexpf(fmaxf(__uint_as_float(item_max), 1e-30f), &eg);
    const float sG = ldexpf(1.0f, 14 - eg);           // |G sG| <= 2^14
    const float inv_s = ldexpf(1.0f, eg - 14) * (1.0f / kSplitB);
    const uint32_t n = wi.end - wi.begin;
    const uint32_t per = (n + kMfmaWaves - 1) / kMfmaWaves;
    const uint32_t w0 = wi.begin + min(n, wave * per);
    const uint32_t w1 = wi.begin + min(n, (wave + 1) * per);
    if (w0 < w1) {
      half8_t ah[2], al[2];
      // the run's box block as A operands.  Every lane reads its W cells
      // from a valid address (row 0 of the block where its row is outside
      // the block or the kernel), so the reads issue together; the values
      // of such rows are then zero.
      auto load_a = [&](int key) {
        const int du = key & 255, dv = (key >> 8) & 255;
        const int q = ((key >> 16) & 0x7fff) - sub_lo + a_t;
        const bool row_ok = a_t < W && q >= 0 && q < sub_n;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int l = 4 * h + og;
          const bool ok = row_ok && l < W;
          const float2* src = sub + ((ok ? q * E + dv + l : 0) * E + (ok ? du : 0));
          float2 g[W];
#pragma unroll
          for (int e = 0; e < W; ++e) g[e] = src[e];
          float x[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) x[e] = (e < W && ok) ? (a_c ? g[e < W ? e : 0].y
                                                                   : g[e < W ? e : 0].x)
                                                          : 0.f;
          const float4 s4 = make_float4(sG, sG, sG, sG);
          SplitProducts8(make_float4(x[0], x[1], x[2], x[3]), make_float4(x[4], x[5], x[6], x[7]),
                         s4, s4, ah[h], al[h]);
        }
      };
      // lanes past the range: their key marks the end and their outputs
      // are not stored
      float4 pc = coord[StageIndex(w0, w1, lane)];
      int prev_last = -1;
      for (uint32_t bb = w0; bb < w1; bb += kStage) {
        const float4 cc = pc;
        const int key = StageKey(bb, w1, lane, cc);
        const uint64_t sm = RunStarts(key, prev_last, lane);
        prev_last = __builtin_amdgcn_readlane(key, 63);
        if (bb + kStage < w1) pc = coord[StageIndex(bb + kStage, w1, lane)];
        float bu[2], bv[2], bw[2];
        HalfBroadcast(cc.x, bu[0], bu[1]);
        HalfBroadcast(cc.y, bv[0], bv[1]);
        HalfBroadcast(cc.z, bw[0], bw[1]);
#pragma unroll
        for (int half = 0; half < 2; ++half) {
          const uint32_t gb = bb + 32u * uint32_t(half);
          if (gb >= w1) break;
          const uint32_t glen = min(32u, w1 - gb);
          // outputs of this group: lanes og = 0 / 1 write column group
          // cg = og (visibility 16 og + jc of the group)
          const uint32_t vout = gb + uint32_t(16 * (og & 1) + jc);
          const bool store = og < 2 && vout < w1;
          uint32_t pw = 0u;
          float2 ph = make_float2(0.f, 0.f), av = make_float2(0.f, 0.f);
          if (store) {
            if constexpr (FUSED) {
              pw = perm[vout];
              ph = phasor[vout];
            } else {
              av = acc[vout];
            }
          }
          float sre[2], sim[2];
          // kernel values of slot es into the table: lanes eh = 0 / 1 take
          // every tap of u / v from the even and odd parts of the tap
          // polynomials (as in the gridder), and each half evaluates its
          // share of the w tap pairs
          {
            const float g1 = __builtin_fmaf(eh ? bv[half] : bu[half], 2.f, -1.f);
            const float gw = __builtin_fmaf(bw[half], 2.f, -1.f);
            const float y1 = g1 * g1, yw = gw * gw;
            float* row = tbl + es * kDgPitch;
            float* r1 = row + (eh ? 8 : 0);
#pragma unroll
            for (int q = 0; q < H; ++q) {
              float c[D + 1];
#pragma unroll
              for (int d = 0; d <= D; ++d) c[d] = dp.es_poly[q][d];
              float E, O;
              EsPolyEvenOdd<D>(c, y1, E, O);
              r1[q] = __builtin_fmaf(g1, O, E);
              r1[W - 1 - q] = __builtin_fmaf(-g1, O, E);
            }
#pragma unroll
            for (int j = 0; j < HW; ++j) {
              const int q = j + (eh ? HW : 0);
              float E, O;
              EsPolyEvenOdd<D>(cw[j], yw, E, O);
              if (q < H) {
                row[16 + q] = __builtin_fmaf(gw, O, E);
                row[16 + W - 1 - q] = __builtin_fmaf(-gw, O, E);
              }
            }
          }
          WaveLdsSync();
          const uint32_t gmask = glen >= 32 ? ~0u : ((1u << glen) - 1u);
          const uint32_t starts = uint32_t(sm >> (32 * half)) & gmask;
#pragma unroll
          for (int cg = 0; cg < 2; ++cg) {
            // B of column jc: kv[4 h + og] ku[0..7] of slot 16 cg + jc
            const float* ent = tbl + (16 * cg + jc) * kDgPitch;
            const float4 u0 = *reinterpret_cast<const float4*>(ent);
            const float4 u1 = *reinterpret_cast<const float4*>(ent + 4);
            const float kv0 = ent[8 + og] * kSplitB, kv1 = ent[12 + og] * kSplitB;
            const float2 kw = *reinterpret_cast<const float2*>(ent + 16 + 2 * og);
            half8_t bh[2], bl[2];
            SplitProducts8(make_float4(kv0, kv0, kv0, kv0), make_float4(kv0, kv0, kv0, kv0), u0,
                           u1, bh[0], bl[0]);
            SplitProducts8(make_float4(kv1, kv1, kv1, kv1), make_float4(kv1, kv1, kv1, kv1), u0,
                           u1, bh[1], bl[1]);
            mfma_f4 dsel = mfma_f4{0.f, 0.f, 0.f, 0.f};
            uint32_t a = 16u * uint32_t(cg);
            const uint32_t end = a + 16u;
            while (a < end) {
              if ((starts >> a) & 1u)
                load_a(__builtin_amdgcn_readlane(key, int(32u * uint32_t(half) + a)));
              const uint32_t after_a = a >= 31u ? 0u : (~0u << (a + 1u));      // bits > a
              const uint32_t before_end = end >= 32u ? ~0u : ((1u << end) - 1u);  // bits < end
              const uint32_t above = starts & after_a & before_end;
              const uint32_t b = above ? uint32_t(__builtin_ctz(above)) : end;
              // cross terms first, from zero (see Mfma3)
              mfma_f4 d = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                  al[0], bh[0], mfma_f4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
              d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0], bl[0], d, 0, 0, 0);
              d = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[1], bh[1], d, 0, 0, 0);
              d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1], bl[1], d, 0, 0, 0);
              d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[0], bh[0], d, 0, 0, 0);
              d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[1], bh[1], d, 0, 0, 0);
              const uint32_t col = 16u * uint32_t(cg) + uint32_t(jc);
              if (col >= a && col < b) dsel = d;
              a = b;
            }
            // rows 4 og + r = (t, c) = (2 og + r / 2, r % 2)
            float re = kw.x * dsel[0] + kw.y * dsel[2];
            float im = kw.x * dsel[1] + kw.y * dsel[3];
            re += __shfl_xor(re, 16);
            im += __shfl_xor(im, 16);
            re += __shfl_xor(re, 32);
            im += __shfl_xor(im, 32);
            sre[cg] = re;
            sim[cg] = im;
          }
          if (store) {
            const float re = (og ? sre[1] : sre[0]) * inv_s;
            const float im = (og ? sim[1] : sim[0]) * inv_s;
            if constexpr (FUSED) {
              float2 V = make_float2(re * ph.x + im * ph.y, im * ph.x - re * ph.y);
              if (pw & kPermFlip) V.y = -V.y;
              acc[pw & ~kPermFlip] = V;
            } else {
              acc[vout] = make_float2(av.x + re, av.y + im);
            }
          }
          WaveLdsSync();
        }
      }
    }
    __syncthreads();
  }
}

template <typename K>
void AllowSplitLds(K kernel, size_t bytes) {
  AllowDynamicLds(reinterpret_cast<const void*>(kernel), bytes);
}

}  // namespace

size_t SplitGridLdsBytes(const DevPlan& dp, int sub_n) {
  const size_t E = size_t(dp.T + dp.W - 1);
  const size_t sub = size_t(sub_n) * E * E * 2 * sizeof(int);
  return ((sub + 15) & ~size_t(15)) + size_t(kMfmaWaves) * kSplitTable * sizeof(float);
}

void LaunchGridSplit(const DevPlan& dp, const VisSet& vs, const float2* vis_sorted,
                     float2* scratch, uint32_t* queue, const float2* bounds, int sub,
                     int sub_lo, int sub_n, hipStream_t s, const float2* vis_raw,
                     const uint32_t* gather_perm) {
  WG_CHECK(dp.W <= 8 && dp.W % 2 == 0, "split-precision gridding covers even W <= 8");
  WG_CHECK(sub_n >= 1 && sub_n <= kMaxSubSlab, "sub-slab depth out of range");
  const size_t ntiles = size_t(dp.ntu) * dp.ntv;
  const int pos_lo = int(sub * ntiles), pos_hi = int((sub + 1) * ntiles);
  const size_t total = SplitGridLdsBytes(dp, sub_n);
  WG_CHECK(total <= 160 * 1024, "gridding LDS exceeds 160 KiB");
  auto go = [&](auto kernel, const float2* v, const uint32_t* pm) {
    AllowSplitLds(kernel, total);
    const int per_cu =
        ResidentBlocksPerCU(reinterpret_cast<const void*>(kernel), 64 * kMfmaWaves, total);
    const int blocks = std::max(1, per_cu) * DeviceCUs();
    LaunchFillZero(queue, sizeof(uint32_t), s);
    kernel<<<blocks, 64 * kMfmaWaves, total, s>>>(dp, vs.items, vs.item_pos, pos_lo, pos_hi,
                                                  vs.coord, v, pm, vs.phasor, scratch, sub_lo,
                                                  sub_n, queue, bounds);
  };
  switch (dp.W) {
    case 4:
      if (vis_raw) go(GridKernelSplit<4, true>, vis_raw, gather_perm);
      else go(GridKernelSplit<4, false>, vis_sorted, nullptr);
      break;
    case 6:
      if (vis_raw) go(GridKernelSplit<6, true>, vis_raw, gather_perm);
      else go(GridKernelSplit<6, false>, vis_sorted, nullptr);
      break;
    default:
      if (vis_raw) go(GridKernelSplit<8, true>, vis_raw, gather_perm);
      else go(GridKernelSplit<8, false>, vis_sorted, nullptr);
  }
  WG_HIP(hipGetLastError());
}

size_t SplitDegridLdsBytes(const DevPlan& dp, int sub_n) {
  const size_t E = size_t(dp.T + dp.W - 1);
  const size_t sub = size_t(sub_n) * E * E * sizeof(float2);
  return ((sub + 15) & ~size_t(15)) + size_t(kMfmaWaves) * kDgTable * sizeof(float);
}

void LaunchDegridSplit(const DevPlan& dp, const VisSet& vs, const float2* box, float2* acc,
                       uint32_t* queue, int sub, int sub_lo, int sub_n, int slab_lo,
                       hipStream_t s, const uint32_t* scatter_perm) {
  WG_CHECK(dp.W <= 8 && dp.W % 2 == 0, "split-precision degridding covers even W <= 8");
  WG_CHECK(sub_n >= 1 && sub_n <= kMaxSubSlab, "sub-slab depth out of range");
  const size_t ntiles = size_t(dp.ntu) * dp.ntv;
  const size_t total = SplitDegridLdsBytes(dp, sub_n);
  WG_CHECK(total <= 160 * 1024, "degridding LDS exceeds 160 KiB");
  auto go = [&](auto kernel) {
    AllowSplitLds(kernel, total);
    const int per_cu =
        ResidentBlocksPerCU(reinterpret_cast<const void*>(kernel), 64 * kMfmaWaves, total);
    LaunchFillZero(queue, sizeof(uint32_t), s);
    kernel<<<std::max(1, per_cu) * DeviceCUs(), 64 * kMfmaWaves, total, s>>>(
        dp, vs.items, vs.item_pos, int(sub * ntiles), int((sub + 1) * ntiles), vs.coord, box, acc,
        scatter_perm, vs.phasor, sub_lo, sub_n, slab_lo, queue);
  };
  switch (dp.W) {
    case 4:
      if (scatter_perm) go(DegridKernelSplit<4, true>);
      else go(DegridKernelSplit<4, false>);
      break;
    case 6:
      if (scatter_perm) go(DegridKernelSplit<6, true>);
      else go(DegridKernelSplit<6, false>);
      break;
    default:
      if (scatter_perm) go(DegridKernelSplit<8, true>);
      else go(DegridKernelSplit<8, false>);
  }
  WG_HIP(hipGetLastError());
}

}  // namespace wg
