// ES-kernel gridding and degridding on LDS-staged tile subgrids.
//
// A work item is one uv tile (T x T first-cells) and a range of its sorted
// visibilities that touch the planes [sub_lo, sub_lo + sub_n) of one
// gridding sub-slab.  Its footprint in the grid is E x E cells per plane
// (E = T + W - 1).  Gridding accumulates the visibilities' W x W x W kernel
// footprints into the LDS subgrid and stores it, with plain coalesced
// stores, into the item's own scratch slot; a gather pass then forms every
// box cell as the compensated sum of the slots of the (at most 4) tiles
// covering it (neighbouring tiles overlap by W - 1 cells), in a fixed order.
// No global atomics, and the box is deterministic.  Degridding stages the
// box subgrid from HBM into LDS and gathers per visibility in registers.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.hpp"
#include "engine.hpp"
#include "mfma_common.hpp"

namespace wg {

namespace {

constexpr int kGridBlock = 256;
// Visibilities per work item: large, so that writing (and gathering) the
// E x E x Sg subgrid of an item is amortised over many visibilities, and
// small enough that the dense uv core spreads over many workgroups.
constexpr uint32_t kChunk = 65536;

// Wave-cooperative gridding.  A wave walks its share of the work item one
// visibility at a time; its lanes cover the W x W kernel footprint (lane ->
// column k = lane % W, rows l = lane / W + r * (64 / W)) and accumulate, in
// registers, the footprint cells of the W planes p0 .. p0 + W - 1 the
// visibility touches (run-relative taps, so the work per visibility does
// not grow with the sub-slab depth).  Visibilities are sorted by exact
// first cell inside each (tile, first plane) bucket, so long runs share one
// footprint and one p0; the registers go to the LDS subgrid with
// collision-free LDS atomics only when the run changes.  With the dense uv
// cores of real arrays, where thousands of visibilities share a cell, this
// replaces ~2 W^3 contended atomics per visibility by 3 kernel evaluations
// and 2 R W FMAs per lane.
template <int W>
__global__ void __launch_bounds__(kGridBlock) GridKernel(
    DevPlan dp, const WorkItem* __restrict__ items, const uint32_t* __restrict__ item_pos,
    int pos_lo, int pos_hi, const float4* __restrict__ coord,
    const float2* __restrict__ vis, float2* __restrict__ scratch, int sub_lo, int sub_n) {
  constexpr int RPP = 64 / W;                 // footprint rows per pass
  constexpr int R = (W + RPP - 1) / RPP;      // row passes per lane
  extern __shared__ float sub_re[];           // [sub_n][E][E] real, then imaginary
  const int E = dp.T + W - 1;
  const int ncell = sub_n * E * E;
  float* sub_im = sub_re + ncell;
  // the wave index is uniform within a wave; readfirstlane tells the
  // compiler so, which keeps the whole visibility loop under scalar control
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int kcol = lane % W, lrow = lane / W;
  // the lane evaluates tap kcol of u, v and w; lane t (< W, row 0) holds
  // the v and w weights of tap t (plane p0 + t)
  constexpr int D = EsPolyDegree(W);
  float ck[D + 1];
  EsPolyLoad<D>(dp, kcol, ck);
  const uint32_t it0 = item_pos[pos_lo], it1 = item_pos[pos_hi];
  for (uint32_t it = it0 + blockIdx.x; it < it1; it += gridDim.x) {
    const WorkItem wi = items[it];
    for (int i = threadIdx.x; i < ncell; i += kGridBlock) {
      sub_re[i] = 0.f;
      sub_im[i] = 0.f;
    }
    __syncthreads();
    const uint32_t n = wi.end - wi.begin;
    const uint32_t per = (n + kGridBlock / 64 - 1) / (kGridBlock / 64);
    const uint32_t w0 = wi.begin + min(n, wave * per);
    const uint32_t w1 = wi.begin + min(n, (wave + 1) * per);
    float are[R][W], aim[R][W];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int t = 0; t < W; ++t) are[r][t] = aim[r][t] = 0.f;
    // packed (du, dv, p0) of the current run, wave-uniform (SGPR), so run
    // changes are scalar branches
    int cur = -1;
    const bool lane_ok = lrow < RPP;

#define WG_FLUSH_RUN()                                                  \
  do {                                                                  \
    const int du_ = cur & 255, dv_ = (cur >> 8) & 255;                  \
    const int q0_ = ((cur >> 16) & 0x7fff) - sub_lo;                    \
    _Pragma("unroll") for (int t = 0; t < W; ++t) {                     \
      const int q = q0_ + t;                                            \
      if (q >= 0 && q < sub_n) {                                        \
        _Pragma("unroll") for (int r = 0; r < R; ++r) {                 \
          const int l = lrow + r * RPP;                                 \
          if (lane_ok && l < W) {                                       \
            const int idx = (q * E + dv_ + l) * E + du_ + kcol;         \
            atomicAdd(&sub_re[idx], are[r][t]);                         \
            atomicAdd(&sub_im[idx], aim[r][t]);                         \
          }                                                             \
        }                                                               \
      }                                                                 \
    }                                                                   \
    _Pragma("unroll") for (int r = 0; r < R; ++r)                       \
      _Pragma("unroll") for (int t = 0; t < W; ++t) are[r][t] = aim[r][t] = 0.f; \
  } while (0)

    for (uint32_t base = w0; base < w1; base += 64) {
      const uint32_t idx = base + lane;
      float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
      float2 V = make_float2(0.f, 0.f);
      if (idx < w1) {
        c = coord[idx];
        V = vis[idx];
      }
      const int cnt = int(min(64u, w1 - base));
      for (int j = 0; j < cnt; ++j) {
        const float fu = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c.x), j));
        const float fv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c.y), j));
        const float fw = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c.z), j));
        const int pk = __builtin_amdgcn_readfirstlane(
            __builtin_amdgcn_readlane(__float_as_int(c.w), j) & 0x7fffffff);
        const float vr = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(V.x), j));
        const float vi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(V.y), j));
        if (pk != cur) {
          if (cur >= 0) WG_FLUSH_RUN();
          cur = pk;
        }
        // per-tap w weights times the visibility, one tap per lane
        const float kwl = EsPoly<D>(ck, fw);
        const float wrl = vr * kwl, wil = vi * kwl;
        const float ku = EsPoly<D>(ck, fu);
        const float kvc = EsPoly<D>(ck, fv);
        float kv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int l = lrow + r * RPP;
          // lane l (< W) evaluated v tap l
          const float kvl = __shfl(kvc, l < W ? l : 0);
          kv[r] = (l < W) ? ku * kvl : 0.f;
        }
#pragma unroll
        for (int t = 0; t < W; ++t) {
          const float wr = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wrl), t));
          const float wi_ = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wil), t));
#pragma unroll
          for (int r = 0; r < R; ++r) {
            are[r][t] += wr * kv[r];
            aim[r][t] += wi_ * kv[r];
          }
        }
      }
    }
    if (cur >= 0) WG_FLUSH_RUN();
#undef WG_FLUSH_RUN
    __syncthreads();
    // the item's partial subgrid goes to its own scratch slot with plain
    // coalesced stores; GatherBoxKernel sums the slots into the box
    float2* slot = scratch + size_t(it - it0) * size_t(ncell);
    for (int i = threadIdx.x; i < ncell; i += kGridBlock)
      slot[i] = make_float2(sub_re[i], sub_im[i]);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Matrix-core gridding for W <= 8.
//
// Inside a run (visibilities with the same first cell and first plane) the
// contribution to the W x W x W footprint is a sum of rank-1 terms,
//   G[t][l][k] = sum_j V_j kw_j[t] kv_j[l] ku_j[k],
// i.e. for a group of 4 visibilities and each plane pair m (t = 2m, 2m+1)
// one FP32 MFMA 16x16x4:  D_m[(l,c)][(k,t')] += A[(l,c)][j] B_m[j][(k,t')]
// with A = kv_j[l] (Re V_j | Im V_j) and B_m = ku_j[k] kw_j[2m + t'].
// Operand layout of v_mfma_f32_16x16x4f32: lane L holds A[L % 16][L / 16],
// B[L / 16][L % 16] and D[4 (L / 16) + r][L % 16], r = 0..3.  So in a group
// lane L works on visibility j = L / 16 and on tap L % 8 of u and v; each
// lane evaluates two kernel values per group (ku or kv, then one kw) and
// exchanges the rest with DPP: row_ror:8 swaps the u/v halves of a row and
// a quad broadcast delivers kw[2m + t'] (lane q of each quad evaluates
// tap 2q + t').  Taps >= W are zero, so W < 8 uses the same layout.
// Runs end at group granularity: a group only takes the leading visibilities
// that share the current run; the run's 4 D blocks go to the LDS subgrid
// with LDS atomics when the run changes.  Coordinates are staged per wave
// through a 2 x 64-visibility LDS ring, loaded one batch ahead, and each
// staged batch carries a 64-bit mask of the visibilities that start a run.

// Stores the lane's staged visibility at ring position `slot` in both views.
__device__ __forceinline__ void StoreStage(float4* ring, int slot, const float4& c,
                                           const float2& v) {
  ring[slot] = make_float4(c.x, c.z, v.x, 0.f);
  ring[2 * kStage + slot] = make_float4(c.y, c.z, v.y, 0.f);
}

// The three staged values one lane uses (16-byte aligned, one 96-bit read).
struct alignas(16) Stage3 {
  float x, y, z;
};
__device__ __forceinline__ Stage3 ReadStage(const float4* view, int slot) {
  return *reinterpret_cast<const Stage3*>(view + slot);
}

// Lane exchanges of the gridding step.  The step runs with all 64 lanes
// active, and every exchange must be executed by all of them: a DPP move
// reads its source lane whatever that lane is doing, and a source lane
// masked off by EXEC reads as invalid.  A per-lane select between an
// exchanged and an own value can be turned by the compiler into a masked
// branch around the move, so the lane choice is made by the bank mask of
// update_dpp instead (lanes of banks outside the mask keep their own
// value).  Every source lane read is a valid lane; bound_ctrl is set
// because it lets the compiler fold a quad broadcast into the multiply that
// consumes it.
__device__ __forceinline__ float DppQuadBcast(float v, int m) {
  // quad_perm [m, m, m, m]
  switch (m) {
    case 0: return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x00, 0xf, 0xf, true));
    case 1: return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x55, 0xf, 0xf, true));
    case 2: return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xaa, 0xf, 0xf, true));
    default: return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xff, 0xf, 0xf, true));
  }
}

// The lanes of the banks (quads of a 16-lane row) in BANKS take the value
// moved by CTRL, the others keep `own`.  CTRL 0x128 rotates a row by 8,
// 0x114 shifts it right by 4 (lane i reads i - 4), 0x104 left by 4 (i + 4).
template <int CTRL, int BANKS>
__device__ __forceinline__ float DppMerge(float own) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(own), __float_as_int(own),
                                                    CTRL, 0xf, BANKS, true));
}

// FUSED: vis is the caller's [row][chan] buffer, read through perm and
// rotated by phasor (no gathered copy); otherwise vis is sorted.
template <int W, bool FUSED>
// LDS (subgrid + rings) allows two workgroups, i.e. four waves per SIMD: let
// the register allocator use the registers that occupancy leaves free
// instead of serialising the staged loads.
__global__ void __launch_bounds__(64 * kMfmaWaves)
    __attribute__((amdgpu_waves_per_eu(4, 4))) GridKernelMfma(
    DevPlan dp, const WorkItem* __restrict__ items, const uint32_t* __restrict__ item_pos,
    int pos_lo, int pos_hi, const float4* __restrict__ coord,
    const float2* __restrict__ vis, const uint32_t* __restrict__ perm,
    const float2* __restrict__ phasor, float2* __restrict__ scratch, int sub_lo, int sub_n,
    uint32_t* __restrict__ queue, const float2* __restrict__ bounds) {
  static_assert(W <= 8 && W % 2 == 0, "matrix-core gridding covers W <= 8");
  __shared__ uint32_t next_item;
  // [sub_n][E][E] fixed-point re, im; then the staging rings.  The dynamic
  // area follows next_item, so its alignment is stated explicitly.
  extern __shared__ __attribute__((aligned(16))) int grid_fx_lds[];
  int* sub_re = grid_fx_lds;
  const int E = dp.T + W - 1;
  const int ncell = sub_n * E * E;
  int* sub_im = sub_re + ncell;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // staging rings after the subgrid, 16-byte aligned
  // per wave two views of 2 x 64 staged visibilities: view 0 = (fu, fw, Re V),
  // view 1 = (fv, fw, Im V); lane L reads view (L % 16 >= 8) only
  float4* ring = reinterpret_cast<float4*>(sub_re + ((2 * ncell + 3) & ~3)) +
                 wave * 2 * (2 * kStage);
  const int j = lane >> 4;          // visibility of the group
  const int r16 = lane & 15;
  const int tap = r16 & 7;          // u tap (B columns) and v tap (A rows)
  const bool hi = r16 >= 8;         // A: imaginary part; first evaluation: v
  const int tsig = 2 * (r16 & 3) + (r16 >> 3);  // w tap this lane evaluates
  const bool tap_ok = tap < W, w_ok = tsig < W;
  const bool q_hi = (r16 & 4) != 0;  // quad 1 or 3 of the 16-lane row
  // kernel polynomials of the lane's u / v tap and of its w tap
  constexpr int D = EsPolyDegree(W);
  float cu[D + 1], cw[D + 1];
  EsPolyLoad<D>(dp, tap, cu);
  EsPolyLoad<D>(dp, tsig, cw);
  // flush coordinates of the lane's D entries: D_m[4 (L/16) + r][r16]
  const int f_c = (lane >> 4) >> 1;                 // row i = 4 (L/16) + r -> c = i / 8
  const int f_k = r16 & 7, f_tp = r16 >> 3;          // column -> k, t'
  const uint32_t it0 = item_pos[pos_lo], it1 = item_pos[pos_hi];
  // persistent workgroups take work items from a queue, in item order
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
    // Fixed-point subgrid: every cell component is a sum of V components x
    // kernel weights (<= 1) over the item's n visibilities, so |cell| <=
    // B = min(n max b_i, sum b_i) <= 2^e for bounds b_i on the components
    // (the item's bounds), and values scaled by 2^(30 - e) fit an int32
    // with a factor 2 of headroom for the rounding of the bound's sum and
    // of each flushed term.  The scale follows the item's own visibilities,
    // so an item of faint ones (a densely sampled region under uniform
    // weighting) is not quantised on the scale of the brightest.  Integer
    // LDS atomics are fast and exact, so the subgrid is also independent of
    // the order of the flushes.
    const float2 bd = bounds[it];
    const float bnd = fminf(float(n) * bd.x, bd.y);
    int e = 0;
    (void)frexpf(fmaxf(bnd, 1e-30f), &e);
    e = max(e, -90);  // all-zero items: keep the scale finite
    const float fx_scale = (bnd < 3.0e38f) ? ldexpf(1.0f, 30 - e) : 0.f;
    // a non-finite bound (NaN or infinite visibilities) writes the subgrid as NaN
    const float fx_inv = (bnd < 3.0e38f) ? ldexpf(1.0f, e - 30) : __builtin_nanf("");
    const uint32_t per = (n + kMfmaWaves - 1) / kMfmaWaves;
    const uint32_t w0 = wi.begin + min(n, wave * per);
    const uint32_t w1 = wi.begin + min(n, (wave + 1) * per);
    mfma_f4 acc[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) acc[m] = mfma_f4{0.f, 0.f, 0.f, 0.f};
    int cur = -1;

#define WG_FLUSH_MFMA()                                                          \
  do {                                                                           \
    const int du_ = cur & 255, dv_ = (cur >> 8) & 255;                           \
    const int q0_ = ((cur >> 16) & 0x7fff) - sub_lo;                             \
    int* dst_ = f_c ? sub_im : sub_re;                                           \
    _Pragma("unroll") for (int m = 0; m < 4; ++m) {                              \
      const int t_ = 2 * m + f_tp;                                               \
      const int q_ = q0_ + t_;                                                   \
      _Pragma("unroll") for (int r = 0; r < 4; ++r) {                            \
        const int l_ = (4 * (lane >> 4) + r) & 7;                                \
        if (t_ < W && l_ < W && f_k < W && q_ >= 0 && q_ < sub_n)                \
          atomicAdd(&dst_[(q_ * E + dv_ + l_) * E + du_ + f_k],                  \
                    __float2int_rn(acc[m][r] * fx_scale));                       \
        acc[m][r] = 0.f;                                                         \
      }                                                                          \
    }                                                                            \
  } while (0)

    if (w0 < w1) {
      // Batch b = [w0 + 64 b, w0 + 64 b + 64) lives in ring slot b & 1.  The
      // group at pos starts in batch b; the ring holds batches b and b + 1,
      // the registers pc/pv hold batch b + 2 (in flight), and kb0/kb1 and
      // sm0/sm1 are the keys and run-start masks of batches b and b + 1.
      // Run lengths then come from scalar bit scans, so the next group's
      // staged data can be read before this group's matrix ops issue.
      float4 pc;
      VisSource<FUSED> src;
      src.vis = vis;
      if constexpr (FUSED) {
        src.perm = perm;
        src.phasor = phasor;
      }
      auto load = [&](uint32_t first) {
        pc = first + uint32_t(lane) < w1 ? coord[first + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
        src.Issue(first, w1, lane);
      };
      src.Prime(w0, w1, lane);
      load(w0);
      StoreStage(ring, lane, pc, src.Value());
      int kb0 = StageKey(w0, w1, lane, pc);
      uint64_t sm0 = RunStarts(kb0, -1, lane);
      load(w0 + kStage);
      StoreStage(ring, kStage + lane, pc, src.Value());
      int kb1 = StageKey(w0 + kStage, w1, lane, pc);
      uint64_t sm1 = RunStarts(kb1, __builtin_amdgcn_readlane(kb0, 63), lane);
      uint32_t bbase = w0, bslot = 0;
      load(w0 + 2 * kStage);
      WaveLdsSync();
      // Each step takes two groups (visibilities pos .. pos + 7) of one run.
      uint32_t pos = w0;
      const float4* view = ring + (hi ? 2 * kStage : 0);
      // the staged values of the current step live in (e0, e1) or (f0, f1)
      // alternately, so the next step's reads stay in flight across the
      // loop edge
      Stage3 e0 = ReadStage(view, j), e1 = ReadStage(view, 4 + j), f0, f1;
      auto step = [&](const Stage3& c0, const Stage3& c1, Stage3& n0, Stage3& n1) {
        const uint32_t o = pos - bbase;  // < 64
        if ((sm0 >> o) & 1ull) {
          if (cur >= 0) WG_FLUSH_MFMA();
          cur = __builtin_amdgcn_readlane(kb0, int(o));
        }
        // bit d - 1 of win: visibility pos + d starts a run
        const uint64_t win = (o < 63 ? (sm0 >> (o + 1)) : 0ull) | (sm1 << (63 - o));
        const int lead = __builtin_ctzll(win | 0x80ull) + 1;
        // The next step's visibilities [npos, npos + 8) lie in batches b and
        // b + 1 (npos - bbase <= 67), which the ring holds and the batch
        // advance below does not overwrite: read them now, a step ahead.
        const uint32_t npos = pos + uint32_t(lead);
        const uint32_t ns = npos - w0 + uint32_t(j);
        n0 = ReadStage(view, int(ns & uint32_t(2 * kStage - 1)));
        n1 = ReadStage(view, int((ns + 4) & uint32_t(2 * kStage - 1)));
        // kernel values: lanes r16 < 8 evaluate ku[tap], r16 >= 8 kv[tap]
        const float k10 = EsPoly<D>(cu, c0.x);
        const float k11 = EsPoly<D>(cu, c1.x);
        // w taps of both groups from one evaluation: quads 0 and 2 of a
        // 16-lane row (r16 & 4 clear) take group 0, quads 1 and 3 group 1
        // (the quads of a pair hold the same taps); each lane then gets the
        // other group's value from the neighbouring quad
        const float kwx = w_ok ? EsPoly<D>(cw, q_hi ? c1.y : c0.y) : 0.f;
        const float kw0 = DppMerge<0x114, 0xA>(kwx);  // quads 1, 3 from quads 0, 2
        const float kw1 = DppMerge<0x104, 0x5>(kwx);  // quads 0, 2 from quads 1, 3
        // A factor: kv (lanes r16 >= 8 own, the others from the rotated
        // row); B factor: ku (lanes r16 < 8 own)
        const float fa0 = DppMerge<0x128, 0x3>(k10), fa1 = DppMerge<0x128, 0x3>(k11);
        const float fb0 = DppMerge<0x128, 0xC>(k10), fb1 = DppMerge<0x128, 0xC>(k11);
        // visibilities past the run contribute zero rows and columns
        const bool take0 = tap_ok && j < lead, take1 = tap_ok && j + 4 < lead;
        const float a0 = take0 ? fa0 * c0.z : 0.f;
        const float a1 = take1 ? fa1 * c1.z : 0.f;
        const float u0 = take0 ? fb0 : 0.f;
        const float u1 = take1 ? fb1 : 0.f;
        if (npos - bbase >= uint32_t(kStage) && npos < w1) {
          // the next group starts in batch b + 1: batch b + 2 replaces b
          StoreStage(ring, bslot * kStage + lane, pc, src.Value());
          const int kb2 = StageKey(bbase + 2 * kStage, w1, lane, pc);
          const uint64_t sm2 = RunStarts(kb2, __builtin_amdgcn_readlane(kb1, 63), lane);
          kb0 = kb1;
          sm0 = sm1;
          kb1 = kb2;
          sm1 = sm2;
          bbase += kStage;
          bslot ^= 1u;
          load(bbase + 2 * kStage);
          WaveLdsSync();
        }
        // B operands next to their use, in the same basic block, so that
        // each quad broadcast folds into its multiply (one DPP-modified op)
#pragma unroll
        for (int m = 0; m < 4; ++m)
          acc[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, DppQuadBcast(kw0, m) * u0, acc[m], 0,
                                                        0, 0);
        if (lead > 4) {
#pragma unroll
          for (int m = 0; m < 4; ++m)
            acc[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, DppQuadBcast(kw1, m) * u1, acc[m],
                                                          0, 0, 0);
        }
        pos = npos;
      };
      for (;;) {
        if (pos >= w1) break;
        step(e0, e1, f0, f1);
        if (pos >= w1) break;
        step(f0, f1, e0, e1);
      }
      if (cur >= 0) WG_FLUSH_MFMA();
    }
#undef WG_FLUSH_MFMA
    __syncthreads();
    float2* slot = scratch + size_t(it - it0) * size_t(ncell);
    for (int i = threadIdx.x; i < ncell; i += 64 * kMfmaWaves)
      slot[i] = make_float2(float(sub_re[i]) * fx_inv, float(sub_im[i]) * fx_inv);
    __syncthreads();
  }
}

// The slots of every tile with more than one work item summed into the
// tile's first slot (compensated, items in order), so that the box gather
// reads one slot per covering tile.  Block (x, y) takes cells
// [256 y, 256 y + 256) of the slot of work item x of the sub-slab, and does
// the work only if item x is the first of a tile with several items; its
// threads read each item's cell in turn (consecutive threads, consecutive
// cells), so a dense tile's items are summed by many blocks at once.
__global__ void __launch_bounds__(256) TileSlotsKernel(const WorkItem* __restrict__ items,
                                                       const uint32_t* __restrict__ item_pos,
                                                       int pos_lo, float2* __restrict__ scratch,
                                                       uint32_t slot) {
  const uint32_t base = item_pos[pos_lo];  // first work item of the sub-slab
  const uint32_t it0 = base + blockIdx.x;
  const uint32_t t = items[it0].tile;
  const uint32_t a = item_pos[pos_lo + t], b = item_pos[pos_lo + t + 1];
  if (it0 != a || b - a < 2) return;  // uniform over the block
  const uint32_t i = blockIdx.y * 256u + threadIdx.x;
  if (i >= slot) return;
  float sx = 0.f, sy = 0.f, cx = 0.f, cy = 0.f;
  for (uint32_t it = a; it < b; ++it) {
    const float2 v = scratch[size_t(it - base) * slot + i];
    const float yx = v.x - cx, yy = v.y - cy;
    const float tx = sx + yx, ty = sy + yy;
    cx = (tx - sx) - yx;
    cy = (ty - sy) - yy;
    sx = tx;
    sy = ty;
  }
  scratch[size_t(a - base) * slot + i] = make_float2(sx, sy);
}

// box[p][r][c] for the planes of one sub-slab: the compensated (Kahan) sum
// of the tile slots (TileSlotsKernel) of every tile whose footprint covers
// (r, c), tiles by index, so the box is deterministic.  Every box cell of
// the sub-slab's planes is written.  Block (x, y, z): columns x * 256 +
// threadIdx.x (grid-strided), box rows y, y + gridDim.y, ..., plane z of
// the sub-slab, so a cell's plane, row and column need no division and a
// row's covering tiles are the same for the whole block.
__global__ void __launch_bounds__(256) GatherBoxKernel(
    DevPlan dp, int W, const uint32_t* __restrict__ item_pos, int pos_lo,
    const float2* __restrict__ scratch, float2* __restrict__ box, int sub_lo, int sub_n,
    int slab_lo) {
  const int E = dp.T + W - 1;
  const int q = blockIdx.z;
  const uint32_t base = item_pos[pos_lo];  // first work item of the sub-slab
  const size_t slot = size_t(sub_n) * E * E;
  for (int r = blockIdx.y; r < dp.BV; r += gridDim.y) {
    // tiles whose footprint [t T, t T + E) contains the row / column
    const int tv0 = max(0, (r - E + dp.T) / dp.T), tv1 = min(dp.ntv - 1, r / dp.T);
    float2* out = box + (size_t(sub_lo + q - slab_lo) * dp.BV + r) * dp.BU;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < dp.BU; c += gridDim.x * blockDim.x) {
      const int tu0 = max(0, (c - E + dp.T) / dp.T), tu1 = min(dp.ntu - 1, c / dp.T);
      float sx = 0.f, sy = 0.f, cx = 0.f, cy = 0.f;
      for (int tv = tv0; tv <= tv1; ++tv) {
        for (int tu = tu0; tu <= tu1; ++tu) {
          const int t = tv * dp.ntu + tu;
          const uint32_t a = item_pos[pos_lo + t], b = item_pos[pos_lo + t + 1];
          if (a == b) continue;
          const size_t off = (size_t(q) * E + (r - tv * dp.T)) * E + (c - tu * dp.T);
          const float2 v = scratch[size_t(a - base) * slot + off];
          const float yx = v.x - cx, yy = v.y - cy;
          const float tx = sx + yx, ty = sy + yy;
          cx = (tx - sx) - yx;
          cy = (ty - sy) - yy;
          sx = tx;
          sy = ty;
        }
      }
      out[c] = make_float2(sx, sy);
    }
  }
}

template <int W>
__global__ void __launch_bounds__(kGridBlock) DegridKernel(
    DevPlan dp, const WorkItem* __restrict__ items, const uint32_t* __restrict__ item_pos,
    int pos_lo, int pos_hi, const float4* __restrict__ coord,
    const float2* __restrict__ box, float2* __restrict__ acc, int sub_lo, int sub_n,
    int slab_lo) {
  extern __shared__ float2 sub[];
  const int E = dp.T + W - 1;
  const int ncell = sub_n * E * E;
  const uint32_t it0 = item_pos[pos_lo], it1 = item_pos[pos_hi];
  for (uint32_t it = it0 + blockIdx.x; it < it1; it += gridDim.x) {
    const WorkItem wi = items[it];
    const int tu = int(wi.tile % uint32_t(dp.ntu)), tv = int(wi.tile / uint32_t(dp.ntu));
    for (int i = threadIdx.x; i < ncell; i += kGridBlock) {
      const int p = i / (E * E), r = (i / E) % E, cc = i % E;
      sub[i] = box[(size_t(sub_lo + p - slab_lo) * dp.BV + size_t(tv * dp.T + r)) * dp.BU +
                   size_t(tu * dp.T + cc)];
    }
    __syncthreads();
    for (uint32_t i = wi.begin + threadIdx.x; i < wi.end; i += kGridBlock) {
      const float4 c = coord[i];
      const uint32_t pk = __float_as_uint(c.w);
      const int du = pk & 255u, dv = (pk >> 8) & 255u, p0 = (pk >> 16) & 0x7fffu;
      float ku[W], kv[W];
      EsTaps<W>(dp, c.x, ku);
      EsTaps<W>(dp, c.y, kv);
      float sx = 0.f, sy = 0.f;
      // w taps in a wave-uniform loop, so each tap's coefficients are
      // scalar loads; lanes skip the planes outside the sub-slab
#pragma unroll 1
      for (int t = 0; t < W; ++t) {
        const int p = p0 + t;
        if (p < sub_lo || p >= sub_lo + sub_n) continue;
        float cw[EsPolyDegree(W) + 1];
        EsPolyLoad<EsPolyDegree(W)>(dp, t, cw);
        const float kw = EsPoly<EsPolyDegree(W)>(cw, c.z);
        const float2* base = sub + ((p - sub_lo) * E + dv) * E + du;
        float px = 0.f, py = 0.f;
#pragma unroll
        for (int l = 0; l < W; ++l) {
          const float2* row = base + l * E;
          float rx = 0.f, ry = 0.f;
#pragma unroll
          for (int k = 0; k < W; ++k) {
            const float2 g = row[k];
            rx += g.x * ku[k];
            ry += g.y * ku[k];
          }
          px += rx * kv[l];
          py += ry * kv[l];
        }
        sx += px * kw;
        sy += py * kw;
      }
      float2 a = acc[i];
      a.x += sx;
      a.y += sy;
      acc[i] = a;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Matrix-core degridding for W <= 8.
//
// Inside a run the visibilities read the same W x W x W block G of the box:
//   V_j = sum_t kw_j[t] sum_l kv_j[l] sum_k G[t][l][k] ku_j[k].
// The u and v contractions together are one GEMM with the block as the A
// operand: for a group of 16 visibilities, 16 FP32 MFMA 16x16x4 steps over
// K = (k, l) give  D[(t,c)][j] = sum_{k,l} G[t][l][k].c ku_j[k] kv_j[l],
// with A[(t,c)][(k,l)] = G[t][l][k].c and B[(k,l)][j] = ku_j[k] kv_j[l].
// Step s covers (k, l) = (4 (s % 2) + g, s / 2) for g = 0..3.  Lane L holds
// D[4 (L/16) + r][L % 16]: visibility j = L % 16 and rows (t, c) =
// (2 (L/16) + r/2, r % 2), so only the w contraction (two planes per lane)
// remains in registers, and the four lanes of a visibility add their parts
// with lane exchanges.  The 16 A registers are reloaded from the LDS box
// block only when the run changes.  Kernel values: each lane evaluates two
// u taps (g, 4 + g), two v taps and the two w taps (2g, 2g + 1) of its own
// rows, and gets the other six v taps of its visibility from the lanes of
// the same visibility.
// Coordinates and the accumulated visibilities are staged per wave through
// an LDS ring with run-start masks, as in the gridder.
constexpr int kDegridGroup = 16;

// FUSED (every visibility in exactly one sub-slab): the prediction goes
// straight to the caller's out[row][chan] through perm, rotated back by the
// conjugate phasor and flipped, instead of into the sorted accumulator.
template <int W, bool FUSED>
__global__ void __launch_bounds__(64 * kMfmaWaves)
    __attribute__((amdgpu_waves_per_eu(4, 4))) DegridKernelMfma(
    DevPlan dp, const WorkItem* __restrict__ items, const uint32_t* __restrict__ item_pos,
    int pos_lo, int pos_hi, const float4* __restrict__ coord, const float2* __restrict__ box,
    float2* __restrict__ acc, const uint32_t* __restrict__ perm, const float2* __restrict__ phasor,
    int sub_lo, int sub_n, int slab_lo, uint32_t* __restrict__ queue) {
  static_assert(W <= 8 && W % 2 == 0, "matrix-core degridding covers W <= 8");
  __shared__ uint32_t next_item;
  extern __shared__ __attribute__((aligned(16))) float2 degrid_lds[];
  float2* sub = degrid_lds;  // [sub_n][E][E] box block of the item's tile
  const int E = dp.T + W - 1;
  const int ncell = sub_n * E * E;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // per wave 2 x 64 staged (coordinate, accumulated visibility) entries
  float4* ring_c = reinterpret_cast<float4*>(degrid_lds + ((ncell + 1) & ~1)) +
                   wave * (2 * kStage);
  float2* ring_a = reinterpret_cast<float2*>(reinterpret_cast<float4*>(
                       degrid_lds + ((ncell + 1) & ~1)) + kMfmaWaves * (2 * kStage)) +
                   wave * (2 * kStage);
  const int j = lane & 15;        // visibility of the group (B column, D column)
  const int g = lane >> 4;        // B row (k, l) = (4 (s % 2) + g, s / 2); D rows 4 g + r
  // the lane evaluates u taps g and 4 + g, and v and w taps 2 g and
  // 2 g + 1 (the coefficient rows of taps >= W are zero, so those taps
  // weigh 0)
  constexpr int D = EsPolyDegree(W);
  float ca[D + 1], cb[D + 1], c2a[D + 1], c2b[D + 1];
  EsPolyLoad<D>(dp, g, ca);
  EsPolyLoad<D>(dp, 4 + g, cb);
  EsPolyLoad<D>(dp, 2 * g, c2a);
  EsPolyLoad<D>(dp, 2 * g + 1, c2b);
  // A operand lane mapping: A[i = L % 16][g] of step s =
  // G[t = i / 2][l = s / 2][k = 4 (s % 2) + g].(i % 2)
  const int a_t = (lane & 15) >> 1, a_c = lane & 1;
  const uint32_t it0 = item_pos[pos_lo], it1 = item_pos[pos_hi];
  for (;;) {
    if (threadIdx.x == 0) next_item = atomicAdd(queue, 1u);
    __syncthreads();
    const uint32_t it = it0 + next_item;
    if (it >= it1) break;
    const WorkItem wi = items[it];
    const int tu = int(wi.tile % uint32_t(dp.ntu)), tv = int(wi.tile / uint32_t(dp.ntu));
    for (int i = threadIdx.x; i < ncell; i += 64 * kMfmaWaves) {
      const int p = i / (E * E), r = (i / E) % E, c = i % E;
      sub[i] = box[(size_t(sub_lo + p - slab_lo) * dp.BV + size_t(tv * dp.T + r)) * dp.BU +
                   size_t(tu * dp.T + c)];
    }
    __syncthreads();
    const uint32_t n = wi.end - wi.begin;
    const uint32_t per = (n + kMfmaWaves - 1) / kMfmaWaves;
    const uint32_t w0 = wi.begin + min(n, wave * per);
    const uint32_t w1 = wi.begin + min(n, (wave + 1) * per);
    if (w0 < w1) {
      auto stage = [&](uint32_t first, float4& c, float2& a) {
        c = make_float4(0.f, 0.f, 0.f, 0.f);
        a = make_float2(0.f, 0.f);
        if (first + uint32_t(lane) < w1) {
          c = coord[first + lane];
          if constexpr (!FUSED) a = acc[first + lane];
        }
      };
      float4 pc;
      float2 pa;
      stage(w0, pc, pa);
      ring_c[lane] = pc;
      ring_a[lane] = pa;
      int kb0 = StageKey(w0, w1, lane, pc);
      uint64_t sm0 = RunStarts(kb0, -1, lane);
      stage(w0 + kStage, pc, pa);
      ring_c[kStage + lane] = pc;
      ring_a[kStage + lane] = pa;
      int kb1 = StageKey(w0 + kStage, w1, lane, pc);
      uint64_t sm1 = RunStarts(kb1, __builtin_amdgcn_readlane(kb0, 63), lane);
      uint32_t bbase = w0, bslot = 0;
      stage(w0 + 2 * kStage, pc, pa);
      WaveLdsSync();
      float areg[16];  // A operands of the current run, one per step
      int cur = -1;
      uint32_t pos = w0;
      float4 cv = ring_c[j];
      float2 av = ring_a[j];
      while (pos < w1) {
        // fused scatter: this group's permutation words and phasors, loaded
        // now and used after the matrix ops
        uint32_t pw_f = 0;
        float2 ph_f = make_float2(0.f, 0.f);
        if constexpr (FUSED) {
          if (g == 0 && pos + uint32_t(j) < w1) {
            pw_f = perm[pos + j];
            ph_f = phasor[pos + j];
          }
        }
        const uint32_t o = pos - bbase;  // < 64
        if ((sm0 >> o) & 1ull) {
          cur = __builtin_amdgcn_readlane(kb0, int(o));
          const int du = cur & 255, dv = (cur >> 8) & 255;
          const int q0 = ((cur >> 16) & 0x7fff) - sub_lo;
          const int q = q0 + a_t;
          const bool q_ok = a_t < W && q >= 0 && q < sub_n;
#pragma unroll
          for (int st = 0; st < 16; ++st) {
            const int l = st >> 1, k = 4 * (st & 1) + g;
            float v = 0.f;
            if (q_ok && l < W && k < W) {
              const float2 gv = sub[(q * E + dv + l) * E + du + k];
              v = a_c ? gv.y : gv.x;
            }
            areg[st] = v;
          }
        }
        // bit d - 1 of win: visibility pos + d starts a run
        const uint64_t win = (o < 63 ? (sm0 >> (o + 1)) : 0ull) | (sm1 << (63 - o));
        const int lead = __builtin_ctzll(win | (1ull << (kDegridGroup - 1))) + 1;
        const uint32_t npos = pos + uint32_t(lead);
        // next group's staged data (batches b, b + 1; see the gridder)
        const uint32_t ns = (npos - w0 + uint32_t(j)) & uint32_t(2 * kStage - 1);
        const float4 cn = ring_c[ns];
        const float2 an = ring_a[ns];
        const bool take = j < lead;
        // B operands: u taps g (even steps) and 4 + g (odd steps) of
        // visibility j, times its v tap of the step
        const float bu0 = take ? EsPoly<D>(ca, cv.x) : 0.f;
        const float bu1 = take ? EsPoly<D>(cb, cv.x) : 0.f;
        // v taps 2 g, 2 g + 1 here; all eight from the four lanes of j
        // (lane 16 p + j holds taps 2 p, 2 p + 1 of visibility j)
        float kv[8];
        {
          const float v0 = EsPoly<D>(c2a, cv.y), v1 = EsPoly<D>(c2b, cv.y);
#pragma unroll
          for (int pp = 0; pp < 4; ++pp) {
            kv[2 * pp] = __shfl(v0, 16 * pp + j);
            kv[2 * pp + 1] = __shfl(v1, 16 * pp + j);
          }
        }
        // w taps of this lane's rows: planes 2 g, 2 g + 1
        const float kw0 = EsPoly<D>(c2a, cv.z), kw1 = EsPoly<D>(c2b, cv.z);
        // four accumulators (steps mod 4): four independent chains
        mfma_f4 dq[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) dq[c] = mfma_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int st = 0; st < 16; ++st)
          dq[st & 3] = __builtin_amdgcn_mfma_f32_16x16x4f32(
              areg[st], ((st & 1) ? bu1 : bu0) * kv[st >> 1], dq[st & 3], 0, 0, 0);
        const mfma_f4 d = (dq[0] + dq[1]) + (dq[2] + dq[3]);
        // rows 4 g + r = (t, c) = (2 g + r / 2, r % 2)
        float part = kw0 * d[0] + kw1 * d[2];
        float im = kw0 * d[1] + kw1 * d[3];
        // the four lanes of visibility j (L, L ^ 16, L ^ 32, L ^ 48)
        part += __shfl_xor(part, 16);
        im += __shfl_xor(im, 16);
        part += __shfl_xor(part, 32);
        im += __shfl_xor(im, 32);
        if constexpr (FUSED) {
          if (g == 0 && take) {
            float2 V = make_float2(part * ph_f.x + im * ph_f.y, im * ph_f.x - part * ph_f.y);
            if (pw_f & kPermFlip) V.y = -V.y;
            acc[pw_f & ~kPermFlip] = V;
          }
        } else {
          if (g == 0 && take) acc[pos + j] = make_float2(av.x + part, av.y + im);
        }
        if (npos - bbase >= uint32_t(kStage) && npos < w1) {
          ring_c[bslot * kStage + lane] = pc;
          ring_a[bslot * kStage + lane] = pa;
          const int kb2 = StageKey(bbase + 2 * kStage, w1, lane, pc);
          const uint64_t sm2 = RunStarts(kb2, __builtin_amdgcn_readlane(kb1, 63), lane);
          kb0 = kb1;
          sm0 = sm1;
          kb1 = kb2;
          sm1 = sm2;
          bbase += kStage;
          bslot ^= 1u;
          stage(bbase + 2 * kStage, pc, pa);
          WaveLdsSync();
        }
        cv = cn;
        av = an;
        pos = npos;
      }
    }
    __syncthreads();
  }
}

// Work-item counts for every (sub-slab, tile); with `sub_vis` non-null also
// the number of visibility passes of every sub-slab (sum over its tiles).
__global__ void ItemCountKernel(DevPlan dp, const uint32_t* __restrict__ off, int Sg, int nsub,
                                uint32_t* __restrict__ counts,
                                unsigned long long* __restrict__ sub_vis) {
  const size_t ntiles = size_t(dp.ntu) * dp.ntv;
  const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool in = idx < ntiles * nsub;
  const int g = in ? int(idx / ntiles) : nsub;
  uint32_t n = 0;
  if (in) {
    const size_t t = idx % ntiles;
    const int sub_lo = g * Sg;
    const int lo = max(0, sub_lo - dp.W + 1);
    const int hi = min(dp.P0 - 1, sub_lo + Sg - 1);
    if (lo <= hi) n = off[t * dp.P0 + hi + 1] - off[t * dp.P0 + lo];
    counts[idx] = (n + kChunk - 1) / kChunk;
  }
  if (!sub_vis) return;
  // one atomic per wave when the whole wave lies in one sub-slab
  const int g0 = __builtin_amdgcn_readfirstlane(g);
  if (__all(g == g0)) {
    unsigned long long s = n;
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
    if ((threadIdx.x & 63) == 0 && g0 < nsub) atomicAdd(&sub_vis[g0], s);
  } else if (in) {
    atomicAdd(&sub_vis[g], (unsigned long long)n);
  }
}

__global__ void ItemEmitKernel(DevPlan dp, const uint32_t* __restrict__ off, int Sg, int nsub,
                               const uint32_t* __restrict__ pos, WorkItem* __restrict__ items) {
  const size_t ntiles = size_t(dp.ntu) * dp.ntv;
  const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= ntiles * nsub) return;
  const int g = int(idx / ntiles);
  const size_t t = idx % ntiles;
  const int sub_lo = g * Sg;
  const int lo = max(0, sub_lo - dp.W + 1);
  const int hi = min(dp.P0 - 1, sub_lo + Sg - 1);
  if (lo > hi) return;
  const uint32_t b = off[t * dp.P0 + lo], e = off[t * dp.P0 + hi + 1];
  uint32_t p = pos[idx];
  for (uint32_t s = b; s < e; s += kChunk) {
    WorkItem w;
    w.tile = uint32_t(t);
    w.begin = s;
    w.end = min(e, s + kChunk);
    items[p++] = w;
  }
}

template <typename K>
void AllowLds(K kernel, size_t bytes) {
  AllowDynamicLds(reinterpret_cast<const void*>(kernel), bytes);
}

template <typename F>
void DispatchW(int W, F&& f) {
  switch (W) {
    case 4: f(std::integral_constant<int, 4>()); break;
    case 6: f(std::integral_constant<int, 6>()); break;
    case 8: f(std::integral_constant<int, 8>()); break;
    case 10: f(std::integral_constant<int, 10>()); break;
    case 12: f(std::integral_constant<int, 12>()); break;
    case 14: f(std::integral_constant<int, 14>()); break;
    case 16: f(std::integral_constant<int, 16>()); break;
    default: throw Error("unsupported kernel support " + std::to_string(W));
  }
}

int GridBlocks() { return DeviceCUs() * 8; }

}  // namespace

size_t GridLdsBytes(const DevPlan& dp, int sub_n) {
  const size_t E = size_t(dp.T + dp.W - 1);
  return size_t(sub_n) * E * E * sizeof(float2);
}

void BuildItems(const DevPlan& dp, const uint32_t* off, int Sg, int nsub, uint32_t* counts,
                uint32_t* pos, uint32_t* scan_scratch, WorkItem* items,
                unsigned long long* sub_vis, hipStream_t s, bool emit) {
  const size_t n = size_t(dp.ntu) * dp.ntv * nsub;
  const unsigned nb = unsigned((n + 255) / 256);
  if (!emit) {
    ItemCountKernel<<<nb, 256, 0, s>>>(dp, off, Sg, nsub, counts, sub_vis);
    DeviceExclusiveScan(counts, pos, n, scan_scratch, s, pos + n);
  } else {
    ItemEmitKernel<<<nb, 256, 0, s>>>(dp, off, Sg, nsub, pos, items);
  }
  WG_HIP(hipGetLastError());
}

void LaunchGrid(const DevPlan& dp, const VisSet& vs, const float2* vis_sorted, float2* scratch,
                uint32_t* queue, const float2* bounds, int sub, int sub_lo, int sub_n,
                hipStream_t s, const float2* vis_raw, const uint32_t* gather_perm) {
  WG_CHECK(dp.W > 8 || bounds, "the matrix-core gridders need the items' bounds");
  if (dp.split && dp.W <= 8) {
    LaunchGridSplit(dp, vs, vis_sorted, scratch, queue, bounds, sub, sub_lo, sub_n, s, vis_raw,
                    gather_perm);
    return;
  }
  const size_t ntiles = size_t(dp.ntu) * dp.ntv;
  const size_t lds = GridLdsBytes(dp, sub_n);
  WG_CHECK(sub_n >= 1 && sub_n <= kMaxSubSlab, "sub-slab depth out of range");
  const int pos_lo = int(sub * ntiles), pos_hi = int((sub + 1) * ntiles);
  DispatchW(dp.W, [&](auto Wc) {
    constexpr int W = decltype(Wc)::value;
    if constexpr (W <= 8) {
      // subgrid (16-byte aligned) + per-wave staging rings
      constexpr size_t ring = size_t(kMfmaWaves) * 2 * (2 * kStage) * sizeof(float4);
      static_assert(ring + 16 <= size_t(kGridLdsReserve), "staging rings exceed the LDS reserve");
      const size_t total = ((lds + 15) & ~size_t(15)) + ring;
      WG_CHECK(total <= 160 * 1024, "gridding LDS exceeds 160 KiB");
      auto launch = [&](auto kernel, const float2* v, const uint32_t* pm) {
        AllowLds(kernel, total);
        // one persistent workgroup per resident slot; they drain the queue
        const int per_cu =
            ResidentBlocksPerCU(reinterpret_cast<const void*>(kernel), 64 * kMfmaWaves, total);
        const int blocks = std::max(1, per_cu) * DeviceCUs();
        LaunchFillZero(queue, sizeof(uint32_t), s);
        kernel<<<blocks, 64 * kMfmaWaves, total, s>>>(dp, vs.items, vs.item_pos, pos_lo, pos_hi,
                                                      vs.coord, v, pm, vs.phasor, scratch, sub_lo,
                                                      sub_n, queue, bounds);
      };
      if (vis_raw)
        launch(GridKernelMfma<W, true>, vis_raw, gather_perm);
      else
        launch(GridKernelMfma<W, false>, vis_sorted, nullptr);
    } else {
      WG_CHECK(!vis_raw, "the fused gather needs the matrix-core gridder (W <= 8)");
      AllowLds(GridKernel<W>, lds);
      GridKernel<W><<<GridBlocks(), kGridBlock, lds, s>>>(dp, vs.items, vs.item_pos, pos_lo,
                                                           pos_hi, vs.coord, vis_sorted, scratch,
                                                           sub_lo, sub_n);
    }
  });
  WG_HIP(hipGetLastError());
}

void LaunchGatherBox(const DevPlan& dp, const VisSet& vs, float2* box, float2* scratch,
                     int sub, int sub_lo, int sub_n, int slab_lo, hipStream_t s) {
  const size_t ntiles = size_t(dp.ntu) * dp.ntv;
  const int pos_lo = int(sub * ntiles);
  const size_t E = size_t(dp.T + dp.W - 1);
  WG_CHECK(ntiles < 0x80000000ull && size_t(sub_n) * E * E < 0xFFFFFFFFull,
           "tile grid too large for the box gather");
  const uint32_t slot = uint32_t(size_t(sub_n) * E * E);
  WG_CHECK(size_t(sub) < vs.sub_items.size(), "sub-slab out of range");
  if (vs.sub_items[sub] > 0)
    TileSlotsKernel<<<dim3(vs.sub_items[sub], (slot + 255) / 256), 256, 0, s>>>(
        vs.items, vs.item_pos, pos_lo, scratch, slot);
  const dim3 grid(unsigned((dp.BU + 255) / 256), unsigned(std::min(dp.BV, 65535)),
                  unsigned(sub_n));
  GatherBoxKernel<<<grid, 256, 0, s>>>(dp, dp.W, vs.item_pos, pos_lo, scratch, box, sub_lo, sub_n,
                                       slab_lo);
  WG_HIP(hipGetLastError());
}

size_t GridScratchBytes(const DevPlan& dp, int Sg, size_t max_items_per_sub) {
  const size_t E = size_t(dp.T + dp.W - 1);
  return max_items_per_sub * size_t(Sg) * E * E * sizeof(float2);
}

void LaunchDegrid(const DevPlan& dp, const VisSet& vs, const float2* box, float2* acc,
                  uint32_t* queue, int sub, int sub_lo, int sub_n, int slab_lo, hipStream_t s,
                  const uint32_t* scatter_perm) {
  if (dp.split && dp.W <= 8) {
    LaunchDegridSplit(dp, vs, box, acc, queue, sub, sub_lo, sub_n, slab_lo, s, scatter_perm);
    return;
  }
  const size_t ntiles = size_t(dp.ntu) * dp.ntv;
  const size_t lds = GridLdsBytes(dp, sub_n);
  WG_CHECK(sub_n >= 1 && sub_n <= kMaxSubSlab, "sub-slab depth out of range");
  DispatchW(dp.W, [&](auto Wc) {
    constexpr int W = decltype(Wc)::value;
    if constexpr (W <= 8) {
      // box block (16-byte aligned) + per-wave rings of float4 + float2
      constexpr size_t ring = size_t(kMfmaWaves) * (2 * kStage) * 24;
      static_assert(ring + 16 <= size_t(kGridLdsReserve), "staging rings exceed the reserve");
      const size_t total = ((lds + 15) & ~size_t(15)) + ring;
      WG_CHECK(total <= 160 * 1024, "degridding LDS exceeds 160 KiB");
      auto launch = [&](auto kernel) {
        AllowLds(kernel, total);
        const int per_cu =
            ResidentBlocksPerCU(reinterpret_cast<const void*>(kernel), 64 * kMfmaWaves, total);
        LaunchFillZero(queue, sizeof(uint32_t), s);
        kernel<<<std::max(1, per_cu) * DeviceCUs(), 64 * kMfmaWaves, total, s>>>(
            dp, vs.items, vs.item_pos, int(sub * ntiles), int((sub + 1) * ntiles), vs.coord, box,
            acc, scatter_perm, vs.phasor, sub_lo, sub_n, slab_lo, queue);
      };
      if (scatter_perm)
        launch(DegridKernelMfma<W, true>);
      else
        launch(DegridKernelMfma<W, false>);
    } else {
      WG_CHECK(!scatter_perm, "the fused scatter needs the matrix-core degridder (W <= 8)");
      AllowLds(DegridKernel<W>, lds);
      DegridKernel<W><<<GridBlocks(), kGridBlock, lds, s>>>(
          dp, vs.items, vs.item_pos, int(sub * ntiles), int((sub + 1) * ntiles), vs.coord, box,
          acc, sub_lo, sub_n, slab_lo);
    }
  });
  WG_HIP(hipGetLastError());
}

}  // namespace wg
