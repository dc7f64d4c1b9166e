// Device helpers shared by the matrix-core gridding kernels (FP32 and
// split-precision): launch shape, the staged visibility source with the
// fused gather, run keys and run-start masks, and the wave-local LDS sync.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"

namespace wg {
namespace {

constexpr int kMfmaWaves = 8;
constexpr int kStage = 64;

// Source of the visibilities the gridding kernel stages, one batch of 64 in
// flight.  VisSource<false> reads the gathered (sorted, flipped, shifted)
// visibilities.  VisSource<true> fuses the gather: it reads the caller's
// visibility through the sorted permutation and applies the flip and the
// sorted phasor itself.  The permutation word of a batch is loaded one
// batch before its visibility, so the dependent random load never waits.
template <bool FUSED>
struct VisSource;

template <>
struct VisSource<false> {
  const float2* vis;
  float2 v;
  __device__ __forceinline__ void Prime(uint32_t, uint32_t, int) {}
  __device__ __forceinline__ void Issue(uint32_t first, uint32_t end, int lane) {
    v = first + uint32_t(lane) < end ? vis[first + lane] : make_float2(0.f, 0.f);
  }
  __device__ __forceinline__ float2 Value() const { return v; }
};

template <>
struct VisSource<true> {
  const float2* vis;
  const uint32_t* perm;
  const float2* phasor;
  uint32_t pw_next, pw;
  float2 raw, ph;
  // permutation word of the batch the next Issue() loads
  __device__ __forceinline__ void Prime(uint32_t first, uint32_t end, int lane) {
    pw_next = first + uint32_t(lane) < end ? perm[first + lane] : 0u;
  }
  __device__ __forceinline__ void Issue(uint32_t first, uint32_t end, int lane) {
    pw = pw_next;
    const bool in = first + uint32_t(lane) < end;
    raw = in ? vis[pw & ~kPermFlip] : make_float2(0.f, 0.f);
    ph = in ? phasor[first + lane] : make_float2(0.f, 0.f);
    const uint32_t nxt = first + uint32_t(kStage) + uint32_t(lane);
    pw_next = nxt < end ? perm[nxt] : 0u;
  }
  __device__ __forceinline__ float2 Value() const {
    const float2 V = make_float2(raw.x, (pw & kPermFlip) ? -raw.y : raw.y);
    return make_float2(V.x * ph.x - V.y * ph.y, V.x * ph.y + V.y * ph.x);
  }
};

// Run key (packed cell and first plane) of the lane's staged visibility;
// -2 past the end of the range, so the end starts a "run".
__device__ __forceinline__ int StageKey(uint32_t first, uint32_t end, int lane, const float4& c) {
  return first + uint32_t(lane) < end ? (__float_as_int(c.w) & 0x7fffffff) : -2;
}

// Bit i set: staged visibility i starts a run (its key differs from that of
// the visibility before it; prev_last is the key before lane 0).
__device__ __forceinline__ uint64_t RunStarts(int key, int prev_last, int lane) {
  int prev = __shfl_up(key, 1);
  if (lane == 0) prev = prev_last;
  return __ballot(key != prev);
}

// Orders one wave's LDS stores before its lanes read each other's values.
__device__ __forceinline__ void WaveLdsSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef float mfma_f4 __attribute__((ext_vector_type(4)));

}  // namespace
}  // namespace wg
