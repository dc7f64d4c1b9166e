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

// Index of lane `lane` of the batch at `first`, clamped to the range's last
// element (first < end): every lane loads in bounds without a branch, and
// lanes past the end are masked where their values are used.
__device__ __forceinline__ uint32_t StageIndex(uint32_t first, uint32_t end, int lane) {
  return min(first + uint32_t(lane), end - 1u);
}

// Visibilities of lanes past the end of the range are 0.
template <>
struct VisSource<false> {
  const float2* vis;
  float2 v;
  bool in;
  __device__ __forceinline__ void Prime(uint32_t, uint32_t, int) {}
  __device__ __forceinline__ void Issue(uint32_t first, uint32_t end, int lane) {
    in = first + uint32_t(lane) < end;
    v = vis[StageIndex(first, end, lane)];
  }
  __device__ __forceinline__ float2 Value() const {
    return in ? v : make_float2(0.f, 0.f);
  }
};

template <>
struct VisSource<true> {
  const float2* vis;
  const uint32_t* perm;
  const float2* phasor;
  uint32_t pw_next, pw;
  float2 raw, ph;
  bool in;
  // permutation word of the batch the next Issue() loads
  __device__ __forceinline__ void Prime(uint32_t first, uint32_t end, int lane) {
    pw_next = perm[StageIndex(first, end, lane)];
  }
  __device__ __forceinline__ void Issue(uint32_t first, uint32_t end, int lane) {
    pw = pw_next;
    in = first + uint32_t(lane) < end;
    raw = vis[pw & ~kPermFlip];
    ph = phasor[StageIndex(first, end, lane)];
    pw_next = perm[min(first + uint32_t(kStage) + uint32_t(lane), end - 1u)];
  }
  __device__ __forceinline__ float2 Value() const {
    const float2 V = make_float2(raw.x, (pw & kPermFlip) ? -raw.y : raw.y);
    return in ? make_float2(V.x * ph.x - V.y * ph.y, V.x * ph.y + V.y * ph.x)
              : make_float2(0.f, 0.f);
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
