// ES-kernel gridding and degridding on LDS-staged tile subgrids.
//
// A work item is one uv tile (T x T first-cells) and a range of its sorted
// visibilities that touch the planes [sub_lo, sub_lo + sub_n) of one
// gridding sub-slab.  Its footprint in the grid is E x E cells per plane
// (E = T + W - 1).  Gridding accumulates every visibility's W x W x W
// kernel footprint into the LDS subgrid with LDS float atomics and flushes
// the non-zero cells to the HBM plane box with global float atomics
// (neighbouring tiles overlap by W - 1 cells).  Degridding stages the same
// subgrid from HBM into LDS and gathers per visibility in registers.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "engine.hpp"

namespace wg {

namespace {

constexpr int kGridBlock = 256;
constexpr uint32_t kChunk = 2048;  // visibilities per work item

template <int W>
__global__ void __launch_bounds__(kGridBlock) GridKernel(
    DevPlan dp, const WorkItem* __restrict__ items, const uint32_t* __restrict__ item_pos,
    int pos_lo, int pos_hi, const float4* __restrict__ coord,
    const float2* __restrict__ vis, float2* __restrict__ box, int sub_lo, int sub_n,
    int slab_lo) {
  extern __shared__ float2 sub[];
  const int E = dp.T + W - 1;
  const int ncell = sub_n * E * E;
  const uint32_t it0 = item_pos[pos_lo], it1 = item_pos[pos_hi];
  for (uint32_t it = it0 + blockIdx.x; it < it1; it += gridDim.x) {
    const WorkItem wi = items[it];
    for (int i = threadIdx.x; i < ncell; i += kGridBlock) sub[i] = make_float2(0.f, 0.f);
    __syncthreads();
    for (uint32_t i = wi.begin + threadIdx.x; i < wi.end; i += kGridBlock) {
      const float4 c = coord[i];
      const float2 V = vis[i];
      const uint32_t pk = __float_as_uint(c.w);
      const int du = pk & 255u, dv = (pk >> 8) & 255u, p0 = (pk >> 16) & 0x7fffu;
      float ku[W], kv[W];
      EsTaps<W>(c.x, dp.beta, ku);
      EsTaps<W>(c.y, dp.beta, kv);
      const int plo = max(p0, sub_lo), phi = min(p0 + W, sub_lo + sub_n);
      for (int p = plo; p < phi; ++p) {
        const float kw =
            EsKernel((float(p - p0 - W / 2 + 1) - c.z) * dp.two_over_w, dp.beta);
        const float2 vw = make_float2(V.x * kw, V.y * kw);
        float2* base = sub + ((p - sub_lo) * E + dv) * E + du;
#pragma unroll
        for (int l = 0; l < W; ++l) {
          const float2 vl = make_float2(vw.x * kv[l], vw.y * kv[l]);
          float2* row = base + l * E;
#pragma unroll
          for (int k = 0; k < W; ++k) {
            atomicAdd(&row[k].x, vl.x * ku[k]);
            atomicAdd(&row[k].y, vl.y * ku[k]);
          }
        }
      }
    }
    __syncthreads();
    const int tu = int(wi.tile % uint32_t(dp.ntu)), tv = int(wi.tile / uint32_t(dp.ntu));
    for (int i = threadIdx.x; i < ncell; i += kGridBlock) {
      const float2 val = sub[i];
      if (val.x != 0.f || val.y != 0.f) {
        const int p = i / (E * E), r = (i / E) % E, cc = i % E;
        float2* dst = box + (size_t(sub_lo + p - slab_lo) * dp.BV + size_t(tv * dp.T + r)) *
                                dp.BU + size_t(tu * dp.T + cc);
        atomicAdd(&dst->x, val.x);
        atomicAdd(&dst->y, val.y);
      }
    }
    __syncthreads();
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
      EsTaps<W>(c.x, dp.beta, ku);
      EsTaps<W>(c.y, dp.beta, kv);
      const int plo = max(p0, sub_lo), phi = min(p0 + W, sub_lo + sub_n);
      float sx = 0.f, sy = 0.f;
      for (int p = plo; p < phi; ++p) {
        const float kw =
            EsKernel((float(p - p0 - W / 2 + 1) - c.z) * dp.two_over_w, dp.beta);
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

// Work-item counts for every (sub-slab, tile).
__global__ void ItemCountKernel(DevPlan dp, const uint32_t* __restrict__ off, int Sg, int nsub,
                                uint32_t* __restrict__ counts) {
  const size_t ntiles = size_t(dp.ntu) * dp.ntv;
  const size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= ntiles * nsub) return;
  const int g = int(idx / ntiles);
  const size_t t = idx % ntiles;
  const int sub_lo = g * Sg;
  const int lo = max(0, sub_lo - dp.W + 1);
  const int hi = min(dp.P0 - 1, sub_lo + Sg - 1);
  uint32_t n = 0;
  if (lo <= hi) n = off[t * dp.P0 + hi + 1] - off[t * dp.P0 + lo];
  counts[idx] = (n + kChunk - 1) / kChunk;
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
  if (bytes > 48 * 1024)
    WG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
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

int GridBlocks() {
  static int n = [] {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return cus * 8;
  }();
  return n;
}

}  // namespace

size_t GridLdsBytes(const DevPlan& dp, int sub_n) {
  const size_t E = size_t(dp.T + dp.W - 1);
  return size_t(sub_n) * E * E * sizeof(float2);
}

void BuildItems(const DevPlan& dp, const uint32_t* off, int Sg, int nsub, uint32_t* counts,
                uint32_t* pos, uint32_t* scan_scratch, WorkItem* items, hipStream_t s,
                bool emit) {
  const size_t n = size_t(dp.ntu) * dp.ntv * nsub;
  const unsigned nb = unsigned((n + 255) / 256);
  if (!emit) {
    ItemCountKernel<<<nb, 256, 0, s>>>(dp, off, Sg, nsub, counts);
    DeviceExclusiveScan(counts, pos, n, scan_scratch, s, pos + n);
  } else {
    ItemEmitKernel<<<nb, 256, 0, s>>>(dp, off, Sg, nsub, pos, items);
  }
  WG_HIP(hipGetLastError());
}

void LaunchGrid(const DevPlan& dp, const VisSet& vs, const float2* vis_sorted, float2* box,
                int sub, int sub_lo, int sub_n, int slab_lo, hipStream_t s) {
  const size_t ntiles = size_t(dp.ntu) * dp.ntv;
  const size_t lds = GridLdsBytes(dp, sub_n);
  DispatchW(dp.W, [&](auto Wc) {
    constexpr int W = decltype(Wc)::value;
    AllowLds(GridKernel<W>, lds);
    GridKernel<W><<<GridBlocks(), kGridBlock, lds, s>>>(
        dp, vs.items, vs.item_pos, int(sub * ntiles), int((sub + 1) * ntiles), vs.coord,
        vis_sorted, box, sub_lo, sub_n, slab_lo);
  });
  WG_HIP(hipGetLastError());
}

void LaunchDegrid(const DevPlan& dp, const VisSet& vs, const float2* box, float2* acc, int sub,
                  int sub_lo, int sub_n, int slab_lo, hipStream_t s) {
  const size_t ntiles = size_t(dp.ntu) * dp.ntv;
  const size_t lds = GridLdsBytes(dp, sub_n);
  DispatchW(dp.W, [&](auto Wc) {
    constexpr int W = decltype(Wc)::value;
    AllowLds(DegridKernel<W>, lds);
    DegridKernel<W><<<GridBlocks(), kGridBlock, lds, s>>>(
        dp, vs.items, vs.item_pos, int(sub * ntiles), int((sub + 1) * ntiles), vs.coord, box,
        acc, sub_lo, sub_n, slab_lo);
  });
  WG_HIP(hipGetLastError());
}

}  // namespace wg
