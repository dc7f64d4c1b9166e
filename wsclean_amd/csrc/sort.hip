// Integer bucket sort of visibilities into (uv tile, first w plane) buckets.
//
// keys are uint32 bucket ids (tile * P0 + p0; nkeys marks skipped zero
// visibilities), values are the flat visibility index row * n_chan + chan.
// The sort is a stable LSD radix sort with 8-bit digits, so the permutation
// is identical to a stable CPU sort (numpy argsort(kind="stable")).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.hpp"
#include "engine.hpp"

namespace wg {

namespace {

constexpr int kRsBlock = 256;
constexpr int kRsRounds = 32;
constexpr int kRsTile = kRsBlock * kRsRounds;

constexpr int kScanBlock = 1024;
constexpr int kScanPer = 8;
constexpr int kScanTile = kScanBlock * kScanPer;

__device__ __forceinline__ uint32_t WaveInclusiveScan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the total.
__device__ uint32_t BlockExclusiveScan(uint32_t v, uint32_t& excl) {
  __shared__ uint32_t wsum[kScanBlock / 64];
  __shared__ uint32_t tot;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwaves = (blockDim.x + 63) >> 6;
  const uint32_t inc = WaveInclusiveScan(v);
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < nwaves; ++w) {
      const uint32_t t = wsum[w];
      wsum[w] = run;
      run += t;
    }
    tot = run;
  }
  __syncthreads();
  excl = inc - v + wsum[wave];
  const uint32_t total = tot;
  __syncthreads();
  return total;
}

__global__ void ScanReduceKernel(const uint32_t* __restrict__ in, size_t n,
                                 uint32_t* __restrict__ bsum) {
  const size_t base = size_t(blockIdx.x) * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    size_t i = base + size_t(k) * kScanBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t ex;
  uint32_t tot = BlockExclusiveScan(s, ex);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Exclusive scan of bsum[0..nb) in one block (sequential chunks).
__global__ void ScanBlockSumsKernel(uint32_t* __restrict__ bsum, uint32_t nb,
                                    uint32_t* __restrict__ total) {
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < nb; base += kScanBlock) {
    uint32_t i = base + threadIdx.x;
    uint32_t v = i < nb ? bsum[i] : 0;
    uint32_t ex;
    uint32_t tot = BlockExclusiveScan(v, ex);
    if (i < nb) bsum[i] = ex + carry;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void ScanDownKernel(const uint32_t* __restrict__ in, size_t n,
                               const uint32_t* __restrict__ bsum,
                               uint32_t* __restrict__ out) {
  const size_t base = size_t(blockIdx.x) * kScanTile + size_t(threadIdx.x) * kScanPer;
  uint32_t v[kScanPer];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    size_t i = base + k;
    v[k] = i < n ? in[i] : 0;
    s += v[k];
  }
  uint32_t ex;
  BlockExclusiveScan(s, ex);
  uint32_t run = ex + bsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    size_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
}

__global__ void RadixHistKernel(const uint32_t* __restrict__ keys, size_t n,
                                int shift, uint32_t* __restrict__ counts,
                                uint32_t nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = size_t(blockIdx.x) * kRsTile;
  for (int r = 0; r < kRsRounds; ++r) {
    size_t i = base + size_t(r) * kRsBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[size_t(threadIdx.x) * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void RadixScatterKernel(const uint32_t* __restrict__ kin,
                                   const uint32_t* __restrict__ vin,
                                   uint32_t* __restrict__ kout,
                                   uint32_t* __restrict__ vout, size_t n,
                                   int shift, const uint32_t* __restrict__ offs,
                                   uint32_t nblocks) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[kRsBlock / 64][256];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  run[t] = offs[size_t(t) * nblocks + blockIdx.x];
  const size_t base = size_t(blockIdx.x) * kRsTile;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int r = 0; r < kRsRounds; ++r) {
#pragma unroll
    for (int w = 0; w < kRsBlock / 64; ++w) wc[w][t] = 0;
    __syncthreads();
    const size_t i = base + size_t(r) * kRsBlock + t;
    const bool act = i < n;
    const uint32_t k = act ? kin[i] : 0u;
    const uint32_t v = act ? vin[i] : 0u;
    const uint32_t d = (k >> shift) & 255u;
    uint64_t peers = __ballot(act);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = __popcll(peers & lt);
    if (act && rank == 0) wc[wave][d] = __popcll(peers);
    __syncthreads();
    uint32_t pos = 0;
    if (act) {
      pos = run[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wc[w][d];
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < kRsBlock / 64; ++w) add += wc[w][t];
    run[t] += add;
    if (act) {
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
  }
}

// off[b] = first sorted position whose key >= b * stride, for b in [0, nb].
__global__ void BucketOffsetsKernel(const uint32_t* __restrict__ keys, size_t n, uint32_t nb,
                                    uint32_t stride, uint32_t* __restrict__ off) {
  const size_t b = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b > nb) return;
  const uint64_t k = uint64_t(b) * stride;
  size_t lo = 0, hi = n;
  while (lo < hi) {
    size_t mid = (lo + hi) >> 1;
    if (uint64_t(keys[mid]) < k) lo = mid + 1; else hi = mid;
  }
  off[b] = uint32_t(lo);
}

}  // namespace

void DeviceExclusiveScan(const uint32_t* in, uint32_t* out, size_t n,
                         uint32_t* scratch, hipStream_t s, uint32_t* total) {
  if (n == 0) {
    if (total) WG_HIP(hipMemsetAsync(total, 0, sizeof(uint32_t), s));
    return;
  }
  const size_t nb = (n + kScanTile - 1) / kScanTile;
  WG_CHECK(nb < (1u << 30), "scan too large");
  ScanReduceKernel<<<dim3(unsigned(nb)), dim3(kScanBlock), 0, s>>>(in, n, scratch);
  ScanBlockSumsKernel<<<1, kScanBlock, 0, s>>>(scratch, uint32_t(nb), total);
  ScanDownKernel<<<dim3(unsigned(nb)), dim3(kScanBlock), 0, s>>>(in, n, scratch, out);
  WG_HIP(hipGetLastError());
}

size_t ScanScratchWords(size_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

size_t RadixSortScratchWords(size_t n) {
  const size_t nblocks = (n + kRsTile - 1) / kRsTile;
  const size_t counts = 256 * std::max<size_t>(nblocks, 1);
  return 2 * counts + ScanScratchWords(counts) + 2 * n;
}

// Sorts (keys, vals) in place by the low key_bits of keys; stable.
void DeviceRadixSort(uint32_t* keys, uint32_t* vals, size_t n, int key_bits,
                     uint32_t* scratch, hipStream_t s) {
  if (n <= 1) return;
  const size_t nblocks = (n + kRsTile - 1) / kRsTile;
  WG_CHECK(nblocks < (1u << 31), "radix sort too large");
  const size_t ncounts = 256 * nblocks;
  uint32_t* counts = scratch;
  uint32_t* offs = counts + ncounts;
  uint32_t* scan_scratch = offs + ncounts;
  uint32_t* kalt = scan_scratch + ScanScratchWords(ncounts);
  uint32_t* valt = kalt + n;
  uint32_t* kin = keys;
  uint32_t* vin = vals;
  uint32_t* ko = kalt;
  uint32_t* vo = valt;
  int passes = 0;
  for (int shift = 0; shift < key_bits; shift += 8) {
    RadixHistKernel<<<dim3(unsigned(nblocks)), dim3(kRsBlock), 0, s>>>(
        kin, n, shift, counts, uint32_t(nblocks));
    DeviceExclusiveScan(counts, offs, ncounts, scan_scratch, s, nullptr);
    RadixScatterKernel<<<dim3(unsigned(nblocks)), dim3(kRsBlock), 0, s>>>(
        kin, vin, ko, vo, n, shift, offs, uint32_t(nblocks));
    std::swap(kin, ko);
    std::swap(vin, vo);
    ++passes;
  }
  WG_HIP(hipGetLastError());
  if (passes & 1) {
    WG_HIP(hipMemcpyAsync(keys, kin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    WG_HIP(hipMemcpyAsync(vals, vin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
}

void DeviceBucketOffsets(const uint32_t* sorted_keys, size_t n, uint32_t nbuckets,
                         uint32_t stride, uint32_t* off, hipStream_t s) {
  const size_t cnt = size_t(nbuckets) + 1;
  BucketOffsetsKernel<<<dim3(unsigned((cnt + 255) / 256)), dim3(256), 0, s>>>(
      sorted_keys, n, nbuckets, stride, off);
  WG_HIP(hipGetLastError());
}

}  // namespace wg
