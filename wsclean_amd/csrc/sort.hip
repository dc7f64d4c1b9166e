// Integer bucket sort of visibilities into (uv tile, first w plane) buckets.
//
// keys are uint32 bucket ids (tile * P0 + p0; nkeys marks skipped zero
// visibilities), values are the flat visibility index row * n_chan + chan.
// The sort is a stable LSD radix sort (digits of up to 9 bits, tiles of
// 8192 elements reordered in LDS before the global writes, runs of 8 tiles
// per block), so the
// permutation is identical to a stable CPU sort (numpy argsort(kind="stable")).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.hpp"
#include "engine.hpp"

namespace wg {

namespace {

// Radix sort tiles: 512 threads x 16 elements; digits of at most 9 bits
// (27-bit keys sort in three passes).
constexpr int kRsThreads = 512;
constexpr int kRsWaves = kRsThreads / 64;
constexpr int kRsPerThread = 16;
constexpr int kRsTile = kRsThreads * kRsPerThread;
constexpr int kRsMaxBits = 9;
constexpr int kRsMaxBuckets = 1 << kRsMaxBits;
static_assert(kRsMaxBuckets <= kRsThreads, "one thread per bucket in the tile scan");
// Each block sorts a run of kRsSub consecutive tiles (its histogram and its
// global offsets cover the run), so there are 8x fewer per-block counts to
// write, scan and read back.
constexpr int kRsSub = 8;
constexpr int kRsRun = kRsTile * kRsSub;

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

// Adds one to h[d] for every active lane, one LDS atomic per run of equal
// digits in consecutive active lanes: the run's first lane adds its length.
// Equal digits in separate runs add separately.  Keys in uv order put long
// runs of equal digits into a wave, which per-lane atomics on one address
// would serialise; finding the runs takes a neighbour exchange and one
// ballot, where grouping all equal digits of a wave takes a ballot per
// digit bit and a 64-bit select per lane and bit.
__device__ __forceinline__ void WaveCountRuns(uint32_t* h, uint32_t d, uint64_t active,
                                              int lane) {
  const uint32_t prev = __shfl_up(d, 1);
  const bool on = (active >> lane) & 1ull;
  const bool prev_on = lane > 0 && ((active >> (lane - 1)) & 1ull);
  const bool head = on && (!prev_on || prev != d);
  // a run ends before the next head or the next inactive lane
  const uint64_t stop = __ballot(head) | ~active;
  if (head) {
    const uint64_t after = lane < 63 ? (stop >> (lane + 1)) | (1ull << (63 - lane)) : 1ull;
    atomicAdd(&h[d], uint32_t(__builtin_ctzll(after)) + 1u);
  }
}

// Digit histogram of one run of kRsSub tiles: counts[d * nblocks + block].
// Wave w reads elements [tile + w * 1024, tile + (w + 1) * 1024) of every
// tile of the run, four keys per 16-byte load in full runs (tiles start
// 32 KiB apart, so the loads are aligned); the ragged last run reads keys
// one by one.
__global__ void __launch_bounds__(kRsThreads) RadixHistKernel(
    const uint32_t* __restrict__ keys, size_t n, int shift, uint32_t mask,
    uint32_t* __restrict__ counts, uint32_t nblocks) {
  __shared__ uint32_t h[kRsWaves][kRsMaxBuckets];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int R = int(mask) + 1;
  for (int i = threadIdx.x; i < kRsWaves * kRsMaxBuckets; i += kRsThreads) (&h[0][0])[i] = 0;
  __syncthreads();
  constexpr int kVec = kRsPerThread / 4;
  const size_t run0 = size_t(blockIdx.x) * kRsRun + size_t(wave) * (kRsTile / kRsWaves);
  auto addr = [&](int sub, int it) {
    return run0 + size_t(sub) * kRsTile + (size_t(it) * 64 + lane) * 4;
  };
  if (size_t(blockIdx.x + 1) * kRsRun <= n) {
    // full run: branch-free, the next tile's keys in flight while the
    // current ones are counted
    uint4 q[kVec], qn[kVec];
#pragma unroll
    for (int it = 0; it < kVec; ++it) q[it] = *reinterpret_cast<const uint4*>(keys + addr(0, it));
    for (int sub = 0; sub < kRsSub; ++sub) {
      if (sub + 1 < kRsSub) {
#pragma unroll
        for (int it = 0; it < kVec; ++it)
          qn[it] = *reinterpret_cast<const uint4*>(keys + addr(sub + 1, it));
      }
#pragma unroll
      for (int it = 0; it < kVec; ++it) {
        WaveCountRuns(h[wave], (q[it].x >> shift) & mask, ~0ull, lane);
        WaveCountRuns(h[wave], (q[it].y >> shift) & mask, ~0ull, lane);
        WaveCountRuns(h[wave], (q[it].z >> shift) & mask, ~0ull, lane);
        WaveCountRuns(h[wave], (q[it].w >> shift) & mask, ~0ull, lane);
      }
#pragma unroll
      for (int it = 0; it < kVec; ++it) q[it] = qn[it];
    }
  } else {
    for (int sub = 0; sub < kRsSub; ++sub) {
      if (addr(sub, 0) - size_t(lane) * 4 >= n) break;
#pragma unroll
      for (int it = 0; it < kVec; ++it) {
        const size_t i = addr(sub, it);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const bool act = i + c < n;
          const uint32_t d = act ? (keys[i + c] >> shift) & mask : 0u;
          WaveCountRuns(h[wave], d, __ballot(act), lane);
        }
      }
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += kRsThreads) {
    uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < kRsWaves; ++w) s += h[w][d];
    counts[size_t(d) * nblocks + blockIdx.x] = s;
  }
}

// Stable scatter of one run of kRsSub tiles, tile by tile in order.  Each
// wave ranks its 1024 elements of a tile by digit (lanes with equal digits
// found by ballots over the digit bits; wave-private running counts in LDS,
// so the order is element order), the block turns the per-wave counts into
// tile positions, reorders the keys and then the values through LDS, and
// writes each digit's run of the tile with consecutive threads to
// consecutive addresses.  run_base[d] (the run's global offset of digit d)
// advances by each tile's digit count.
//
// IDX: the values are the element indices (vin is not read), which saves
// writing and reading the identity permutation in the first pass.
// Shared memory of one scatter block.
struct ScatterLds {
  uint32_t cnt[kRsWaves][kRsMaxBuckets];  // per-wave counts, then offsets
  uint32_t start[kRsMaxBuckets + 1];      // tile-local start of each digit
  uint32_t gbase[kRsMaxBuckets];          // global position - tile position
  uint32_t run_base[kRsMaxBuckets];       // next global position of each digit
  uint32_t buf[kRsTile];                  // keys, then values, in tile order
  uint16_t dig[kRsTile];                  // digit of each tile position
  uint8_t tag[kRsWaves][kRsMaxBuckets];   // lane of a run head, per digit
};

// One tile of the scatter (see RadixScatterKernel).  FULL: the tile holds
// kRsTile elements, so no element needs a bounds test.  BITS: the digit
// width when known at compile time (0: the runtime `bits`).
// Coordinates epilogue of the last pass (SortCoords): the sorted entry p of
// renumbered element e gets its gridding coordinates and shift phasor, and
// its permutation word becomes the caller's flat index with the flip bit
// (and, with co.vrow, its renumbered row is kept).
// The elements of one tile come from 8192 consecutive renumbered elements,
// so their rows are read from a few cache lines.
__device__ __forceinline__ uint32_t SortedCoords(const SortCoords& co, uint32_t e, uint32_t p) {
  const uint32_t nc = uint32_t(co.n_chan);
  const uint32_t r = DivSmall(e, nc, co.inv_chan), ch = e - r * nc;
  const double2 uv = reinterpret_cast<const double2*>(co.uvw4)[2 * size_t(r)];
  const double2 wt = reinterpret_cast<const double2*>(co.uvw4)[2 * size_t(r) + 1];
  const VisLoc L = LocateVis(uv.x, uv.y, wt.x, co.scale[ch], co.dp);
  co.coord[p] = make_float4(L.fu, L.fv, L.fw, __uint_as_float(PackLoc(L)));
  co.phasor[p] = ShiftPhasor(co.dp, L, -1.f);
  if (co.vrow) co.vrow[p] = r;
  return (RowTag(wt) * nc + ch) | (L.flip ? kPermFlip : 0u);
}

template <bool IDX, bool FULL, int BITS, bool COORDS>
__device__ __forceinline__ void ScatterTile(ScatterLds& L, const uint32_t* __restrict__ kin,
                                            const uint32_t* __restrict__ vin,
                                            uint32_t* __restrict__ kout,
                                            uint32_t* __restrict__ vout, uint32_t n,
                                            uint32_t tile, int shift, int bits_rt,
                                            const SortCoords& co) {
  const int bits = BITS > 0 ? BITS : bits_rt;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t mask = (1u << bits) - 1u;
  const int R = 1 << bits;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int i = threadIdx.x; i < kRsWaves * kRsMaxBuckets; i += kRsThreads)
    (&L.cnt[0][0])[i] = 0;
  __syncthreads();
  const uint32_t base = tile + uint32_t(wave) * (kRsTile / kRsWaves);
  auto in = [&](int it) { return FULL || base + uint32_t(it) * 64 + lane < n; };
  uint32_t key[kRsPerThread], pos[kRsPerThread], val[kRsPerThread];
#pragma unroll
  for (int it = 0; it < kRsPerThread; ++it) {
    const uint32_t i = base + uint32_t(it) * 64 + lane;
    key[it] = in(it) ? kin[i] : 0u;
  }
  // the values are loaded now, so their latency overlaps the ranking
#pragma unroll
  for (int it = 0; it < kRsPerThread; ++it) {
    const uint32_t i = base + uint32_t(it) * 64 + lane;
    val[it] = IDX ? i : (in(it) ? vin[i] : 0u);
  }
  // Rank of each element among the wave's equal digits (r), and at the
  // digit's first lane (lead) their number (tot).  Fast path: when every
  // digit of the wave forms one run of consecutive lanes (keys in uv order
  // mostly do), the runs give both: run heads from a neighbour exchange,
  // and a digit with two runs is detected by its heads writing their lane
  // to tag[d] and reading it back.  Otherwise the equal digits are found by
  // one ballot per digit bit.
#pragma unroll
  for (int it = 0; it < kRsPerThread; ++it) {
    const bool act = in(it);
    const uint32_t d = (key[it] >> shift) & mask;
    const uint64_t active = FULL ? ~0ull : __ballot(act);
    const uint32_t prev = __shfl_up(d, 1);
    const bool prev_on = lane > 0 && ((active >> (lane - 1)) & 1ull);
    const bool head = act && (!prev_on || prev != d);
    const uint64_t heads = __ballot(head);
    if (head) L.tag[wave][d] = uint8_t(lane);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bool clash = head && L.tag[wave][d] != uint8_t(lane);
    uint32_t r, tot;
    bool lead;
    if (__ballot(clash) == 0ull) {
      // this lane's run starts at the highest head at or below it
      const uint64_t upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
      r = uint32_t(lane - (63 - __builtin_clzll(upto | 1ull)));
      // a run ends before the next head or the next inactive lane
      const uint64_t stop = heads | ~active;
      const uint64_t after = lane < 63 ? (stop >> (lane + 1)) | (1ull << (63 - lane)) : 1ull;
      tot = uint32_t(__builtin_ctzll(after)) + 1u;
      lead = head;
    } else {
      uint64_t peers = active;
      auto narrow = [&](int b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
      };
      if constexpr (BITS > 0) {
#pragma unroll
        for (int b = 0; b < BITS; ++b) narrow(b);
      } else {
        for (int b = 0; b < bits; ++b) narrow(b);
      }
      r = uint32_t(__popcll(peers & lt));
      tot = uint32_t(__popcll(peers));
      lead = r == 0;
    }
    const uint32_t c = act ? L.cnt[wave][d] : 0u;
    pos[it] = c + r;
    if (act && lead) L.cnt[wave][d] = c + tot;
  }
  __syncthreads();
  // wave offsets inside each digit, digit totals, then tile starts
  for (int d = threadIdx.x; d < R; d += kRsThreads) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kRsWaves; ++w) {
      const uint32_t t = L.cnt[w][d];
      L.cnt[w][d] = run;
      run += t;
    }
    L.start[d] = run;
  }
  __syncthreads();
  {
    const int d = threadIdx.x;
    const uint32_t v = d < R ? L.start[d] : 0u;
    uint32_t ex = 0;
    const uint32_t total = BlockExclusiveScan(v, ex);
    if (d < R) {
      L.start[d] = ex;
      L.gbase[d] = L.run_base[d] - ex;
    }
    if (d == 0) L.start[R] = total;
  }
  __syncthreads();
  const int count = FULL ? kRsTile : int(min(uint32_t(kRsTile), n - tile));
#pragma unroll
  for (int it = 0; it < kRsPerThread; ++it) {
    if (in(it)) {
      const uint32_t d = (key[it] >> shift) & mask;
      const uint32_t p = L.start[d] + L.cnt[wave][d] + pos[it];
      pos[it] = p;
      L.buf[p] = key[it];
      L.dig[p] = uint16_t(d);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < count; i += kRsThreads) kout[L.gbase[L.dig[i]] + i] = L.buf[i];
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kRsPerThread; ++it)
    if (in(it)) L.buf[pos[it]] = val[it];
  __syncthreads();
  if constexpr (COORDS) {
    for (int i = threadIdx.x; i < count; i += kRsThreads) {
      const uint32_t p = L.gbase[L.dig[i]] + i;
      vout[p] = SortedCoords(co, L.buf[i], p);
    }
  } else {
    for (int i = threadIdx.x; i < count; i += kRsThreads) vout[L.gbase[L.dig[i]] + i] = L.buf[i];
  }
  // the next tile's digits follow this tile's
  for (int d = threadIdx.x; d < R; d += kRsThreads) L.run_base[d] += L.start[d + 1] - L.start[d];
  __syncthreads();
}

// Two blocks per CU (LDS) need four waves per SIMD, i.e. at most 128 VGPRs.
template <bool IDX, int BITS, bool COORDS>
__global__ void __launch_bounds__(kRsThreads)
    __attribute__((amdgpu_waves_per_eu(4, 4))) RadixScatterKernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, int shift,
    int bits, const uint32_t* __restrict__ offs, uint32_t nblocks, SortCoords co) {
  __shared__ ScatterLds L;
  const int R = 1 << (BITS > 0 ? BITS : bits);
  for (int d = threadIdx.x; d < R; d += kRsThreads)
    L.run_base[d] = offs[size_t(d) * nblocks + blockIdx.x];
  for (int sub = 0; sub < kRsSub; ++sub) {
    const uint32_t tile = (blockIdx.x * kRsSub + sub) * kRsTile;
    if (tile >= n) break;  // uniform over the block
    if (n - tile >= uint32_t(kRsTile))
      ScatterTile<IDX, true, BITS, COORDS>(L, kin, vin, kout, vout, n, tile, shift, bits, co);
    else
      ScatterTile<IDX, false, BITS, COORDS>(L, kin, vin, kout, vout, n, tile, shift, bits, co);
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
  const size_t nblocks = (n + kRsRun - 1) / kRsRun;
  const size_t counts = size_t(kRsMaxBuckets) * std::max<size_t>(nblocks, 1);
  return 2 * counts + ScanScratchWords(counts);
}

// Sorts (keys, vals) by the low key_bits of keys; stable.  index_vals: vals
// is not read, the values are the element indices 0 .. n - 1.  The digit width
// is ceil(key_bits / passes) with passes = ceil(key_bits / 9).  The passes
// alternate between (keys, vals) and (keys_alt, vals_alt); the return value
// says which pair holds the result (true: the alt pair).
bool DeviceRadixSort(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                     size_t n, int key_bits, uint32_t* scratch, hipStream_t s,
                     bool index_vals, const SortCoords* coords) {
  WG_CHECK(!coords || (index_vals && n < 0x80000000ull),
           "the coordinates epilogue sorts renumbered element indices");
  if (n <= 1 && !coords) {
    if (n == 1 && index_vals) LaunchFillZero(vals, sizeof(uint32_t), s);
    return false;
  }
  WG_CHECK(key_bits >= 1 && key_bits <= 32, "key width out of range");
  WG_CHECK(((reinterpret_cast<uintptr_t>(keys) | reinterpret_cast<uintptr_t>(keys_alt)) & 15) == 0,
           "radix sort keys must be 16-byte aligned");
  const size_t nblocks = (n + kRsRun - 1) / kRsRun;
  // 32-bit element indices inside the scatter kernel
  WG_CHECK(n <= 0xFFFFFFFFull - kRsRun, "radix sort too large");
  if (n == 0) return false;
  const int npass = (key_bits + kRsMaxBits - 1) / kRsMaxBits;
  const int digit = (key_bits + npass - 1) / npass;
  uint32_t* counts = scratch;
  uint32_t* offs = counts + size_t(kRsMaxBuckets) * nblocks;
  uint32_t* scan_scratch = offs + size_t(kRsMaxBuckets) * nblocks;
  uint32_t* kin = keys;
  uint32_t* vin = vals;
  uint32_t* ko = keys_alt;
  uint32_t* vo = vals_alt;
  bool in_alt = false;
  for (int shift = 0; shift < key_bits; shift += digit) {
    const int bits = std::min(digit, key_bits - shift);
    const size_t ncounts = (size_t(1) << bits) * nblocks;
    RadixHistKernel<<<dim3(unsigned(nblocks)), dim3(kRsThreads), 0, s>>>(
        kin, n, shift, (1u << bits) - 1u, counts, uint32_t(nblocks));
    DeviceExclusiveScan(counts, offs, ncounts, scan_scratch, s, nullptr);
    const bool idx = index_vals && shift == 0;
    const bool last = shift + digit >= key_bits;
    const SortCoords co = coords ? *coords : SortCoords{};
    auto launch = [&](auto kernel) {
      kernel<<<dim3(unsigned(nblocks)), dim3(kRsThreads), 0, s>>>(
          kin, idx ? nullptr : vin, ko, vo, uint32_t(n), shift, bits, offs, uint32_t(nblocks),
          co);
    };
    auto pick = [&](auto full_bits) {
      constexpr int B = decltype(full_bits)::value;
      if (coords && last)
        idx ? launch(RadixScatterKernel<true, B, true>) : launch(RadixScatterKernel<false, B, true>);
      else
        idx ? launch(RadixScatterKernel<true, B, false>)
            : launch(RadixScatterKernel<false, B, false>);
    };
    if (bits == kRsMaxBits)
      pick(std::integral_constant<int, kRsMaxBits>());
    else
      pick(std::integral_constant<int, 0>());
    std::swap(kin, ko);
    std::swap(vin, vo);
    in_alt = !in_alt;
  }
  WG_HIP(hipGetLastError());
  return in_alt;
}

void DeviceBucketOffsets(const uint32_t* sorted_keys, size_t n, uint32_t nbuckets,
                         uint32_t stride, uint32_t* off, hipStream_t s) {
  const size_t cnt = size_t(nbuckets) + 1;
  BucketOffsetsKernel<<<dim3(unsigned((cnt + 255) / 256)), dim3(256), 0, s>>>(
      sorted_keys, n, nbuckets, stride, off);
  WG_HIP(hipGetLastError());
}

}  // namespace wg
