// Binning: prefix scan, stable LSD radix sort, instance expansion and tile ranges.
//
// Ordering contract (reference rasterizer_impl.cu:94-167, :622-679): the
// instances of every tile must come out ordered by (depth bits, Gaussian id),
// tiles ascending -- the order of a stable sort of key = tile<<32 | depth
// bits over instances emitted in Gaussian order.  This library reaches the
// same order in two cheaper steps:
//   1. a stable 32-bit radix sort of the P Gaussians by depth bits
//      (ties keep Gaussian-id order),
//   2. instance expansion in that depth order, then a stable radix sort of
//      the R instances by tile id over only getHigherMsb(tiles) bits.
// Stability of step 2 keeps each tile's instances in step-1 order, which is
// exactly (depth bits, id).  The R-sized sort shrinks from 32+13 key bits
// (6 passes of 12-byte pairs) to 13 bits (2 passes of 8-byte pairs).
#include "kernels.hpp"

namespace omnigs {

__device__ __forceinline__ uint32_t load_in(const uint32_t* __restrict__ in, const uint32_t* __restrict__ idx,
                                            size_t i) {
  return idx ? in[idx[i]] : in[i];
}

// Exclusive scan of a short array by one block, kScanTile values at a time:
// the chunk is loaded into LDS coalesced, thread t scans its kScanItems
// consecutive values, the block scans the run sums (plus the carry of earlier
// chunks), and the chunk leaves coalesced.
__device__ void block_scan_array(const uint32_t* __restrict__ in, size_t n, uint32_t* __restrict__ out) {
  __shared__ uint32_t s_v[kScanTile];
  uint32_t carry = 0;
  for (size_t c0 = 0; c0 < n; c0 += kScanTile) {
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads) s_v[k] = c0 + k < n ? in[c0 + k] : 0u;
    __syncthreads();
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) run += s_v[threadIdx.x * kScanItems + k];
    uint32_t tot;
    uint32_t pre = carry + block_exclusive_scan(run, &tot);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      const uint32_t v = s_v[threadIdx.x * kScanItems + k];
      s_v[threadIdx.x * kScanItems + k] = pre;
      pre += v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads)
      if (c0 + k < n) out[c0 + k] = s_v[k];
    carry += tot;
    __syncthreads();
  }
}

// Block sums; with a gather index the gathered values are also written to
// `gathered` so that the downsweep reads them contiguously.  One block past
// the last (launched when side_in is set) scans the unrelated short array
// side_in -> side_out (the per-block gradient slot counts) while the others
// reduce, so its serial chunks cost no launch and no time of their own.
// kRect: the values are the areas of the expansion records rect_in,
// gathered through idx; each record's rect goes to rect_out[i].
template <bool kRect>
__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in,
                                                                   const uint32_t* __restrict__ idx, size_t n,
                                                                   uint32_t* __restrict__ partials,
                                                                   uint32_t* __restrict__ gathered,
                                                                   const uint32_t* __restrict__ side_in, size_t side_n,
                                                                   uint32_t* __restrict__ side_out,
                                                                   const uint2* __restrict__ rect_in,
                                                                   uint2* __restrict__ rect_out) {
  if (side_in && blockIdx.x == gridDim.x - 1) {
    block_scan_array(side_in, side_n, side_out);
    return;
  }
  const size_t base = (size_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t i = base + (size_t)k * kScanThreads + threadIdx.x;
    if (i < n) {
      uint32_t v;
      if constexpr (kRect) {
        const uint2 r = rect_in[idx[i]];
        v = dup_rect_area(r);
        rect_out[i] = dup_rect_xy(r);
      } else {
        v = load_in(in, idx, i);
      }
      if (gathered) gathered[i] = v;
      s += v;
    }
  }
  uint32_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Exclusive scan of the block partials in place, one block looping over
// chunks.  Their total goes to partials[nb] and an overflow flag to
// partials[nb + 1]: the scan is 32-bit, so the total is also summed in 64
// bits, and a total past 2^32 - 1 leaves {0, 1} there (no usable instance,
// overflow) instead of a wrapped count.  A single block's partial cannot
// wrap: its kScanTile values are each at most kMaxTiles.
__device__ void scan_partials_block(uint32_t* __restrict__ partials, size_t nb) {
  __shared__ unsigned long long s_total;
  if (threadIdx.x == 0) s_total = 0ull;
  __syncthreads();
  uint32_t carry = 0;
  for (size_t c = 0; c < nb; c += kScanThreads) {
    const size_t i = c + threadIdx.x;
    const uint32_t v = i < nb ? partials[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(v, &tot);
    if (i < nb) partials[i] = carry + ex;
    carry += tot;
    unsigned long long w = v;
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) w += __shfl_xor(w, o);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(&s_total, w);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const bool ovf = s_total > 0xFFFFFFFFull;
    partials[nb] = ovf ? 0u : carry;
    partials[nb + 1] = ovf ? 1u : 0u;
  }
}

__global__ __launch_bounds__(kScanThreads) void scan_partials_kernel(uint32_t* __restrict__ partials, size_t nb) {
  scan_partials_block(partials, nb);
}

// Each thread scans kScanItems consecutive elements (blocked layout); the
// tile enters and leaves through LDS (rows padded to kScanItems + 1 words) so
// that the global loads and stores stay coalesced.
__global__ __launch_bounds__(kScanThreads) void scan_downsweep_kernel(const uint32_t* __restrict__ in,
                                                                      const uint32_t* __restrict__ idx, size_t n,
                                                                      const uint32_t* __restrict__ partials,
                                                                      uint32_t* __restrict__ out, int exclusive) {
  constexpr int kRow = kScanItems + 1;
  __shared__ uint32_t s_v[kScanThreads * kRow];
  const size_t base = (size_t)blockIdx.x * kScanTile;
  const int tid = threadIdx.x;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int il = k * kScanThreads + tid;
    const size_t i = base + (size_t)il;
    s_v[(il / kScanItems) * kRow + il % kScanItems] = i < n ? load_in(in, idx, i) : 0u;
  }
  __syncthreads();
  uint32_t* row = s_v + tid * kRow;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = row[k];
    sum += v[k];
  }
  uint32_t tot;
  uint32_t run = block_exclusive_scan(sum, &tot) + partials[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t inc = run + v[k];
    row[k] = exclusive ? run : inc;
    run = inc;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int il = k * kScanThreads + tid;
    const size_t i = base + (size_t)il;
    if (i < n) out[i] = s_v[(il / kScanItems) * kRow + il % kScanItems];
  }
}

hipError_t launch_scan(const uint32_t* in, const uint32_t* idx, uint32_t* out, size_t n, bool exclusive,
                       uint32_t* partials, hipStream_t s, uint32_t* gather_buf, const uint32_t* side_in,
                       size_t side_n, uint32_t* side_out, const uint2* rect_in, uint2* rect_out) {
  if (n == 0) return hipSuccess;
  const size_t nb = scan_blocks(n);
  uint32_t* gathered = idx ? gather_buf : nullptr;
  const bool side = side_in && side_n;
  const dim3 grid(nb + (side ? 1 : 0));
  if (rect_in && idx && gathered) {
    hipLaunchKernelGGL(scan_reduce_kernel<true>, grid, dim3(kScanThreads), 0, s, in, idx, n, partials, gathered,
                       side ? side_in : nullptr, side_n, side_out, rect_in, rect_out);
  } else {
    hipLaunchKernelGGL(scan_reduce_kernel<false>, grid, dim3(kScanThreads), 0, s, in, idx, n, partials, gathered,
                       side ? side_in : nullptr, side_n, side_out, (const uint2*)nullptr, (uint2*)nullptr);
  }
  hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanThreads), 0, s, partials, nb);
  const uint32_t* din = gathered ? gathered : in;
  const uint32_t* didx = gathered ? nullptr : idx;
  hipLaunchKernelGGL(scan_downsweep_kernel, dim3(nb), dim3(kScanThreads), 0, s, din, didx, n,
                     (const uint32_t*)partials, out, exclusive ? 1 : 0);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.  Per pass three kernels:
//   count   -- digit histogram of every block's tile of keys, digit-major table;
//   rowscan -- one block per digit: exclusive scan of that digit's row over the
//              blocks, and the digit's total;
//   scatter -- every block turns the 256 digit totals into digit bases (LDS
//              scan), then ranks its keys in input order with 64-lane ballots
//              and writes each to base[d] + row prefix + rank (stable).
// counts holds kRadix * blocks row entries followed by kRadix digit totals.
// ---------------------------------------------------------------------------
// Which sort tile block b of a grid of nb takes.  Blocks are dispatched
// round-robin over the 8 XCDs, each with its own L2, and the digit-count
// table holds one word per tile in each digit's row: with tile = b, the 32
// words of a 128-byte line come from 32 blocks on all eight XCDs, and every
// line is written back (and read again) eight times over.  Here each run of
// 32 consecutive tiles goes to blocks on one XCD; the runs of a group of 256
// tiles alternate over the XCDs, so live tiles stay spread over the chip when
// the grid is sized for a capacity.  A partial last group keeps tile = b.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
  constexpr uint32_t kXcds = 8, kRun = 32, kGroup = kXcds * kRun;
  const uint32_t g = b / kGroup;
  if ((g + 1) * kGroup > nb) return b;
  const uint32_t r = b % kGroup;
  return g * kGroup + (r % kXcds) * kRun + r / kXcds;
}

// n: element count, or the capacity when n_dev holds the count on the device
// (min(*n_dev, n) elements are sorted; blocks past it count nothing).
__device__ __forceinline__ size_t live_count(size_t n, const uint32_t* n_dev) {
  return n_dev ? min((size_t)*n_dev, n) : n;
}

__global__ __launch_bounds__(kSortThreads) void radix_count_kernel(const uint32_t* __restrict__ keys, size_t n,
                                                                   const uint32_t* __restrict__ n_dev, int shift,
                                                                   uint32_t mask, uint32_t* __restrict__ counts,
                                                                   uint32_t nblocks) {
  __shared__ uint32_t hist[kRadix];
  n = live_count(n, n_dev);
  for (int d = threadIdx.x; d < kRadix; d += kSortThreads) hist[d] = 0;
  __syncthreads();
  const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
  const size_t base = (size_t)t * kSortTile;
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const size_t i = base + (size_t)k * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&hist[(keys[i] >> shift) & mask], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadix; d += kSortThreads) counts[(size_t)d * nblocks + t] = hist[d];
}

// One block per digit.  A round covers kRowItems consecutive entries per
// thread (one 64-byte line each), summed serially around a single block scan,
// so that a row of up to 4096 tiles takes one round instead of one per 256.
constexpr int kRowItems = 16;
__global__ __launch_bounds__(kSortThreads) void radix_rowscan_kernel(uint32_t* __restrict__ counts,
                                                                     uint32_t nblocks) {
  const uint32_t d = blockIdx.x;
  uint32_t* row = counts + (size_t)d * nblocks;
  uint32_t carry = 0;
  for (uint32_t c = 0; c < nblocks; c += kSortThreads * kRowItems) {
    const uint32_t i0 = c + threadIdx.x * kRowItems;
    uint32_t v[kRowItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kRowItems; ++k) {
      v[k] = i0 + k < nblocks ? row[i0 + k] : 0u;
      sum += v[k];
    }
    uint32_t tot;
    uint32_t run = carry + block_exclusive_scan(sum, &tot);
#pragma unroll
    for (int k = 0; k < kRowItems; ++k) {
      if (i0 + k < nblocks) row[i0 + k] = run;
      run += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) counts[(size_t)kRadix * nblocks + d] = carry;
}

// Scatter of one pass.  Wave w of a block owns keys [512 w, 512 w + 512) of
// the block's tile, lane l holding keys 512 w + 64 k + l; ranking item by item
// with per-wave digit counters in LDS (no block barrier inside the ranking)
// keeps the keys' input order inside each digit, which makes the sort stable.
// The keys are then grouped by digit in LDS and written run by run
// (consecutive threads, consecutive addresses).
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, size_t n, const uint32_t* __restrict__ n_dev, int shift, int bits,
    const uint32_t* __restrict__ counts, uint32_t nblocks) {
  constexpr int kWaves = kSortThreads / kWave;
  constexpr int kPerWave = kSortItems * kWave;
  static_assert(kRadix == kSortThreads, "one digit per thread");
  __shared__ uint32_t s_key[kSortTile];
  __shared__ uint32_t s_val[kSortTile];
  __shared__ uint32_t wcnt[kWaves][kRadix];  // per-wave digit counts, then per-wave offsets inside the digit
  __shared__ uint32_t s_local[kRadix];       // first slot of digit d in the tile's digit-grouped order
  __shared__ int32_t s_out[kRadix];          // global position of the key in slot i: s_out[d] + i
  const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
  const uint32_t mask = (1u << bits) - 1u;
  const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
  const size_t base = (size_t)t * kSortTile;
  n = live_count(n, n_dev);
  if (base >= n) return;  // uniform across the block
  const uint32_t nloc = (uint32_t)min((size_t)kSortTile, n - base);
#pragma unroll
  for (int ww = 0; ww < kWaves; ++ww) wcnt[ww][tid] = 0;

  uint32_t keys[kSortItems], vals[kSortItems], rank[kSortItems];
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const uint32_t li = (uint32_t)(w * kPerWave + k * kWave + lane);
    const size_t i = base + li;
    keys[k] = li < nloc ? kin[i] : 0u;
    vals[k] = li < nloc ? (vin ? vin[i] : (uint32_t)i) : 0u;
  }
  __syncthreads();
  const uint64_t lanes_below = (1ull << lane) - 1ull;
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const bool valid = (uint32_t)(w * kPerWave + k * kWave + lane) < nloc;
    const uint32_t d = (keys[k] >> shift) & mask;
    uint64_t same = __ballot(valid);  // lanes holding a key with the same digit
    for (int b = 0; b < bits; ++b) {
      const bool set = (d >> b) & 1u;
      const uint64_t bb = __ballot(set);
      same &= set ? bb : ~bb;
    }
    const uint32_t before = wcnt[w][d];
    const uint32_t r = (uint32_t)__popcll(same & lanes_below);
    rank[k] = before + r;
    if (valid && r == 0) wcnt[w][d] = before + (uint32_t)__popcll(same);
  }
  __syncthreads();
  // thread d: the waves' offsets inside digit d, the digit's first slot in
  // the tile, and where the digit's run of this tile goes globally
  const uint32_t d = (uint32_t)tid;
  uint32_t here = 0;
#pragma unroll
  for (int ww = 0; ww < kWaves; ++ww) {
    const uint32_t c = wcnt[ww][d];
    wcnt[ww][d] = here;
    here += c;
  }
  uint32_t tot;
  const uint32_t local = block_exclusive_scan(here, &tot);
  const uint32_t digit_base = block_exclusive_scan(counts[(size_t)kRadix * nblocks + d], &tot);
  s_local[d] = local;
  s_out[d] = (int32_t)(digit_base + counts[(size_t)d * nblocks + t]) - (int32_t)local;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    if ((uint32_t)(w * kPerWave + k * kWave + lane) < nloc) {
      const uint32_t dk = (keys[k] >> shift) & mask;
      const uint32_t slot = s_local[dk] + wcnt[w][dk] + rank[k];
      s_key[slot] = keys[k];
      s_val[slot] = vals[k];
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < nloc; i += kSortThreads) {
    const uint32_t key = s_key[i];
    const size_t pos = (size_t)((int64_t)s_out[(key >> shift) & mask] + (int64_t)i);
    kout[pos] = key;
    vout[pos] = s_val[i];
  }
}

uint32_t radix_first_mask(int bits) {
  if (bits < 1) bits = 1;
  const int passes = radix_passes(bits);
  const int width = (bits + passes - 1) / passes;
  return (1u << (bits < width ? bits : width)) - 1u;
}

hipError_t launch_radix_sort(uint32_t* keys, uint32_t* keys_alt, const uint32_t* vals_in, uint32_t* vals,
                             uint32_t* vals_alt, size_t n, int bits, uint32_t* counts, uint32_t* /*partials*/,
                             hipStream_t s, SortResult* out, const uint32_t* n_dev, bool first_counted) {
  // Pass p reads (src) and writes (dst); the first pass reads vals_in (or
  // indices) and writes vals_alt, so an odd pass count ends in the *_alt arrays.
  uint32_t* ksrc = keys;
  uint32_t* kdst = keys_alt;
  const uint32_t* vsrc = vals_in;
  uint32_t* vdst = vals_alt;
  uint32_t* vother = vals;
  if (n == 0) {
    out->keys = keys;
    out->vals = vals;
    return hipSuccess;
  }
  if (bits < 1) bits = 1;
  const uint32_t nb = (uint32_t)sort_blocks(n);
  // the passes split the key bits evenly (13 bits: 7 + 6), so that the
  // first pass's digit runs are longer than with 8 + 5
  const int passes = radix_passes(bits);
  const int width = (bits + passes - 1) / passes;
  for (int shift = 0; shift < bits; shift += width) {
    const int pb = bits - shift < width ? bits - shift : width;
    const uint32_t mask = (1u << pb) - 1u;
    if (!(first_counted && shift == 0))
      hipLaunchKernelGGL(radix_count_kernel, dim3(nb), dim3(kSortThreads), 0, s, (const uint32_t*)ksrc, n, n_dev,
                         shift, mask, counts, nb);
    hipLaunchKernelGGL(radix_rowscan_kernel, dim3(kRadix), dim3(kSortThreads), 0, s, counts, nb);
    hipLaunchKernelGGL(radix_scatter_kernel, dim3(nb), dim3(kSortThreads), 0, s, (const uint32_t*)ksrc, vsrc, kdst,
                       vdst, n, n_dev, shift, pb, (const uint32_t*)counts, nb);
    uint32_t* kt = ksrc;
    ksrc = kdst;
    kdst = kt;
    vsrc = vdst;
    uint32_t* vt = vdst;
    vdst = vother;
    vother = vt;
  }
  out->keys = ksrc;
  out->vals = const_cast<uint32_t*>(vsrc);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Instance expansion in depth order, balanced over instances: block b owns
// instances [kDupChunk b, kDupChunk (b+1)).  Thread 0 finds the first and last
// depth-ordered Gaussian owning them (binary search over the inclusive
// offsets); their offsets and rects are staged in LDS (every Gaussian in that
// span covers >= 1 tile, so the span is <= kDupChunk long) and each instance
// finds its owner by binary search in LDS.  Near-camera Gaussians covering
// thousands of tiles therefore spread over many blocks instead of serialising
// one.  Tiles of one Gaussian are emitted row-major over its rect, as the
// reference's duplicateWithKeys does (rasterizer_impl.cu:127-138).
// ---------------------------------------------------------------------------
constexpr int kDupThreads = 256;
constexpr int kDupRun = 8;  // consecutive instances per thread
constexpr int kDupChunk = kDupThreads * kDupRun;

__device__ __forceinline__ uint32_t first_greater(const uint32_t* __restrict__ a, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n;  // first i with a[i] > v (a non-decreasing)
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] > v) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// first_greater for two values at once, by the whole 256-thread block: each
// half runs a 128-ary search (threads [0, 128) for v[0], [128, 256) for
// v[1]), one load per thread and round, so that 1.5 M entries take three
// rounds instead of 21 dependent loads by one thread.  a is non-decreasing;
// out[h] = the first i < n with a[i] > v[h], or n.  Every thread calls it.
__device__ void block_first_greater2(const uint32_t* __restrict__ a, uint32_t n, uint32_t v0, uint32_t v1,
                                     uint32_t* out) {
  constexpr uint32_t kWays = kDupThreads / 2;
  __shared__ uint32_t s_lo[2], s_hi[2], s_below[2];
  const int tid = threadIdx.x, h = tid / (int)kWays;
  const uint32_t j = (uint32_t)tid % kWays, v = h ? v1 : v0;
  if (tid < 2) {
    s_lo[tid] = 0;
    s_hi[tid] = n;  // the answer lies in [lo, hi]
  }
  __syncthreads();
  while (s_lo[0] < s_hi[0] || s_lo[1] < s_hi[1]) {  // block-uniform: shared values after a barrier
    const uint32_t lo = s_lo[h], hi = s_hi[h];
    const uint32_t stride = (hi - lo + kWays - 1) / kWays;  // 0 once this half is done
    const uint32_t q = lo + j * stride;
    // probes at or below v form a prefix of the half's 128 probes
    const bool below = stride > 0 && q < hi && a[q] <= v;
    const uint64_t bal = __ballot(below);
    __syncthreads();  // every thread has read lo / hi
    if (tid < 2) s_below[tid] = 0;
    __syncthreads();
    if ((tid & (kWave - 1)) == 0 && bal) atomicAdd(&s_below[h], (uint32_t)__popcll(bal));
    __syncthreads();
    if (j == 0 && stride > 0) {
      const uint32_t nb = s_below[h];  // probes 0 .. nb-1 are <= v
      s_lo[h] = nb == 0 ? lo : lo + (nb - 1) * stride + 1;
      s_hi[h] = min(hi, lo + nb * stride);
    }
    __syncthreads();
  }
  if (tid < 2) out[tid] = s_lo[tid];
  __syncthreads();
}

// kMulti: items of a batch (item i = Gaussian i % vt.P in view i / vt.P), each
// keyed by its view's global tile id.  R_dev: instance count on the device
// (clamped to R, the capacity); status[0] receives the count needed and
// status[1] = 1 when it does not fit.
template <bool kMulti>
__global__ __launch_bounds__(kDupThreads) void duplicate_kernel(int P, uint32_t R, const uint32_t* __restrict__ R_dev,
                                                                uint32_t* __restrict__ status,
                                                                const uint32_t* __restrict__ order,
                                                                const uint32_t* __restrict__ offsets,
                                                                const uint2* __restrict__ rect_depth, TileGrid grid,
                                                                ViewTable vt, uint32_t* __restrict__ tile_key,
                                                                uint32_t* __restrict__ gid,
                                                                uint32_t* __restrict__ no_record,
                                                                uint32_t* __restrict__ counts, uint32_t count_mask,
                                                                uint32_t nblocks, uint32_t* __restrict__ zero,
                                                                uint32_t zero_n, uint32_t* __restrict__ blend_ok,
                                                                const uint32_t* __restrict__ P_dev) {
  static_assert(kDupChunk == kSortTile, "expansion blocks are the tile sort's tiles");
  constexpr int kViewSlots = kMulti ? kDupChunk : 1;
  __shared__ uint32_t s_hist[kRadix];
  __shared__ uint32_t s_end[kDupChunk];
  __shared__ uint32_t s_gid[kDupChunk];
  __shared__ int16_t s_x0[kDupChunk];  // |x0| <= gx + 1 <= kMaxGridCols + 1 fits 16 bits
  __shared__ uint16_t s_y0[kDupChunk], s_w[kDupChunk];
  __shared__ uint8_t s_view[kViewSlots];
  __shared__ uint32_t s_span[2];
  const int tid = threadIdx.x;
  // the tile ranges, cleared here for the ranges pass (which sets the
  // occupied tiles only) instead of by a launch of their own
  for (uint32_t i = blockIdx.x * kDupThreads + tid; i < zero_n; i += gridDim.x * kDupThreads) zero[i] = 0u;
  // the forward's per-wave blend bits are stale until a render forward of
  // this binning writes them
  if (blend_ok && blockIdx.x == 0 && tid == 0) *blend_ok = 0u;
  if (R_dev) {
    // R_dev = {count, 32-bit overflow flag} from the offset scan's partials
    const uint32_t want = R_dev[0];
    const bool wrapped = R_dev[1] != 0u;
    if (status && blockIdx.x == 0 && tid == 0) {
      status[0] = wrapped ? 0xFFFFFFFFu : want;
      status[1] = (wrapped || want > R) ? 1u : 0u;
    }
    R = min(R, want);
  }
  // the sort tile this block expands (xcd_tile: its count column shares
  // lines with tiles on the same XCD)
  const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t k0 = t * (uint32_t)kDupChunk;
  if (k0 >= R) {  // past the live instances (capacity-sized grids): an empty sort tile
    if (counts)
      for (int d = tid; d < kRadix; d += kDupThreads) counts[(size_t)d * nblocks + t] = 0u;
    return;
  }
  const uint32_t k1 = min(R, k0 + (uint32_t)kDupChunk);
  for (int d = tid; d < kRadix; d += kDupThreads) s_hist[d] = 0u;
  // P_dev: the owner count on the device (a band's compacted list)
  const uint32_t Pn = P_dev ? min((uint32_t)P, *P_dev) : (uint32_t)P;
  block_first_greater2(offsets, Pn, k0, k1 - 1, s_span);
  const uint32_t sb = s_span[0];
  // <= kDupChunk: every owner in the span holds an instance (the clamp only
  // keeps the LDS staging in bounds should a caller break that)
  const uint32_t ns = min(s_span[1] - sb + 1, (uint32_t)kDupChunk);
  const uint32_t begin0 = sb == 0 ? 0u : offsets[sb - 1];
  for (uint32_t j = tid; j < ns; j += kDupThreads) {
    const uint32_t s = sb + j;
    const uint32_t gg = order[s];
    s_end[j] = offsets[s];
    s_gid[j] = gg;
    if constexpr (kMulti) s_view[j] = (uint8_t)(gg / (uint32_t)vt.P);
    // the rect in depth order, as the offset scan gathered it
    const uint2 rd = rect_depth[s];
    s_x0[j] = (int16_t)(int)rd.x;
    s_y0[j] = (uint16_t)(rd.y & 0xFFFFu);
    s_w[j] = (uint16_t)(rd.y >> 16);
  }
  __syncthreads();
  // Each thread expands kDupRun consecutive instances: one search finds the
  // owner of the first, then owners and rect positions (row q, column r)
  // advance one instance at a time; a Gaussian's first instance is its rect's
  // (0, 0), and owners with no instance are stepped over.
  const uint32_t kb = k0 + (uint32_t)tid * kDupRun;
  uint32_t keys[kDupRun], ids[kDupRun];
#pragma unroll
  for (int u = 0; u < kDupRun; ++u) keys[u] = ids[u] = 0u;
  if (kb < k1) {
    uint32_t j = first_greater(s_end, ns, kb);
    uint32_t w = s_w[j];
    const uint32_t local = kb - (j == 0 ? begin0 : s_end[j - 1]);
    uint32_t q = local / w, r = local - q * w;
#pragma unroll
    for (int u = 0; u < kDupRun; ++u) {
      const uint32_t k = kb + (uint32_t)u;
      if (k >= k1) continue;
      while (k >= s_end[j]) {  // stops at the span's last owner: s_end[ns - 1] >= k1
        ++j;
        w = s_w[j];
        q = r = 0;
      }
      const uint32_t y = (uint32_t)s_y0[j] + q;
      if constexpr (kMulti) {
        const int v = s_view[j];
        const TileGrid gv = vt.grid(v);
        keys[u] = vt.key_base(v) + y * (uint32_t)gv.gx + (uint32_t)wrap_col(s_x0[j] + (int)r, gv);
      } else {
        keys[u] = y * (uint32_t)grid.gx + (uint32_t)wrap_col(s_x0[j] + (int)r, grid);
      }
      ids[u] = s_gid[j];
      if (++r == w) {
        r = 0;
        ++q;
      }
    }
  }
  // The tile sort's first digit, counted here for its first pass (this block
  // is that pass's tile), in place of a count kernel re-reading the keys.
  if (counts) {
#pragma unroll
    for (int u = 0; u < kDupRun; ++u)
      if (kb + (uint32_t)u < k1) atomicAdd(&s_hist[keys[u] & count_mask], 1u);
    __syncthreads();
    for (int d = tid; d < kRadix; d += kDupThreads) counts[(size_t)d * nblocks + t] = s_hist[d];
  }
  if (kb >= k1) return;
  // The gradient slots number as many as the instances: this pass also marks
  // every one as holding no record (kNoRecord in its opacity word) for the
  // backward, which overwrites the slots it stages.  no_record is NULL in a
  // binning without gradient records (a forward whose backward bins again).
  if (kb + kDupRun <= k1) {  // whole run: 16-byte stores (kb is a multiple of kDupRun)
    uint4* tk = reinterpret_cast<uint4*>(tile_key + kb);
    uint4* gd = reinterpret_cast<uint4*>(gid + kb);
    tk[0] = make_uint4(keys[0], keys[1], keys[2], keys[3]);
    tk[1] = make_uint4(keys[4], keys[5], keys[6], keys[7]);
    gd[0] = make_uint4(ids[0], ids[1], ids[2], ids[3]);
    gd[1] = make_uint4(ids[4], ids[5], ids[6], ids[7]);
    if (no_record) {
      uint4* nr = reinterpret_cast<uint4*>(no_record + kb);
      nr[0] = nr[1] = make_uint4(kNoRecord, kNoRecord, kNoRecord, kNoRecord);
    }
  } else {
#pragma unroll
    for (int u = 0; u < kDupRun; ++u) {
      if (kb + (uint32_t)u < k1) {
        tile_key[kb + u] = keys[u];
        gid[kb + u] = ids[u];
        if (no_record) no_record[kb + u] = kNoRecord;
      }
    }
  }
}

hipError_t launch_duplicate(int P, uint32_t R, const uint32_t* R_dev, const uint32_t* order, const uint32_t* offsets,
                            const uint2* rect_depth, TileGrid grid, uint32_t* tile_key, uint32_t* gid,
                            uint32_t* no_record, uint32_t* counts, int key_bits, uint2* ranges, uint32_t* blend_ok,
                            hipStream_t s, const uint32_t* P_dev) {
  if (P == 0 || R == 0) return hipSuccess;
  const uint32_t blocks = (R + kDupChunk - 1) / kDupChunk;
  ViewTable none{};
  hipLaunchKernelGGL(duplicate_kernel<false>, dim3(blocks), dim3(kDupThreads), 0, s, P, R, R_dev,
                     (uint32_t*)nullptr, order, offsets, rect_depth, grid, none, tile_key, gid, no_record, counts,
                     radix_first_mask(key_bits), (uint32_t)sort_blocks(R), reinterpret_cast<uint32_t*>(ranges),
                     (uint32_t)(2 * grid.tiles()), blend_ok, P_dev);
  return hipGetLastError();
}

hipError_t launch_duplicate_views(int items, uint32_t capacity, const uint32_t* R_dev, uint32_t* status,
                                  const uint32_t* order, const uint32_t* offsets, const uint2* rect_depth,
                                  const ViewTable& vt, uint32_t* tile_key, uint32_t* gid, uint32_t* no_record,
                                  uint32_t* counts, int key_bits, uint2* ranges, uint32_t* blend_ok, hipStream_t s) {
  if (items == 0 || capacity == 0) return hipSuccess;
  const uint32_t blocks = (capacity + kDupChunk - 1) / kDupChunk;
  hipLaunchKernelGGL(duplicate_kernel<true>, dim3(blocks), dim3(kDupThreads), 0, s, items, capacity, R_dev, status,
                     order, offsets, rect_depth, vt.grid(0), vt, tile_key, gid, no_record, counts,
                     radix_first_mask(key_bits), (uint32_t)sort_blocks(capacity), reinterpret_cast<uint32_t*>(ranges),
                     2u * vt.key_count, blend_ok, (const uint32_t*)nullptr);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tile-row bands (the checkpointed backward bins one band of tile rows at a
// time).
// ---------------------------------------------------------------------------
// Instances per tile row as a difference array: a rect of width w over rows
// [y0, y0 + h) adds w at y0 and takes it off at y0 + h (mod 2^32).  With kLds
// each block counts its Gaussians (grid-stride) in LDS and adds its nonzero
// entries once.
constexpr int kRowLdsRows = 8192;  // rows + 1 counted in LDS (32 KiB)
template <bool kLds>
__global__ __launch_bounds__(256) void row_counts_kernel(int P, const uint2* __restrict__ dup_rect, int gy,
                                                         uint32_t* __restrict__ row_diff) {
  extern __shared__ uint32_t s_diff[];
  if constexpr (kLds) {
    for (int y = threadIdx.x; y <= gy; y += blockDim.x) s_diff[y] = 0u;
    __syncthreads();
  }
  uint32_t* d = kLds ? s_diff : row_diff;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
    const uint2 r = dup_rect[i];
    const uint32_t h = r.x >> 16, w = r.y >> 16, y0 = r.y & 0xFFFFu;
    if (h == 0) continue;  // culled: the empty record
    atomicAdd(&d[y0], w);
    atomicAdd(&d[y0 + h], 0u - w);
  }
  if constexpr (kLds) {
    __syncthreads();
    for (int y = threadIdx.x; y <= gy; y += blockDim.x)
      if (s_diff[y]) atomicAdd(&row_diff[y], s_diff[y]);
  }
}

hipError_t launch_row_counts(int P, const uint2* dup_rect, int gy, uint32_t* row_diff, hipStream_t s) {
  const hipError_t e = launch_fill_u32(row_diff, (size_t)gy + 1, 0u, s);
  if (e != hipSuccess || P == 0) return e;
  const size_t want = ((size_t)P + 255) / 256;
  const unsigned blocks = (unsigned)(want < 1024 ? want : 1024);
  if (gy + 1 <= kRowLdsRows)
    hipLaunchKernelGGL(row_counts_kernel<true>, dim3(blocks), dim3(256), (size_t)(gy + 1) * sizeof(uint32_t), s, P,
                       dup_rect, gy, row_diff);
  else
    hipLaunchKernelGGL(row_counts_kernel<false>, dim3(blocks), dim3(256), 0, s, P, dup_rect, gy, row_diff);
  return hipGetLastError();
}

// The whole rects in depth order (rect_d[i] = dup_rect[order[i]]), gathered
// once for all bands.
__global__ void gather_rects_kernel(const uint2* __restrict__ dup_rect, const uint32_t* __restrict__ order, int P,
                                    uint2* __restrict__ rect_d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P) rect_d[i] = dup_rect[order[i]];
}

hipError_t launch_gather_rects(const uint2* dup_rect, const uint32_t* order, int P, uint2* rect_d, hipStream_t s) {
  if (P == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_rects_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, dup_rect, order, P,
                     rect_d);
  return hipGetLastError();
}

// A band's owners: the depth-ordered Gaussians with rect rows in the band,
// compacted in order with their inclusive instance offsets and band rects,
// so that the expansion sees no owner without an instance.  Thread t of a
// block takes kScanItems consecutive entries of the block's scan tile (so
// that exclusive scans of the threads' sums keep them in order).
__device__ __forceinline__ void band_items(const uint2* __restrict__ rect_d, size_t n, size_t base, uint32_t by0,
                                           uint32_t by1, uint2 r[kScanItems], uint32_t& area, uint32_t& count) {
  area = count = 0u;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t i = base + (size_t)k;
    r[k] = band_dup_rect(i < n ? rect_d[i] : make_uint2(0u, 1u << 16), by0, by1);
    const uint32_t a = dup_rect_area(r[k]);
    area += a;
    count += a ? 1u : 0u;
  }
}

// Pass 1: per scan tile, the band's instances and owners.
__global__ __launch_bounds__(kScanThreads) void band_reduce_kernel(const uint2* __restrict__ rect_d, size_t n,
                                                                   uint32_t by0, uint32_t by1,
                                                                   uint32_t* __restrict__ part_area,
                                                                   uint32_t* __restrict__ part_count) {
  uint2 r[kScanItems];
  uint32_t area, count, tot_a, tot_c;
  band_items(rect_d, n, (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems, by0, by1, r, area, count);
  block_exclusive_scan(area, &tot_a);
  block_exclusive_scan(count, &tot_c);
  if (threadIdx.x == 0) {
    part_area[blockIdx.x] = tot_a;
    part_count[blockIdx.x] = tot_c;
  }
}

// Pass 2: both partial arrays scanned at once (block 0: areas, block 1:
// owner counts); totals at [nb] with the overflow flag at [nb + 1].
__global__ __launch_bounds__(kScanThreads) void band_partials_kernel(uint32_t* __restrict__ part_area,
                                                                     uint32_t* __restrict__ part_count, size_t nb) {
  scan_partials_block(blockIdx.x == 0 ? part_area : part_count, nb);
}

// Pass 3: the owners' entries, compacted, and each owner's first slot.
__global__ __launch_bounds__(kScanThreads) void band_owners_kernel(
    const uint2* __restrict__ rect_d, const uint32_t* __restrict__ order, size_t n, uint32_t by0, uint32_t by1,
    const uint32_t* __restrict__ part_area, const uint32_t* __restrict__ part_count, uint32_t* __restrict__ order_c,
    uint32_t* __restrict__ offsets_c, uint2* __restrict__ rect_c, uint32_t* __restrict__ slot_words) {
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  uint2 r[kScanItems];
  uint32_t area, count, tot;
  band_items(rect_d, n, base, by0, by1, r, area, count);
  uint32_t off = part_area[blockIdx.x] + block_exclusive_scan(area, &tot);
  uint32_t pos = part_count[blockIdx.x] + block_exclusive_scan(count, &tot);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t a = dup_rect_area(r[k]);
    if (a) {
      const uint32_t g = order[base + (size_t)k];
      slot_words[g] = off;
      off += a;
      order_c[pos] = g;
      offsets_c[pos] = off;
      rect_c[pos] = dup_rect_xy(r[k]);
      ++pos;
    }
  }
}

hipError_t launch_band_owners(const uint2* rect_d, const uint32_t* order, size_t n, uint32_t by0, uint32_t by1,
                              uint32_t* part_area, uint32_t* part_count, uint32_t* order_c, uint32_t* offsets_c,
                              uint2* rect_c, uint32_t* slot_words, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t nb = scan_blocks(n);
  hipLaunchKernelGGL(band_reduce_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s, rect_d, n, by0, by1,
                     part_area, part_count);
  hipLaunchKernelGGL(band_partials_kernel, dim3(2), dim3(kScanThreads), 0, s, part_area, part_count, nb);
  hipLaunchKernelGGL(band_owners_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s, rect_d, order, n, by0, by1,
                     (const uint32_t*)part_area, (const uint32_t*)part_count, order_c, offsets_c, rect_c,
                     slot_words);
  return hipGetLastError();
}

// The view of every item of a batch in its current order (item i is in view
// i / P), as keys for the stable sort that regroups the depth order by view.
__global__ void view_keys_kernel(const uint32_t* __restrict__ order, size_t n, uint32_t P, uint32_t* __restrict__ keys) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keys[i] = order[i] / P;
}

hipError_t launch_view_keys(const uint32_t* order, size_t n, uint32_t P, uint32_t* keys, hipStream_t s) {
  if (n == 0 || P == 0) return hipSuccess;
  hipLaunchKernelGGL(view_keys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, order, n, P, keys);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tile ranges [start, end) from the sorted tile keys (reference
// identifyTileRanges, rasterizer_impl.cu:145-167); empty tiles stay (0, 0).
// ---------------------------------------------------------------------------
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

// Four consecutive keys per thread (one 16-byte load) and the key before
// them; a tile's range starts where its key differs from the previous one.
constexpr int kRangeKeys = 4;
__global__ void tile_ranges_kernel(const uint32_t* __restrict__ keys, uint32_t R, const uint32_t* __restrict__ R_dev,
                                   uint2* __restrict__ ranges) {
  const uint32_t k0 = (blockIdx.x * blockDim.x + threadIdx.x) * (uint32_t)kRangeKeys;
  if (R_dev) R = min(R, *R_dev);
  if (k0 >= R) return;
  uint32_t t[kRangeKeys];
  if (k0 + kRangeKeys <= R) {
    const uint4 q = *reinterpret_cast<const uint4*>(keys + k0);  // keys is 16-byte aligned
    t[0] = q.x;
    t[1] = q.y;
    t[2] = q.z;
    t[3] = q.w;
  } else {
#pragma unroll
    for (int j = 0; j < kRangeKeys; ++j) t[j] = k0 + j < R ? keys[k0 + j] : 0u;
  }
  uint32_t prev = k0 > 0 ? keys[k0 - 1] : 0u;
#pragma unroll
  for (int j = 0; j < kRangeKeys; ++j) {
    const uint32_t k = k0 + (uint32_t)j;
    if (k >= R) break;
    if (k == 0) {
      ranges[t[j]].x = 0;
    } else if (prev != t[j]) {
      ranges[prev].y = k;
      ranges[t[j]].x = k;
    }
    if (k == R - 1) ranges[t[j]].y = R;
    prev = t[j];
  }
}

hipError_t launch_tile_ranges(const uint32_t* sorted_tile_keys, size_t R, uint2* ranges, int tiles,
                              hipStream_t s, const uint32_t* R_dev, bool zeroed) {
  if (!zeroed) {
    const hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(ranges), 2 * (size_t)tiles, 0u, s);
    if (e != hipSuccess) return e;
  }
  if (R == 0) return hipSuccess;
  const size_t threads = (R + kRangeKeys - 1) / kRangeKeys;
  hipLaunchKernelGGL(tile_ranges_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, sorted_tile_keys,
                     (uint32_t)R, R_dev, ranges);
  return hipGetLastError();
}

}  // namespace omnigs
