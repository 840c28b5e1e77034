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

// ---------------------------------------------------------------------------
// Wave / block scans (64-lane wavefronts, 4 waves per 256-thread block).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v, int lane) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t n = __shfl_up(v, o, kWave);
    if (lane >= o) v += n;
  }
  return v;
}

// Exclusive scan of one value per thread across a 256-thread block.
// Returns the exclusive prefix; *total receives the block sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kScanThreads / kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const uint32_t inc = wave_inclusive_scan(v, lane);
  if (lane == kWave - 1) wsum[w] = inc;
  __syncthreads();
  uint32_t woff = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanThreads / kWave; ++i) {
    const uint32_t s = wsum[i];
    woff += i < w ? s : 0u;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return woff + inc - v;
}

__device__ __forceinline__ uint32_t load_in(const uint32_t* __restrict__ in, const uint32_t* __restrict__ idx,
                                            size_t i) {
  return idx ? in[idx[i]] : in[i];
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in,
                                                                   const uint32_t* __restrict__ idx, size_t n,
                                                                   uint32_t* __restrict__ partials) {
  const size_t base = (size_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t i = base + (size_t)k * kScanThreads + threadIdx.x;
    if (i < n) s += load_in(in, idx, i);
  }
  uint32_t tot;
  block_exclusive_scan(s, &tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Exclusive scan of the block partials in place, one block looping over chunks.
__global__ __launch_bounds__(kScanThreads) void scan_partials_kernel(uint32_t* __restrict__ partials, size_t nb) {
  uint32_t carry = 0;
  for (size_t c = 0; c < nb; c += kScanThreads) {
    const size_t i = c + threadIdx.x;
    const uint32_t v = i < nb ? partials[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(v, &tot);
    if (i < nb) partials[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) partials[nb] = carry;
}

// Each thread owns kScanItems consecutive elements (blocked layout).
__global__ __launch_bounds__(kScanThreads) void scan_downsweep_kernel(const uint32_t* __restrict__ in,
                                                                      const uint32_t* __restrict__ idx, size_t n,
                                                                      const uint32_t* __restrict__ partials,
                                                                      uint32_t* __restrict__ out, int exclusive) {
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t i = base + k;
    v[k] = i < n ? load_in(in, idx, i) : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = block_exclusive_scan(s, &tot) + partials[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const size_t i = base + k;
    const uint32_t inc = run + v[k];
    if (i < n) out[i] = exclusive ? run : inc;
    run = inc;
  }
}

hipError_t launch_scan(const uint32_t* in, const uint32_t* idx, uint32_t* out, size_t n, bool exclusive,
                       uint32_t* partials, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t nb = scan_blocks(n);
  hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(kScanThreads), 0, s, in, idx, n, partials);
  hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanThreads), 0, s, partials, nb);
  hipLaunchKernelGGL(scan_downsweep_kernel, dim3(nb), dim3(kScanThreads), 0, s, in, idx, n,
                     (const uint32_t*)partials, out, exclusive ? 1 : 0);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.  Per pass: digit counts per block,
// an exclusive scan of the digit-major count table, then a stable scatter in
// which every block ranks its 4096 keys in input order with 64-lane ballots.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSortThreads) void radix_count_kernel(const uint32_t* __restrict__ keys, size_t n,
                                                                   int shift, uint32_t mask,
                                                                   uint32_t* __restrict__ counts, uint32_t nblocks) {
  __shared__ uint32_t hist[kRadix];
  for (int d = threadIdx.x; d < kRadix; d += kSortThreads) hist[d] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kSortTile;
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const size_t i = base + (size_t)k * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&hist[(keys[i] >> shift) & mask], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadix; d += kSortThreads) counts[(size_t)d * nblocks + blockIdx.x] = hist[d];
}

__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, size_t n, int shift, int bits, const uint32_t* __restrict__ offs,
    uint32_t nblocks) {
  constexpr int kWaves = kSortThreads / kWave;
  __shared__ uint32_t run[kRadix];
  __shared__ uint32_t wcnt[kWaves][kRadix];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
  const uint32_t mask = (1u << bits) - 1u;
  for (int d = tid; d < kRadix; d += kSortThreads) {
    run[d] = offs[(size_t)d * nblocks + blockIdx.x];
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) wcnt[ww][d] = 0;
  }
  __syncthreads();
  const uint64_t lanes_below = (1ull << lane) - 1ull;
  const size_t base = (size_t)blockIdx.x * kSortTile;
  for (int k = 0; k < kSortItems; ++k) {
    const size_t i = base + (size_t)k * kSortThreads + tid;
    if (base + (size_t)k * kSortThreads >= n) break;  // uniform across the block
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u;
    const uint32_t val = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
    const uint32_t d = (key >> shift) & mask;
    // lanes of this wave holding the same digit
    uint64_t same = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const bool set = (d >> b) & 1u;
      const uint64_t bb = __ballot(set);
      same &= set ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(same & lanes_below);
    if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(same);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (int ww = 0; ww < w; ++ww) pos += wcnt[ww][d];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    for (int dd = tid; dd < kRadix; dd += kSortThreads) {
      uint32_t add = 0;
#pragma unroll
      for (int ww = 0; ww < kWaves; ++ww) {
        add += wcnt[ww][dd];
        wcnt[ww][dd] = 0;
      }
      run[dd] += add;
    }
    __syncthreads();
  }
}

hipError_t launch_radix_sort(uint32_t* keys, uint32_t* keys_alt, const uint32_t* vals_in, uint32_t* vals,
                             uint32_t* vals_alt, size_t n, int bits, uint32_t* counts, uint32_t* partials,
                             hipStream_t s, SortResult* out) {
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
  const size_t ncounts = (size_t)kRadix * nb;
  for (int shift = 0; shift < bits; shift += kRadixBits) {
    const int pb = bits - shift < kRadixBits ? bits - shift : kRadixBits;
    const uint32_t mask = (1u << pb) - 1u;
    hipLaunchKernelGGL(radix_count_kernel, dim3(nb), dim3(kSortThreads), 0, s, (const uint32_t*)ksrc, n, shift,
                       mask, counts, nb);
    hipError_t e = launch_scan(counts, nullptr, counts, ncounts, true, partials, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(radix_scatter_kernel, dim3(nb), dim3(kSortThreads), 0, s, (const uint32_t*)ksrc, vsrc, kdst,
                       vdst, n, shift, pb, (const uint32_t*)counts, nb);
    // swap
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
// Instance expansion in depth order.  Block b owns sorted positions
// [256 b, 256 b + 256); its instances [begin, end) are spread over all
// threads (load-balanced regardless of how many tiles each Gaussian covers),
// each finding its owner by binary search over the block's inclusive offsets.
// Tiles of one Gaussian are emitted row-major over its rect, as the reference's
// duplicateWithKeys does (rasterizer_impl.cu:127-138).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void duplicate_kernel(int P, const uint32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ offsets,
                                                        const float2* __restrict__ xy, const int* __restrict__ radii,
                                                        TileGrid grid, uint32_t* __restrict__ tile_key,
                                                        uint32_t* __restrict__ gid) {
  __shared__ uint32_t s_end[256];
  __shared__ uint32_t s_gid[256];
  __shared__ int s_x0[256], s_y0[256], s_w[256];
  const int tid = threadIdx.x;
  const int s0 = blockIdx.x * 256;
  const int nloc = min(256, P - s0);
  if (tid < nloc) {
    const uint32_t gg = order[s0 + tid];
    s_end[tid] = offsets[s0 + tid];
    s_gid[tid] = gg;
    const int r = radii[gg];
    if (r > 0) {
      const Rect rc = tile_rect(xy[gg], r, grid);
      s_x0[tid] = rc.x0;
      s_y0[tid] = rc.y0;
      s_w[tid] = rc.x1 - rc.x0;
    } else {
      s_x0[tid] = 0;
      s_y0[tid] = 0;
      s_w[tid] = 1;
    }
  }
  __syncthreads();
  const uint32_t begin = s0 == 0 ? 0u : offsets[s0 - 1];
  const uint32_t end = s_end[nloc - 1];
  for (uint32_t k = begin + tid; k < end; k += 256) {
    int lo = 0, hi = nloc - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_end[mid] > k) hi = mid; else lo = mid + 1;
    }
    const uint32_t first = lo == 0 ? begin : s_end[lo - 1];
    const uint32_t local = k - first;
    const uint32_t w = (uint32_t)s_w[lo];
    const uint32_t y = (uint32_t)s_y0[lo] + local / w;
    const uint32_t x = (uint32_t)s_x0[lo] + local % w;
    tile_key[k] = y * (uint32_t)grid.gx + x;
    gid[k] = s_gid[lo];
  }
}

hipError_t launch_duplicate(int P, const uint32_t* order, const uint32_t* offsets, const float2* xy,
                            const int* radii, TileGrid grid, uint32_t* tile_key, uint32_t* gid, hipStream_t s) {
  if (P == 0) return hipSuccess;
  hipLaunchKernelGGL(duplicate_kernel, dim3((P + 255) / 256), dim3(256), 0, s, P, order, offsets, xy, radii, grid,
                     tile_key, gid);
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

__global__ void tile_ranges_kernel(const uint32_t* __restrict__ keys, uint32_t R, uint2* __restrict__ ranges) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= R) return;
  const uint32_t t = keys[k];
  if (k == 0) {
    ranges[t].x = 0;
  } else {
    const uint32_t pt = keys[k - 1];
    if (pt != t) {
      ranges[pt].y = k;
      ranges[t].x = k;
    }
  }
  if (k == R - 1) ranges[t].y = R;
}

hipError_t launch_tile_ranges(const uint32_t* sorted_tile_keys, size_t R, uint2* ranges, int tiles,
                              hipStream_t s) {
  hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(ranges), 2 * (size_t)tiles, 0u, s);
  if (e != hipSuccess || R == 0) return e;
  hipLaunchKernelGGL(tile_ranges_kernel, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, s, sorted_tile_keys,
                     (uint32_t)R, ranges);
  return hipGetLastError();
}

}  // namespace omnigs
