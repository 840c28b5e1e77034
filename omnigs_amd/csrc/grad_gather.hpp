// Per-Gaussian (per-item) sum of the render backward's gradient records,
// shared by the gather kernel and the fused preprocess backward.
//
// Item i owns the slots [base, base + tiles_touched) with base =
// slot_base[block of its Gaussian] + slot_local[i]; consecutive items' slots
// are adjacent, so neighbouring lanes read neighbouring records.  A slot the
// backward did not write (its tile never staged the instance) holds kNoRecord
// in its opacity word.  Slots are fetched kGatherGroup per round trip (the
// opacity words, then the records that exist).  Items with more than
// kGatherSerial instances (a few Gaussians cover thousands of tiles) are
// summed one at a time by the whole wave, lanes striding over the slots, so
// no lane walks a long list alone.  nslots (with R_dev, the live count)
// bounds the slots read.
#pragma once
#include "common.hpp"
#include "wave_ops.hpp"

namespace omnigs {

constexpr uint32_t kGatherSerial = 16;
constexpr int kGatherGroup = 8;  // slots fetched per round trip

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

// acc[0..8] += the sums of the records in slots [base, base + cnt) (below
// nslots): dL/dcolor (3), dL/dmean2D (2), dL/dconic (a, b, c), dL/dopacity
// (kInit: acc is zeroed first).  Every lane of the wave must call it (the
// wide path is wave-cooperative); lanes with cnt = 0 add nothing and help.
template <bool kInit = true>
__device__ __forceinline__ void gather_range(const float4* __restrict__ grad_rec,
                                             const float* __restrict__ grad_opacity, uint32_t nslots, uint32_t base,
                                             uint32_t cnt, float acc[kGradTerms]) {
  const int lane = threadIdx.x & (kWave - 1);

  // adds the records of slots base + l for the group's l (l0, l0 + stride, ...)
  auto add_group = [&](uint32_t b, uint32_t n, uint32_t l0, uint32_t stride, float* a) {
    uint32_t jj[kGatherGroup];
    bool in[kGatherGroup];
#pragma unroll
    for (int u = 0; u < kGatherGroup; ++u) {
      const uint32_t l = l0 + (uint32_t)u * stride;
      in[u] = l < n && b + l < nslots;
      jj[u] = in[u] ? b + l : 0u;
    }
    uint32_t op[kGatherGroup];
#pragma unroll
    for (int u = 0; u < kGatherGroup; ++u)
      op[u] = in[u] ? __float_as_uint(grad_opacity[jj[u]]) : kNoRecord;
    float4 r0[kGatherGroup], r1[kGatherGroup];
#pragma unroll
    for (int u = 0; u < kGatherGroup; ++u) {
      r0[u] = r1[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (op[u] != kNoRecord) {
        r0[u] = grad_rec[2 * (size_t)jj[u]];
        r1[u] = grad_rec[2 * (size_t)jj[u] + 1];
      }
    }
#pragma unroll
    for (int u = 0; u < kGatherGroup; ++u) {
      a[0] += r0[u].x; a[1] += r0[u].y; a[2] += r0[u].z; a[3] += r0[u].w;
      a[4] += r1[u].x; a[5] += r1[u].y; a[6] += r1[u].z; a[7] += r1[u].w;
      a[8] += op[u] != kNoRecord ? __uint_as_float(op[u]) : 0.f;
    }
  };

  if constexpr (kInit) {
#pragma unroll
    for (int q = 0; q < kGradTerms; ++q) acc[q] = 0.f;
  }
  if (cnt <= kGatherSerial)
    for (uint32_t l0 = 0; l0 < cnt; l0 += kGatherGroup) add_group(base, cnt, l0, 1, acc);
  uint64_t big = __ballot(cnt > kGatherSerial);
  while (big) {
    const int from = __ffsll((unsigned long long)big) - 1;
    big &= big - 1;
    const uint32_t bcnt = readlane_u32(cnt, from), bbase = readlane_u32(base, from);
    float part[kGradTerms];
#pragma unroll
    for (int q = 0; q < kGradTerms; ++q) part[q] = 0.f;
    for (uint32_t l0 = lane; l0 < bcnt; l0 += kWave * kGatherGroup) add_group(bbase, bcnt, l0, kWave, part);
    const float grp = wave_reduce8(part, lane);  // term q in lanes 8q .. 8q + 7
    const float sw = wave_reduce1(part[8]);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float tot = __uint_as_float(readlane_u32(__float_as_uint(grp), 8 * q));
      if (lane == from) acc[q] = kInit ? tot : acc[q] + tot;
    }
    if (lane == from) acc[8] = kInit ? sw : acc[8] + sw;
  }
}

// acc[0..8] = the sums of item i's records (its slots: see the top of this
// file).  Every lane of the wave must call it; lanes with !valid get zeros
// and help.
__device__ __forceinline__ void gather_sum(const GatherSrc& src, int i, bool valid, float acc[kGradTerms]) {
  uint32_t nslots = src.nslots;
  if (src.R_dev) nslots = min(nslots, *src.R_dev);
  const uint32_t cnt = valid ? src.tiles_touched[i] : 0u;
  const uint32_t base = cnt > 0 ? src.slot_base[(uint32_t)(i % src.P) / kSlotBlock] + src.slot_local[i] : 0u;
  gather_range<true>(src.grad_rec, src.grad_opacity, nslots, base, cnt, acc);
}

// Reads back what store_gathered wrote for item i.
__device__ __forceinline__ void load_gathered(size_t i, float acc[kGradTerms], const float* __restrict__ dL_dmean2D,
                                              const float* __restrict__ dL_dconic,
                                              const float* __restrict__ dL_dopacity,
                                              const float* __restrict__ dL_dcolors) {
  acc[0] = dL_dcolors[3 * i];
  acc[1] = dL_dcolors[3 * i + 1];
  acc[2] = dL_dcolors[3 * i + 2];
  acc[3] = dL_dmean2D[3 * i];
  acc[4] = dL_dmean2D[3 * i + 1];
  acc[5] = dL_dconic[4 * i];
  acc[6] = dL_dconic[4 * i + 1];
  acc[7] = dL_dconic[4 * i + 3];
  acc[8] = dL_dopacity[i];
}

// Writes item i's sums in the reference's layouts: dL_dmean2D [.,3] (z = 0),
// dL_dconic [.,2,2] ((1,0) = 0), dL_dopacity [.], dL_dcolors [.,3].
__device__ __forceinline__ void store_gathered(size_t i, const float acc[kGradTerms], float* __restrict__ dL_dmean2D,
                                               float* __restrict__ dL_dconic, float* __restrict__ dL_dopacity,
                                               float* __restrict__ dL_dcolors) {
  dL_dcolors[3 * i] = acc[0];
  dL_dcolors[3 * i + 1] = acc[1];
  dL_dcolors[3 * i + 2] = acc[2];
  dL_dmean2D[3 * i] = acc[3];
  dL_dmean2D[3 * i + 1] = acc[4];
  dL_dmean2D[3 * i + 2] = 0.f;
  dL_dconic[4 * i] = acc[5];
  dL_dconic[4 * i + 1] = acc[6];
  dL_dconic[4 * i + 2] = 0.f;
  dL_dconic[4 * i + 3] = acc[7];
  dL_dopacity[i] = acc[8];
}

}  // namespace omnigs
