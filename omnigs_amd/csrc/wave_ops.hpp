// Wave-level reductions on gfx950: DPP row moves and the permlane swaps
// (64-lane wavefronts).
#pragma once
#include "common.hpp"

namespace omnigs {

// Cross-lane helpers (gfx950): DPP moves, and the permlane swaps that trade
// the upper 32 lanes of one VGPR with the lower 32 of another (32) or the odd
// 16-lane rows of one with the even rows of another (16).
template <int Ctrl>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), Ctrl, 0xf, 0xf, false));
}
// x + y of a 32-lane swap: lanes 0-31 hold x[i] + x[i+32], lanes 32-63 hold y[i-32] + y[i].
__device__ __forceinline__ float fold32(float x, float y) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(x), __float_as_int(y), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}
// Rows 0,2 hold x summed over row pairs (0+1, 2+3); rows 1,3 hold y likewise.
__device__ __forceinline__ float fold16(float x, float y) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_int(x), __float_as_int(y), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}

// Reduces eight per-lane terms t[0..7] over the wave by recursive halving:
// on return every lane of 8-lane group g = lane >> 3 holds sum(t[g]).
__device__ __forceinline__ float wave_reduce8(const float t[8], int lane) {
  // lanes 0-31 carry t[0..3], lanes 32-63 carry t[4..7]
  const float a0 = fold32(t[0], t[4]), a1 = fold32(t[1], t[5]);
  const float a2 = fold32(t[2], t[6]), a3 = fold32(t[3], t[7]);
  // bit 4 of the lane picks between (a0, a2) and (a1, a3)
  const float b0 = fold16(a0, a2), b1 = fold16(a1, a3);
  // bit 3 of the lane keeps b1 (else b0); the partner 8 lanes away gives the other
  const bool hi = (lane >> 3) & 1;
  const float keep = hi ? b1 : b0, give = hi ? b0 : b1;
  float c = keep + dpp<0x128>(give);  // row_ror:8 (lane i <-> i^8 within a 16-lane row)
  c += dpp<0x141>(c);                 // row_half_mirror: i + (7 - i) within 8 lanes
  c += dpp<0x4e>(c);                  // quad_perm [2,3,0,1]
  c += dpp<0xb1>(c);                  // quad_perm [1,0,3,2]
  return c;
}

// Full wave sum of one term; every lane receives it.
__device__ __forceinline__ float wave_reduce1(float v) {
  float c = fold32(v, v);
  c = fold16(c, c);
  c += dpp<0x128>(c);
  c += dpp<0x141>(c);
  c += dpp<0x4e>(c);
  c += dpp<0xb1>(c);
  return c;
}

// Orders one wave's LDS accesses at this point for the compiler (the
// hardware executes a wave's LDS operations in order), for data that lanes of
// the same wave exchange through LDS.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace omnigs
