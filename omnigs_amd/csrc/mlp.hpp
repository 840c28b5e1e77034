// Learned colour decoder: per-Gaussian 16-dim latent plus the 9 degree-2 SH
// basis values of the view direction -> 64 -> 64 -> 3 (ReLU, ReLU, sigmoid),
// BF16 operands on the gfx950 MFMA unit (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation.  It stands in for the degree-3 SH evaluation of the
// reference's computeColorFromSH (cuda_rasterizer/forward.cu:30-83) through
// the colour slot of the render record; it has no counterpart in the reference.
//
// Everything runs in transposed form, H = W X, with one matrix column per
// Gaussian, so that a wave evaluates 64 Gaussians as four 16-column blocks.
// For the 16x16x32 MFMA, lane l = c + 16 q (c < 16, q < 4) holds
//   A[row c][k = 8q + e]  and  B[k = 8q + e][col c]  (e < 8, BF16),
//   D[row 4q + e][col c]                               (e < 4, FP32).
// A D tile therefore feeds the next layer's B operand without moving data if
// the K index of that layer is permuted: B slot (q, e) of K block kb holds
// hidden unit pi(kb, q, e) = 32 kb + 4 q + e (e < 4) or 32 kb + 16 + 4 q + e - 4
// (e >= 4), the units lane (c, q) holds in D tiles 2 kb and 2 kb + 1.  The
// packed weight image stores every A operand in lane order with that
// permutation applied (mlp.hip, pack kernel).
#pragma once

#include "common.hpp"

namespace omnigs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kLatent = 16, kHidden = 64, kMlpIn = 25, kMlpOut = 3;

// FP32 master parameter vector (theta), row-major weights [out][in]:
constexpr int kThW1 = 0;                            // W1 [64][25]
constexpr int kThB1 = kThW1 + kHidden * kMlpIn;     // b1 [64]
constexpr int kThW2 = kThB1 + kHidden;              // W2 [64][64]
constexpr int kThB2 = kThW2 + kHidden * kHidden;    // b2 [64]
constexpr int kThW3 = kThB2 + kHidden;              // W3 [3][64]
constexpr int kThB3 = kThW3 + kMlpOut * kHidden;    // b3 [3]
constexpr int kThetaSize = kThB3 + kMlpOut;         // 6019
constexpr int kThetaStride = 6144;                  // per-wave partial row (padded)

// Packed image (bytes).  Forward operands first so a forward-only kernel
// stages only kMlpFwdBytes into LDS.
constexpr int kPkW1A = 0;                   // [4 rt][64 lanes] bf16x8
constexpr int kPkW2A = kPkW1A + 4 * 64 * 16;  // [4 rt][2 kb][64] bf16x8
constexpr int kPkW3A = kPkW2A + 8 * 64 * 16;  // [2 kb][64] bf16x8
constexpr int kPkB1 = kPkW3A + 2 * 64 * 16;   // f32 [64]
constexpr int kPkB2 = kPkB1 + 64 * 4;         // f32 [64]
constexpr int kPkB3 = kPkB2 + 64 * 4;         // f32 [4] (b3, 0)
constexpr int kMlpFwdBytes = kPkB3 + 16;      // 14864
constexpr int kPkW3T = 15104;                 // [4 rt][64] bf16x8: W3^T (K = 3 outputs)
constexpr int kPkW2T = kPkW3T + 4 * 64 * 16;  // [4 rt][2 kb][64] bf16x8: W2^T, K permuted
constexpr int kPkW1T = kPkW2T + 8 * 64 * 16;  // [2 rt][2 kb][64] bf16x8: W1^T (rows = inputs)
constexpr int kMlpPackedBytes = kPkW1T + 4 * 64 * 16;  // 31488

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short short2v __attribute__((ext_vector_type(2)));

// Two FP32 values as two BF16 (round to nearest even) in one dword, element
// 0 in the low half: one v_cvt_pk_bf16_f32.
__device__ __forceinline__ uint32_t cvt_pk_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

__device__ __forceinline__ bf16x8 pack_bf16(f32x4 lo, f32x4 hi) {
  return __builtin_bit_cast(bf16x8, make_uint4(cvt_pk_bf16(lo[0], lo[1]), cvt_pk_bf16(lo[2], lo[3]),
                                               cvt_pk_bf16(hi[0], hi[1]), cvt_pk_bf16(hi[2], hi[3])));
}

// ReLU of two BF16 in a dword: the bits as signed 16-bit integers, maxed
// with 0 (v_pk_max_i16), so every negative value and -0 become +0.  Rounding
// to BF16 keeps the sign, so relu(bf16(v)) = bf16(relu(v)).
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t v) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(short2v, v), (short2v){0, 0}));
}

// bf16(relu(lo)), bf16(relu(hi)) in pack_bf16's order.
__device__ __forceinline__ bf16x8 relu_pack_bf16(f32x4 lo, f32x4 hi) {
  const uint4 u = __builtin_bit_cast(uint4, pack_bf16(lo, hi));
  return __builtin_bit_cast(bf16x8, make_uint4(relu_bf16x2(u.x), relu_bf16x2(u.y), relu_bf16x2(u.z), relu_bf16x2(u.w)));
}

__device__ __forceinline__ bf16x8 zero_bf16x8() {
  bf16x8 r;
#pragma unroll
  for (int e = 0; e < 8; ++e) r[e] = (__bf16)0.f;
  return r;
}

// The 9 degree-2 real SH basis values the decoder sees (same constants and
// signs as the colour path's basis, auxiliary.h:32-49).
__device__ __forceinline__ void sh2_basis(float x, float y, float z, float b[9]) {
  b[0] = 0.28209479177387814f;
  b[1] = -0.4886025119029199f * y;
  b[2] = 0.4886025119029199f * z;
  b[3] = -0.4886025119029199f * x;
  b[4] = 1.0925484305920792f * x * y;
  b[5] = -1.0925484305920792f * y * z;
  b[6] = 0.31539156525252005f * (2.f * z * z - x * x - y * y);
  b[7] = -1.0925484305920792f * x * z;
  b[8] = 0.5462742152960396f * (x * x - y * y);
}

// d(sum_k g_k b_k)/d(x, y, z) for the basis above.
__device__ __forceinline__ float3 sh2_basis_vjp(float x, float y, float z, const float g[9]) {
  const float c1 = 0.4886025119029199f, c20 = 1.0925484305920792f, c22 = 0.31539156525252005f,
              c24 = 0.5462742152960396f;
  float3 r;
  r.x = -c1 * g[3] + c20 * y * g[4] - c20 * z * g[7] - 2.f * c22 * x * g[6] + 2.f * c24 * x * g[8];
  r.y = -c1 * g[1] + c20 * x * g[4] - c20 * z * g[5] - 2.f * c22 * y * g[6] - 2.f * c24 * y * g[8];
  r.z = c1 * g[2] - c20 * y * g[5] - c20 * x * g[7] + 4.f * c22 * z * g[6];
  return r;
}

// B operand of layer 1 for lane (c, q): inputs 8q..8q+7 of its Gaussian:
// q = 0, 1: latent; q = 2: basis 0..7; q = 3: basis 8, then zeros.
__device__ __forceinline__ bf16x8 input_operand(int q, const bf16x8& latent_half, float x, float y, float z) {
  float b[9];
  sh2_basis(x, y, z, b);
  const uint4 lat = __builtin_bit_cast(uint4, latent_half);
  const uint4 bas = q == 2 ? make_uint4(cvt_pk_bf16(b[0], b[1]), cvt_pk_bf16(b[2], b[3]), cvt_pk_bf16(b[4], b[5]),
                                        cvt_pk_bf16(b[6], b[7]))
                           : make_uint4(cvt_pk_bf16(b[8], 0.f), 0u, 0u, 0u);
  // selects, not branches: the operand is built by every lane alike
  return __builtin_bit_cast(bf16x8, q < 2 ? lat : bas);
}

// Activations of one 16-Gaussian column block: the post-ReLU layers as the
// next layer's BF16 B operands (units permuted as above), and the logits.
struct MlpAct {
  bf16x8 x2[2];  // layer 1 (D tiles 0, 1 | 2, 3)
  bf16x8 x3[2];  // layer 2
  f32x4 o;       // output logits: rows 0..2 valid in lanes q = 0
};

// Forward of one column block.  w: packed image (LDS or global), lane: 0..63.
__device__ __forceinline__ void mlp_forward_block(const char* w, int lane, const bf16x8& xin, MlpAct& a) {
  const bf16x8* W1A = reinterpret_cast<const bf16x8*>(w + kPkW1A);
  const bf16x8* W2A = reinterpret_cast<const bf16x8*>(w + kPkW2A);
  const bf16x8* W3A = reinterpret_cast<const bf16x8*>(w + kPkW3A);
  const f32x4* B1 = reinterpret_cast<const f32x4*>(w + kPkB1);
  const f32x4* B2 = reinterpret_cast<const f32x4*>(w + kPkB2);
  const f32x4 b3 = *reinterpret_cast<const f32x4*>(w + kPkB3);
  const int q = lane >> 4;
  f32x4 t[4];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) t[rt] = mfma16(W1A[rt * 64 + lane], xin, B1[rt * 4 + q]);
  a.x2[0] = relu_pack_bf16(t[0], t[1]);
  a.x2[1] = relu_pack_bf16(t[2], t[3]);
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    f32x4 acc = B2[rt * 4 + q];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) acc = mfma16(W2A[(rt * 2 + kb) * 64 + lane], a.x2[kb], acc);
    t[rt] = acc;
  }
  a.x3[0] = relu_pack_bf16(t[0], t[1]);
  a.x3[1] = relu_pack_bf16(t[2], t[3]);
  f32x4 acc = q == 0 ? b3 : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) acc = mfma16(W3A[kb * 64 + lane], a.x3[kb], acc);
  a.o = acc;
}

// Seeded direction k of Gaussian g for distillation: uniform on the sphere
// from a counter-based integer hash (same integers on host and device).
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__host__ __device__ __forceinline__ float3 seeded_direction(uint32_t seed, uint32_t g, uint32_t k) {
  const uint32_t h1 = mix32(seed ^ mix32(g * 0x9E3779B9U + k * 0x85EBCA6BU + 0x27d4eb2fU));
  const uint32_t h2 = mix32(h1 ^ 0x165667b1U);
  const float u = (float)(h1 >> 8) * (1.0f / 16777216.0f), v = (float)(h2 >> 8) * (1.0f / 16777216.0f);
  const float z = 1.f - 2.f * u;
  const float s = sqrtf(fmaxf(0.f, 1.f - z * z));
  const float ph = 6.283185307179586f * v;
  return make_float3(s * cosf(ph), s * sinf(ph), z);
}

// Seeded view centre k of Gaussian g: uniform in the box [lo, hi] (three
// more hash words of the same counter).  The distillation direction is then
// the view direction from that centre, unit(mean - centre), as a render from
// a camera there sees the Gaussian.
__host__ __device__ __forceinline__ float3 seeded_view_centre(uint32_t seed, uint32_t g, uint32_t k, float3 lo,
                                                              float3 hi) {
  const uint32_t h0 = mix32(seed ^ mix32(g * 0x9E3779B9U + k * 0x85EBCA6BU + 0x27d4eb2fU));
  const uint32_t h1 = mix32(h0 ^ 0x5bd1e995U), h2 = mix32(h1 ^ 0x165667b1U), h3 = mix32(h2 ^ 0xc2b2ae35U);
  const float s = 1.0f / 16777216.0f;
  return make_float3(fmaf((float)(h1 >> 8) * s, hi.x - lo.x, lo.x), fmaf((float)(h2 >> 8) * s, hi.y - lo.y, lo.y),
                     fmaf((float)(h3 >> 8) * s, hi.z - lo.z, lo.z));
}

// Direction index a distillation step gives Gaussian g.
__host__ __device__ __forceinline__ uint32_t distill_dir_index(uint32_t seed, int step, uint32_t g, int num_dirs) {
  return mix32(seed * 0x632BE5ABU + (uint32_t)step * 0x9E3779B9U + g) % (uint32_t)num_dirs;
}

// Logistic sigmoid with the hardware reciprocal (1 ulp), the same in every
// decoder kernel, so a training pass sees the colours the renders use.
__device__ __forceinline__ float sigmoidf(float v) { return __builtin_amdgcn_rcpf(1.f + __expf(-v)); }

// Latent inputs 8h..8h+7 of Gaussian G (BF16 [P,16]); zeros when !ok.
__device__ __forceinline__ bf16x8 load_latent_half(const uint16_t* __restrict__ lat, size_t G, int q, bool ok) {
  if (!ok) return zero_bf16x8();
  return *reinterpret_cast<const bf16x8*>(lat + G * kLatent + 8 * q);
}

// Latents of the wave's 64 Gaussians (wave_base + 0..63) as the layer-1 B
// operand halves of the four column blocks: lane (c, q < 2) holds inputs
// 8q..8q+7 of Gaussian wave_base + 16 cb + c, other lanes zeros.  Issued
// ahead of other work so that their HBM latency overlaps it.  P >= 1.
__device__ __forceinline__ void mlp_wave_latents(const uint16_t* __restrict__ latents, int P, size_t wave_base,
                                                 bf16x8 lat[4]) {
  const int lane = threadIdx.x & 63, c = lane & 15, q = lane >> 4;
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    // every lane loads (from a clamped row) and masks afterwards: no branch
    // around the load, so all four stay in flight together
    const size_t G = wave_base + 16 * cb + c;
    const bool ok = G < (size_t)P && q < 2;
    const size_t Gc = G < (size_t)P ? G : (size_t)P - 1;
    const uint4 v = *reinterpret_cast<const uint4*>(latents + Gc * kLatent + 8 * (q & 1));
    const uint32_t m = ok ? 0xFFFFFFFFu : 0u;
    lat[cb] = __builtin_bit_cast(bf16x8, make_uint4(v.x & m, v.y & m, v.z & m, v.w & m));
  }
}

// Decoder colours of the wave's 64 Gaussians, evaluated inside the forward
// preprocess.  lat: from mlp_wave_latents; dir: this lane's own view
// direction.  All 64 lanes must call it.  w: forward part of the packed image
// in LDS.
__device__ __forceinline__ float3 mlp_colour_wave(const char* w, const bf16x8 lat[4], float3 dir) {
  const int lane = threadIdx.x & 63, c = lane & 15, q = lane >> 4;
  // each lane packs the basis of its own direction once; column block cb
  // then fetches Gaussian 16 cb + c's packs (the same bits input_operand
  // would compute there)
  float b[9];
  sh2_basis(dir.x, dir.y, dir.z, b);
  const uint32_t p0 = cvt_pk_bf16(b[0], b[1]), p1 = cvt_pk_bf16(b[2], b[3]), p2 = cvt_pk_bf16(b[4], b[5]),
                 p3 = cvt_pk_bf16(b[6], b[7]), p4 = cvt_pk_bf16(b[8], 0.f);
  float3 mine = make_float3(0.f, 0.f, 0.f);
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const int src = 16 * cb + c;
    const uint32_t v0 = __shfl(p0, src), v1 = __shfl(p1, src), v2 = __shfl(p2, src), v3 = __shfl(p3, src),
                   v4 = __shfl(p4, src);
    const uint4 bas = q == 2 ? make_uint4(v0, v1, v2, v3) : make_uint4(v4, 0u, 0u, 0u);
    const bf16x8 xin = __builtin_bit_cast(bf16x8, q < 2 ? __builtin_bit_cast(uint4, lat[cb]) : bas);
    MlpAct a;
    mlp_forward_block(w, lane, xin, a);
    // lanes (c, 0) hold the logits of Gaussian 16cb + c; lane 16cb + c takes them
    const float o0 = __shfl(a.o[0], c), o1 = __shfl(a.o[1], c), o2 = __shfl(a.o[2], c);
    if (q == cb) mine = make_float3(o0, o1, o2);
  }
  return make_float3(sigmoidf(mine.x), sigmoidf(mine.y), sigmoidf(mine.z));
}


}  // namespace omnigs
