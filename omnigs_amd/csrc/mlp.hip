// Colour decoder kernels: weight packing, standalone evaluation, the training
// pass (distillation from SH colours, or the backward of a render whose
// colours came from the decoder), the deterministic reduction of the weight
// gradient, Adam, and FP32 -> BF16 conversion.  Device math in mlp.hpp.
//
// Training pass layout: every wave walks its own pairs of 16-Gaussian
// column blocks.  Per block it runs the forward and backward chains, turns
// the activations and adjoints into K-over-Gaussians operands on the matrix
// cores (an MFMA against a 0/1 selector transposes a tile exactly), and adds
// the block's weight-gradient tiles into registers it keeps for the whole
// kernel: no LDS staging and no barrier inside the loop.  Bias gradients come
// from MFMAs against all-ones rows.  At the end the four waves of a block add
// their gradients in a fixed order through LDS, each block writes one partial
// row, and a two-stage kernel adds the rows in a fixed order.
#include "kernels.hpp"
#include "lonlat_math.hpp"
#include "mlp.hpp"

namespace omnigs {

namespace {

__device__ __forceinline__ int pi_unit(int kb, int q, int e) {
  return e < 4 ? 32 * kb + 4 * q + e : 32 * kb + 16 + 4 * q + (e - 4);
}

__global__ __launch_bounds__(256) void mlp_pack_kernel(const float* __restrict__ th, char* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  // 1920 lane entries: W1A 256, W2A 512, W3A 128, W3T 256, W2T 512, W1T 256; then 132 bias floats
  if (t < 1920) {
    int u = t;
    const int lane = u & 63, c = lane & 15, q = lane >> 4;
    float v[8];
    int off;
    if (u < 256) {  // W1A[rt]: W1[16rt + c][8q + e]
      const int rt = u >> 6, row = 16 * rt + c;
      for (int e = 0; e < 8; ++e) v[e] = (8 * q + e < kMlpIn) ? th[kThW1 + row * kMlpIn + 8 * q + e] : 0.f;
      off = kPkW1A + u * 16;
    } else if ((u -= 256) < 512) {  // W2A[rt][kb]: W2[16rt + c][pi]
      const int rt = u >> 7, kb = (u >> 6) & 1, row = 16 * rt + c;
      for (int e = 0; e < 8; ++e) v[e] = th[kThW2 + row * kHidden + pi_unit(kb, q, e)];
      off = kPkW2A + u * 16;
    } else if ((u -= 512) < 128) {  // W3A[kb]: W3[c][pi], rows >= 3 zero
      const int kb = u >> 6;
      for (int e = 0; e < 8; ++e) v[e] = c < kMlpOut ? th[kThW3 + c * kHidden + pi_unit(kb, q, e)] : 0.f;
      off = kPkW3A + u * 16;
    } else if ((u -= 128) < 256) {  // W3T[rt]: W3^T[16rt + c][k = 8q + e], k < 3
      const int rt = u >> 6;
      for (int e = 0; e < 8; ++e) v[e] = (q == 0 && e < kMlpOut) ? th[kThW3 + e * kHidden + 16 * rt + c] : 0.f;
      off = kPkW3T + u * 16;
    } else if ((u -= 256) < 512) {  // W2T[rt][kb]: W2[pi][16rt + c]
      const int rt = u >> 7, kb = (u >> 6) & 1;
      for (int e = 0; e < 8; ++e) v[e] = th[kThW2 + pi_unit(kb, q, e) * kHidden + 16 * rt + c];
      off = kPkW2T + u * 16;
    } else {  // W1T[rt][kb]: W1[pi][16rt + c], input index < 25
      u -= 512;
      const int rt = u >> 7, kb = (u >> 6) & 1, col = 16 * rt + c;
      for (int e = 0; e < 8; ++e) v[e] = col < kMlpIn ? th[kThW1 + pi_unit(kb, q, e) * kMlpIn + col] : 0.f;
      off = kPkW1T + u * 16;
    }
    bf16x8 r;
    for (int e = 0; e < 8; ++e) r[e] = (__bf16)v[e];
    *reinterpret_cast<bf16x8*>(out + off) = r;
  } else if (t < 1920 + 132) {
    const int b = t - 1920;
    float* o = reinterpret_cast<float*>(out);
    if (b < 64) o[kPkB1 / 4 + b] = th[kThB1 + b];
    else if (b < 128) o[kPkB2 / 4 + b - 64] = th[kThB2 + b - 64];
    else o[kPkB3 / 4 + b - 128] = (b - 128) < kMlpOut ? th[kThB3 + b - 128] : 0.f;
  }
}

// Copies n bytes (multiple of 16) of the packed image into LDS.
__device__ __forceinline__ void stage_weights(char* dst, const char* __restrict__ src, int n) {
  const uint4* s = reinterpret_cast<const uint4*>(src);
  uint4* d = reinterpret_cast<uint4*>(dst);
  for (int i = threadIdx.x; i < n / 16; i += blockDim.x) d[i] = s[i];
}

// Unit direction of a world-space vector (0, 0, 1 for a zero vector),
// computed by every lane (selects, no branch).
__device__ __forceinline__ float3 unit_dir(float3 v, float* len_out) {
  const float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  *len_out = len;
  const bool pos = len > 0.f;
  const float l = pos ? len : 1.f;
  const float3 d = make_float3(v.x / l, v.y / l, v.z / l);
  return pos ? d : make_float3(0.f, 0.f, 1.f);
}

__device__ __forceinline__ bf16x8 load_latent_half(const float* __restrict__ lat, size_t G, int q, bool ok) {
  if (!ok) return zero_bf16x8();
  const float4 a = *reinterpret_cast<const float4*>(lat + G * kLatent + 8 * q);
  const float4 b = *reinterpret_cast<const float4*>(lat + G * kLatent + 8 * q + 4);
  return pack_bf16(f32x4{a.x, a.y, a.z, a.w}, f32x4{b.x, b.y, b.z, b.w});
}

}  // namespace

// ----------------------------------------------------------------------------
// Standalone evaluation: rgb[i] = sigmoid(MLP(latent[i], Y2(dir[i]))).
__global__ __launch_bounds__(256) void mlp_eval_kernel(int N, const uint16_t* __restrict__ latents,
                                                       const float* __restrict__ dirs, const char* __restrict__ packed,
                                                       float* __restrict__ rgb) {
  __shared__ __attribute__((aligned(16))) char s_w[kMlpFwdBytes];
  stage_weights(s_w, packed, kMlpFwdBytes);
  __syncthreads();
  const int lane = threadIdx.x & 63, c = lane & 15, q = lane >> 4;
  const size_t wave_base = ((size_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63));
  if (wave_base >= (size_t)N) return;
  float3 mine = make_float3(0.f, 0.f, 0.f);
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    const size_t G = wave_base + 16 * cb + c;
    const bool ok = G < (size_t)N;
    float x = 0.f, y = 0.f, z = 1.f;
    if (ok && q >= 2) {
      x = dirs[3 * G];
      y = dirs[3 * G + 1];
      z = dirs[3 * G + 2];
    }
    const bf16x8 xin = input_operand(q, load_latent_half(latents, G, q & 1, ok && q < 2), x, y, z);
    MlpAct a;
    mlp_forward_block(s_w, lane, xin, a);
    // lanes (c, 0) hold the logits of Gaussian 16cb + c; lane 16cb + c takes them
    const float o0 = __shfl(a.o[0], c), o1 = __shfl(a.o[1], c), o2 = __shfl(a.o[2], c);
    if (q == cb) mine = make_float3(o0, o1, o2);
  }
  const size_t G = wave_base + lane;
  if (G < (size_t)N) {
    rgb[3 * G] = sigmoidf(mine.x);
    rgb[3 * G + 1] = sigmoidf(mine.y);
    rgb[3 * G + 2] = sigmoidf(mine.z);
  }
}

// ----------------------------------------------------------------------------
// Training pass.
struct MlpTrainArgs {
  int P;
  const char* packed;
  const float* latent_f32;    // distillation: FP32 master latents [P,16]
  const uint16_t* latent_bf16;  // render backward: BF16 latents [P,16]
  // render backward
  const float* means3D;
  const float* campos;
  const int* radii;
  const float* dL_dcolor;  // [V P, 3]
  float* dL_dmean3D;       // [P,3], += direction term of every view with radius > 0
  int V;                   // views (items v P + g); campos is [V,3]
  // V = 1: the Gaussians with a colour gradient, per 64-Gaussian wave (the
  // preprocess backward's list), or NULL for every Gaussian
  const uint32_t* act_list;
  const uint32_t* act_count;
  // distillation (means3D given: directions from seeded view centres in
  // [view_lo, view_hi]; NULL: seeded directions uniform on the sphere)
  const float* shs;  // [P,M,3] teacher, degree 3
  int M;
  float3 view_lo, view_hi;
  uint32_t seed;
  int step;
  int num_dirs;
  // outputs
  float* dL_dlatent;     // [P,16]
  float* theta_partial;  // [gridDim.x][kThetaStride]; column kThetaSize holds the block's squared error
};

// Weight-gradient operands need the Gaussians on the MFMA K index, while the
// chains above hold them on the lane index (D tiles, one Gaussian per lane
// c).  The matrix cores transpose: mfma(z, S) = M(z) M(S)^T, where M(r) is the
// 16x32 matrix a register holds in A-operand order.  When row j of M(S) has a
// single 1, at slot k(j), lane (j, q) element e of the result holds slot
// k(j) of Gaussian 4q + e: unit j of a tile on the lane, four Gaussians on
// its elements.  The values are BF16 operands times 1 plus zeros, so the
// FP32 result is exact and packs back to BF16 without rounding.  Selectors,
// in B-operand lane order:
enum : int {
  kSelLo,    // hidden pack, even tile: unit j at slot (q = j/4, e = j%4)
  kSelHi,    // hidden pack, odd tile: slot (q = j/4, e = 4 + j%4)
  kSelInLo,  // layer-1 input, inputs 0..15: slot j
  kSelInHi,  // inputs 16..31: slot 16 + j
  kSelDO,    // output adjoint: row j = 11..13 takes output j - 11
  kSelOne9,  // all ones in row 9 (a K = 16 operand: elements 0..3)
  kSelOne10,
  kSelOne14,
  kNumSel
};

__device__ __forceinline__ bf16x8 make_selector(int sel, int lane) {
  const int j = lane & 15, q = lane >> 4;
  int e = -1;
  bool all = false;
  if (sel == kSelLo) e = q == (j >> 2) ? (j & 3) : -1;
  else if (sel == kSelHi) e = q == (j >> 2) ? 4 + (j & 3) : -1;
  else if (sel == kSelInLo) e = q == (j >> 3) ? (j & 7) : -1;
  else if (sel == kSelInHi) e = q == 2 + (j >> 3) ? (j & 7) : -1;
  else if (sel == kSelDO) e = (q == 0 && j >= 11 && j < 14) ? j - 11 : -1;
  else all = j == (sel == kSelOne9 ? 9 : sel == kSelOne10 ? 10 : 14);
  bf16x8 r;
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = (__bf16)((all || k == e) ? 1.f : 0.f);
  return r;
}

// Four BF16-exact FP32 values as BF16 bits (two dwords).
__device__ __forceinline__ uint2 bf16x4_bits(f32x4 v) {
  return make_uint2(cvt_pk_bf16(v[0], v[1]), cvt_pk_bf16(v[2], v[3]));
}

// One column block's transposed tile (four Gaussians per lane).
__device__ __forceinline__ uint2 tr_half(const bf16x8& z, const bf16x8& sel) {
  return bf16x4_bits(mfma16(z, sel, f32x4{0.f, 0.f, 0.f, 0.f}));
}

// A weight-gradient operand over 32 Gaussians: slot (q, e) is Gaussian
// 4q + e of the first column block (e < 4) or 4q + e - 4 of the second.
__device__ __forceinline__ bf16x8 join_halves(uint2 a, uint2 b) {
  return __builtin_bit_cast(bf16x8, make_uint4(a.x, a.y, b.x, b.y));
}

// ReLU derivative on BF16 packs: element e of g where element e of the
// post-ReLU activation h (relu_pack_bf16: +0 or positive) is nonzero, else
// zero.  Per dword, min(h, 1) as unsigned 16-bit halves is the 0/1 mask and
// a 16-bit multiply applies it: two packed ops (v_pk_min_u16,
// v_pk_mul_lo_u16), written out because the compiler turns the same vector
// code into per-half compares and selects.
__device__ __forceinline__ bf16x8 relu_mask(const bf16x8& h, const bf16x8& g) {
  const uint4 hb = __builtin_bit_cast(uint4, h), gb = __builtin_bit_cast(uint4, g);
  const uint32_t hv[4] = {hb.x, hb.y, hb.z, hb.w}, gv[4] = {gb.x, gb.y, gb.z, gb.w};
  const uint32_t one = 0x00010001u;
  uint32_t r[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t m;
    asm("v_pk_min_u16 %0, %1, %2" : "=v"(m) : "v"(hv[i]), "v"(one));
    asm("v_pk_mul_lo_u16 %0, %1, %2" : "=v"(r[i]) : "v"(gv[i]), "v"(m));
  }
  return __builtin_bit_cast(bf16x8, make_uint4(r[0], r[1], r[2], r[3]));
}

// 32-bit cross-row moves: the value lane (c, 1) holds (row), or lane (c, 2)
// (row2), delivered to lane (c, 0).  A permlane swap of a register with
// itself leaves {own, other} in the two results of row 0 (in an order the
// XOR with the own value makes irrelevant).
__device__ __forceinline__ float from_row1(float v) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_int(v), __float_as_int(v), false, false);
  return __int_as_float(r[0] ^ r[1] ^ __float_as_int(v));
}
__device__ __forceinline__ float from_row2(float v) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(v), __float_as_int(v), false, false);
  return __int_as_float(r[0] ^ r[1] ^ __float_as_int(v));
}
// Sum over the four 16-lane rows, on every lane.
__device__ __forceinline__ float row_sum(float v) {
  auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(v), __float_as_int(v), false, false);
  const float a = __int_as_float(r[0]) + __int_as_float(r[1]);
  r = __builtin_amdgcn_permlane16_swap(__float_as_int(a), __float_as_int(a), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}

// Column blocks a wave carries through each step: independent chains that
// hide one another's MFMA and LDS latency (one wave per SIMD holds them).
constexpr int kTrainCB = 4;
// 64-Gaussian segments whose listed Gaussians a wave of the training pass
// holds in LDS at once.
constexpr int kSegBatch = 24;

template <bool kDistill>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void mlp_train_kernel(MlpTrainArgs a) {
  constexpr int NCB = kTrainCB, NP = NCB / 2;  // column blocks, and pairs of them (K = 32 Gaussians)
  __shared__ __attribute__((aligned(16))) char s_w[kMlpPackedBytes];
  __shared__ __attribute__((aligned(16))) bf16x8 s_sel[kNumSel][64];
  __shared__ float s_err[4];
  __shared__ uint32_t s_ids[4][kSegBatch * 64];  // per wave: its listed Gaussians (render backward)

  const int lane = threadIdx.x & 63, c = lane & 15, q = lane >> 4, w = threadIdx.x >> 6;
  const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
  const bf16x8* W1A = reinterpret_cast<const bf16x8*>(s_w + kPkW1A);
  const bf16x8* W2A = reinterpret_cast<const bf16x8*>(s_w + kPkW2A);
  const bf16x8* W3A = reinterpret_cast<const bf16x8*>(s_w + kPkW3A);
  const bf16x8* W3T = reinterpret_cast<const bf16x8*>(s_w + kPkW3T);
  const bf16x8* W2T = reinterpret_cast<const bf16x8*>(s_w + kPkW2T);
  const bf16x8* W1T = reinterpret_cast<const bf16x8*>(s_w + kPkW1T);
  const f32x4* B1 = reinterpret_cast<const f32x4*>(s_w + kPkB1);
  const f32x4* B2 = reinterpret_cast<const f32x4*>(s_w + kPkB2);
  // The wave's whole weight gradient, in D-tile order: gW2[nt][mt] holds
  // dW2 rows 16nt.., columns 16mt..; gW1[nt][it] dW1 rows 16nt.., inputs
  // 16it.. (it = 1: inputs 16..24 in columns 0..8), and in the free columns
  // of gW1[nt][1]: 9 = db1, 10 = db2, 11..13 = dW3^T, and (nt = 0, rows
  // 11..13) 14 = db3.
  f32x4 gW2[4][4], gW1[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int k = 0; k < 4; ++k) gW2[i][k] = z4;
    gW1[i][0] = z4;
    gW1[i][1] = z4;
  }
  float sq_err = 0.f;
  const float inv3P = 1.f / (3.f * (float)a.P);

  const int nwaves = (int)gridDim.x * 4;
  const int nviews = kDistill ? 1 : a.V;
  // Work.  Listed (render backward with the preprocess backward's list):
  // wave wid takes the 64-Gaussian segments [seg0, seg1), an equal share,
  // and copies the listed Gaussians of up to kSegBatch of them at a time into
  // its LDS id array, in index order; groups are 64 consecutive entries, so
  // only a batch's last group is partial.  Otherwise units of 64 consecutive
  // Gaussians, wave wid taking units wid, wid + nwaves, ...
  const bool listed = !kDistill && a.act_list != nullptr;
  const int wid = (int)blockIdx.x * 4 + w;
  const int nseg = (a.P + 63) / 64;
  const int seg0 = (int)((int64_t)wid * nseg / nwaves), seg1 = (int)((int64_t)(wid + 1) * nseg / nwaves);
  const int nunits = (a.P + 63) / 64;
  uint32_t* ids = s_ids[w];
  struct Cursor {
    int u, gi, n;  // listed: first segment of the batch; else the unit.  Group in it, Gaussians in it
  };
  auto more = [&](const Cursor& k) { return listed ? k.u < seg1 : k.u < nunits; };
  // the listed Gaussians of segments [sb, se) (se - sb <= kSegBatch) into ids
  auto list_batch = [&](int sb, int se) -> int {
    const int ns = se - sb;
    const uint32_t cl = lane < ns ? a.act_count[sb + lane] : 0u;
    uint32_t inc = cl;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
      const uint32_t t = __shfl_up(inc, o, 64);
      if (lane >= o) inc += t;
    }
    const uint32_t excl = inc - cl;
#pragma unroll
    for (int i0 = 0; i0 < kSegBatch; i0 += 8) {
      uint32_t v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t ci = (uint32_t)__shfl((int)cl, i0 + u, 64);
        v[u] = (uint32_t)lane < ci ? a.act_list[(size_t)64 * (sb + i0 + u) + lane] : 0u;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const uint32_t ci = (uint32_t)__shfl((int)cl, i0 + u, 64), ei = (uint32_t)__shfl((int)excl, i0 + u, 64);
        if ((uint32_t)lane < ci) ids[ei + lane] = v[u];
      }
    }
    // the wave's later reads of ids (other lanes) follow these writes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return __shfl((int)inc, ns - 1, 64);
  };
  auto unit_open = [&](Cursor& k) {
    if (listed) k.n = list_batch(k.u, min(seg1, k.u + kSegBatch));
    else k.n = min(64, a.P - 64 * k.u);
  };
  auto step = [&](Cursor& k) {
    if (64 * (k.gi + 1) < k.n) {
      ++k.gi;
      return;
    }
    k.gi = 0;
    for (k.u += listed ? kSegBatch : nwaves; more(k); k.u += listed ? kSegBatch : nwaves) {
      unit_open(k);
      if (k.n > 0) return;
    }
  };
  // Per column block: Gaussian G (lane c) and its inputs, loaded a group
  // ahead.  Loads use a valid index for every lane; results of lanes past
  // the group's Gaussians are masked.
  struct GroupIn {
    uint32_t G[NCB];
    bool ok[NCB];
    bf16x8 lat[NCB];
    float3 m[NCB], dm[NCB];
  };
  auto load_group = [&](const Cursor& k, GroupIn& in) {
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      const int j = 64 * k.gi + 16 * cb + c;
      in.ok[cb] = j < k.n;
      in.G[cb] = !in.ok[cb] ? 0u : listed ? ids[j] : (uint32_t)(64 * k.u + j);
      const size_t Gv = in.G[cb];
      if constexpr (kDistill) {
        const float4 x = *reinterpret_cast<const float4*>(a.latent_f32 + Gv * kLatent + 8 * (q & 1));
        const float4 y = *reinterpret_cast<const float4*>(a.latent_f32 + Gv * kLatent + 8 * (q & 1) + 4);
        in.lat[cb] = pack_bf16(f32x4{x.x, x.y, x.z, x.w}, f32x4{y.x, y.y, y.z, y.w});
        in.m[cb] = a.means3D ? make_float3(a.means3D[3 * Gv], a.means3D[3 * Gv + 1], a.means3D[3 * Gv + 2])
                             : make_float3(0.f, 0.f, 0.f);
        in.dm[cb] = make_float3(0.f, 0.f, 0.f);
      } else {
        in.lat[cb] = *reinterpret_cast<const bf16x8*>(a.latent_bf16 + Gv * kLatent + 8 * (q & 1));
        in.m[cb] = make_float3(a.means3D[3 * Gv], a.means3D[3 * Gv + 1], a.means3D[3 * Gv + 2]);
        in.dm[cb] = make_float3(a.dL_dmean3D[3 * Gv], a.dL_dmean3D[3 * Gv + 1], a.dL_dmean3D[3 * Gv + 2]);
      }
    }
  };
  // The first group's list and inputs are requested before the weights are
  // staged, so that their latencies overlap.
  Cursor nk{listed ? seg0 - kSegBatch : wid - nwaves, 0, 0};
  step(nk);
  GroupIn nxt;
  if (more(nk)) load_group(nk, nxt);
  stage_weights(s_w, a.packed, kMlpPackedBytes);
  for (int i = threadIdx.x; i < kNumSel * 64; i += blockDim.x) s_sel[i >> 6][i & 63] = make_selector(i >> 6, i & 63);
  __syncthreads();
  while (more(nk)) {
    const GroupIn cur = nxt;
    step(nk);
    if (more(nk)) load_group(nk, nxt);
    size_t Gs[NCB];
    bool ok[NCB];
    bf16x8 lat[NCB];
    float3 m[NCB], dm[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      ok[cb] = cur.ok[cb];
      Gs[cb] = cur.G[cb];
      lat[cb] = (ok[cb] && q < 2) ? cur.lat[cb] : zero_bf16x8();
      m[cb] = cur.m[cb];
      dm[cb] = cur.dm[cb];
    }
    float4 dlat[NCB];
    float3 dmv[NCB];
    bool any_dir[NCB];
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      dlat[cb] = make_float4(0.f, 0.f, 0.f, 0.f);
      dmv[cb] = make_float3(0.f, 0.f, 0.f);
      any_dir[cb] = false;
    }
    for (int vi = 0; vi < nviews; ++vi) {
      // keeps the weight and selector reads from LDS inside the loop (hoisted,
      // they would hold 150 registers for the whole kernel)
      asm volatile("" ::: "memory");
      // direction (every q lane of Gaussian c holds it), layer-1 input, and
      // the loads the adjoint needs
      float3 d[NCB];
      float len[NCB];
      bf16x8 xin[NCB];
      float3 dcv[NCB];
      int rad[NCB];
      float sh[NCB][12];
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        len[cb] = 1.f;
        rad[cb] = 0;
        dcv[cb] = make_float3(0.f, 0.f, 0.f);
        if constexpr (kDistill) {
          const uint32_t k = distill_dir_index(a.seed, a.step, (uint32_t)Gs[cb], a.num_dirs);
          if (a.means3D) {
            const float3 cv = seeded_view_centre(a.seed, (uint32_t)Gs[cb], k, a.view_lo, a.view_hi);
            d[cb] = unit_dir(make_float3(m[cb].x - cv.x, m[cb].y - cv.y, m[cb].z - cv.z), &len[cb]);
          } else {
            d[cb] = seeded_direction(a.seed, (uint32_t)Gs[cb], k);
          }
          const float* src = a.shs + (Gs[cb] * a.M + 4 * q) * 3;
#pragma unroll
          for (int i = 0; i < 12; ++i) sh[cb][i] = (4 * q + i / 3 < a.M) ? src[i] : 0.f;
        } else {
          const float3 cp = make_float3(a.campos[3 * vi], a.campos[3 * vi + 1], a.campos[3 * vi + 2]);
          d[cb] = unit_dir(make_float3(m[cb].x - cp.x, m[cb].y - cp.y, m[cb].z - cp.z), &len[cb]);
          const size_t item = (size_t)vi * a.P + Gs[cb];  // view vi's copy of Gaussian G
          dcv[cb] = make_float3(a.dL_dcolor[3 * item], a.dL_dcolor[3 * item + 1], a.dL_dcolor[3 * item + 2]);
          rad[cb] = a.radii[item];
        }
        xin[cb] = input_operand(q, lat[cb], d[cb].x, d[cb].y, d[cb].z);
      }

      // forward chains (units on the elements, Gaussian c on the lane); each
      // weight operand read once for all column blocks
      bf16x8 x2[NCB][2], x3[NCB][2];
      f32x4 t[NCB][4];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        const bf16x8 wa = W1A[rt * 64 + lane];
        const f32x4 bias = B1[rt * 4 + q];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) t[cb][rt] = mfma16(wa, xin[cb], bias);
      }
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        x2[cb][0] = relu_pack_bf16(t[cb][0], t[cb][1]);
        x2[cb][1] = relu_pack_bf16(t[cb][2], t[cb][3]);
      }
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        const bf16x8 wa0 = W2A[(rt * 2) * 64 + lane], wa1 = W2A[(rt * 2 + 1) * 64 + lane];
        const f32x4 bias = B2[rt * 4 + q];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) t[cb][rt] = mfma16(wa1, x2[cb][1], mfma16(wa0, x2[cb][0], bias));
      }
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        x3[cb][0] = relu_pack_bf16(t[cb][0], t[cb][1]);
        x3[cb][1] = relu_pack_bf16(t[cb][2], t[cb][3]);
      }
      f32x4 o[NCB];
      {
        const bf16x8 wa0 = W3A[lane], wa1 = W3A[64 + lane];
        const f32x4 b3 = q == 0 ? *reinterpret_cast<const f32x4*>(s_w + kPkB3) : z4;
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) o[cb] = mfma16(wa1, x3[cb][1], mfma16(wa0, x3[cb][0], b3));
      }

      // output adjoints (lanes q = 0, rows 0..2) as the B operand bo
      bf16x8 bo[NCB];
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        const bool own = ok[cb] && q == 0;
        float dO[3];
        if constexpr (kDistill) {
          // teacher: degree-3 SH colour + 0.5, clamped at 0; lane q adds coefficients 4q..4q+3
          float b[16];
          sh_basis(3, d[cb].x, d[cb].y, d[cb].z, b);
          float3 tc = make_float3(0.f, 0.f, 0.f);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            tc.x += b[4 * q + i] * sh[cb][3 * i];
            tc.y += b[4 * q + i] * sh[cb][3 * i + 1];
            tc.z += b[4 * q + i] * sh[cb][3 * i + 2];
          }
          const float tt[3] = {fmaxf(row_sum(tc.x) + 0.5f, 0.f), fmaxf(row_sum(tc.y) + 0.5f, 0.f),
                               fmaxf(row_sum(tc.z) + 0.5f, 0.f)};
#pragma unroll
          for (int e = 0; e < 3; ++e) {
            const float s = sigmoidf(o[cb][e]);
            const float err = own ? s - tt[e] : 0.f;
            sq_err += err * err;
            dO[e] = 2.f * err * inv3P * s * (1.f - s);
          }
        } else {
          const float dc[3] = {dcv[cb].x, dcv[cb].y, dcv[cb].z};
#pragma unroll
          for (int e = 0; e < 3; ++e) {
            const float s = sigmoidf(o[cb][e]);
            dO[e] = own ? dc[e] * s * (1.f - s) : 0.f;
          }
        }
        bo[cb] = q == 0 ? pack_bf16(f32x4{dO[0], dO[1], dO[2], 0.f}, z4) : zero_bf16x8();
      }

      // backward chains, each adjoint's weight-gradient products taken as soon
      // as its transposed tiles exist; column blocks 2p and 2p + 1 make one
      // K = 32 operand (slots e < 4: block 2p)
      bf16x8 b2[NCB][2];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        const bf16x8 wt = W3T[rt * 64 + lane];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) t[cb][rt] = mfma16(wt, bo[cb], z4);
      }
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        b2[cb][0] = relu_mask(x3[cb][0], pack_bf16(t[cb][0], t[cb][1]));
        b2[cb][1] = relu_mask(x3[cb][1], pack_bf16(t[cb][2], t[cb][3]));
      }
      // dW2 = dH2 H1^T, db2 (column 10 of gW1[.][1])
      {
        const bf16x8 sLo = s_sel[kSelLo][lane], sHi = s_sel[kSelHi][lane];
        bf16x8 d2[NP][4];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
#pragma unroll
          for (int nt = 0; nt < 4; ++nt) {
            const bf16x8& s = (nt & 1) ? sHi : sLo;
            d2[p][nt] = join_halves(tr_half(b2[2 * p][nt >> 1], s), tr_half(b2[2 * p + 1][nt >> 1], s));
          }
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
          const bf16x8& s = (mt & 1) ? sHi : sLo;
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            const bf16x8 h1 = join_halves(tr_half(x2[2 * p][mt >> 1], s), tr_half(x2[2 * p + 1][mt >> 1], s));
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) gW2[nt][mt] = mfma16(d2[p][nt], h1, gW2[nt][mt]);
          }
        }
        const bf16x8 one10 = s_sel[kSelOne10][lane];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
#pragma unroll
          for (int nt = 0; nt < 4; ++nt) gW1[nt][1] = mfma16(d2[p][nt], one10, gW1[nt][1]);
        }
      }
      bf16x8 b1[NCB][2];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        const bf16x8 wt0 = W2T[(rt * 2) * 64 + lane], wt1 = W2T[(rt * 2 + 1) * 64 + lane];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) t[cb][rt] = mfma16(wt1, b2[cb][1], mfma16(wt0, b2[cb][0], z4));
      }
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        b1[cb][0] = relu_mask(x2[cb][0], pack_bf16(t[cb][0], t[cb][1]));
        b1[cb][1] = relu_mask(x2[cb][1], pack_bf16(t[cb][2], t[cb][3]));
      }
      f32x4 dx[NCB][2];
#pragma unroll
      for (int rt = 0; rt < 2; ++rt) {
        const bf16x8 wt0 = W1T[(rt * 2) * 64 + lane], wt1 = W1T[(rt * 2 + 1) * 64 + lane];
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) dx[cb][rt] = mfma16(wt1, b1[cb][1], mfma16(wt0, b1[cb][0], z4));
      }
      // dW1 = dH1 X^T, db1 (column 9)
      {
        const bf16x8 sLo = s_sel[kSelLo][lane], sHi = s_sel[kSelHi][lane];
        const bf16x8 sIL = s_sel[kSelInLo][lane], sIH = s_sel[kSelInHi][lane], one9 = s_sel[kSelOne9][lane];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          const bf16x8 xt0 = join_halves(tr_half(xin[2 * p], sIL), tr_half(xin[2 * p + 1], sIL));
          const bf16x8 xt1 = join_halves(tr_half(xin[2 * p], sIH), tr_half(xin[2 * p + 1], sIH));
#pragma unroll
          for (int nt = 0; nt < 4; ++nt) {
            const bf16x8& s = (nt & 1) ? sHi : sLo;
            const bf16x8 d1 = join_halves(tr_half(b1[2 * p][nt >> 1], s), tr_half(b1[2 * p + 1][nt >> 1], s));
            gW1[nt][0] = mfma16(d1, xt0, gW1[nt][0]);
            gW1[nt][1] = mfma16(d1, xt1, gW1[nt][1]);
            gW1[nt][1] = mfma16(d1, one9, gW1[nt][1]);
          }
        }
      }
      // dW3^T = H2 dO^T (columns 11..13), db3 (rows 11..13 of column 14, tile 0)
      {
        const bf16x8 sLo = s_sel[kSelLo][lane], sHi = s_sel[kSelHi][lane], sDO = s_sel[kSelDO][lane];
        const bf16x8 one14 = s_sel[kSelOne14][lane];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          const bf16x8 dot_ = join_halves(tr_half(bo[2 * p], sDO), tr_half(bo[2 * p + 1], sDO));
#pragma unroll
          for (int nt = 0; nt < 4; ++nt) {
            const bf16x8& s = (nt & 1) ? sHi : sLo;
            const bf16x8 h2 = join_halves(tr_half(x3[2 * p][nt >> 1], s), tr_half(x3[2 * p + 1][nt >> 1], s));
            gW1[nt][1] = mfma16(h2, dot_, gW1[nt][1]);
          }
          gW1[0][1] = mfma16(dot_, one14, gW1[0][1]);
        }
      }

      // per-Gaussian adjoints: latent, and (render backward) the direction's
      // term of dL/dmean3D
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        dlat[cb].x += dx[cb][0][0];
        dlat[cb].y += dx[cb][0][1];
        dlat[cb].z += dx[cb][0][2];
        dlat[cb].w += dx[cb][0][3];
        if constexpr (!kDistill) {
          // basis adjoints: lane (c, q) holds basis 4q..4q+3 (q < 2) and basis 8
          // (q = 2, element 0); lane (c, 0) gathers all nine
          float gb[9];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            gb[e] = dx[cb][1][e];
            gb[4 + e] = from_row1(dx[cb][1][e]);
          }
          gb[8] = from_row2(dx[cb][1][0]);
          const float3 gd = sh2_basis_vjp(d[cb].x, d[cb].y, d[cb].z, gb);
          // d(v/|v|)/dv applied to gd (dnormvdv, auxiliary.h:134-144)
          const float dot = d[cb].x * gd.x + d[cb].y * gd.y + d[cb].z * gd.z;
          const float il = __builtin_amdgcn_rcpf(len[cb]);
          const bool use = ok[cb] && q == 0 && rad[cb] > 0 && len[cb] > 0.f;
          dmv[cb].x += use ? (gd.x - d[cb].x * dot) * il : 0.f;
          dmv[cb].y += use ? (gd.y - d[cb].y * dot) * il : 0.f;
          dmv[cb].z += use ? (gd.z - d[cb].z * dot) * il : 0.f;
          any_dir[cb] = any_dir[cb] || use;
        }
      }
    }
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
      if (ok[cb]) {
        *reinterpret_cast<float4*>(a.dL_dlatent + Gs[cb] * kLatent + 4 * q) = dlat[cb];
        if (!kDistill && any_dir[cb]) {
          a.dL_dmean3D[3 * Gs[cb]] = dm[cb].x + dmv[cb].x;
          a.dL_dmean3D[3 * Gs[cb] + 1] = dm[cb].y + dmv[cb].y;
          a.dL_dmean3D[3 * Gs[cb] + 2] = dm[cb].z + dmv[cb].z;
        }
      }
    }
  }

  // Block sum of the four waves' gradients in a fixed order (wave 0, + 1,
  // + 2, + 3), through LDS (the weights are no longer read), then wave 3
  // writes the block's partial row.
  __syncthreads();
  f32x4* red = reinterpret_cast<f32x4*>(s_w);  // 24 tiles x 64 lanes
  auto tile_ref = [&](int t) -> f32x4& {
    return t < 16 ? gW2[t >> 2][t & 3] : gW1[(t - 16) >> 1][t & 1];
  };
  for (int src = 0; src < 3; ++src) {
    if (w == src) {
#pragma unroll
      for (int t = 0; t < 24; ++t) {
        f32x4 v = tile_ref(t);
        if (src > 0) {
          const f32x4 p = red[t * 64 + lane];
#pragma unroll
          for (int e = 0; e < 4; ++e) v[e] = p[e] + v[e];
        }
        red[t * 64 + lane] = v;
      }
    }
    __syncthreads();
  }
  for (int o = 32; o > 0; o >>= 1) sq_err += __shfl_xor(sq_err, o);
  if (lane == 0) s_err[w] = sq_err;
  __syncthreads();
  if (w != 3) return;
  float* row = a.theta_partial + (size_t)blockIdx.x * kThetaStride;
#pragma unroll
  for (int t = 0; t < 24; ++t) {
    const f32x4 p = red[t * 64 + lane];
    f32x4& v = tile_ref(t);
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = p[e] + v[e];
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const int n = 16 * nt + 4 * q + e;  // hidden unit of this lane's D row
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) row[kThW2 + n * kHidden + 16 * mt + c] = gW2[nt][mt][e];
      row[kThW1 + n * kMlpIn + c] = gW1[nt][0][e];
      const float v = gW1[nt][1][e];
      if (16 + c < kMlpIn) row[kThW1 + n * kMlpIn + 16 + c] = v;
      else if (c == 9) row[kThB1 + n] = v;
      else if (c == 10) row[kThB2 + n] = v;
      else if (c < 11 + kMlpOut) row[kThW3 + (c - 11) * kHidden + n] = v;
      else if (c == 14 && nt == 0 && 4 * q + e >= 11 && 4 * q + e < 11 + kMlpOut) row[kThB3 + 4 * q + e - 11] = v;
    }
  }
  if (lane == 0) row[kThetaSize] = (s_err[0] + s_err[1]) + (s_err[2] + s_err[3]);
}
// Sum of rows [r0, r1) of column col with sixteen loads in flight, in a
// fixed order (sixteen strided partial sums, then added pairwise).
// The weight-gradient sum over the nb partial rows, in one launch: block b
// owns columns [64 b, 64 b + 64) (lane = column; column kThetaSize is the
// distillation error), wave w of its 16 sums rows w, w + 16, ... with all of
// its loads in flight together, and wave 0 adds the 16 wave sums.  The order
// is fixed, so the result is deterministic.
constexpr int kReduceWaves = 16;

__global__ __launch_bounds__(64 * kReduceWaves) void theta_reduce_kernel(const float* __restrict__ part, int nrows,
                                                                        float* __restrict__ grad, float err_scale,
                                                                        float* __restrict__ loss) {
  __shared__ float s_part[kReduceWaves][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int col = blockIdx.x * 64 + lane;  // < kThetaStride: every load stays in its row
  float acc = 0.f;
  for (int r0 = w; r0 < nrows; r0 += 16 * kReduceWaves) {
    float v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      // clamped rows and a select after the load: no branch holds a load back
      const int r = r0 + k * kReduceWaves;
      const float x = part[(size_t)min(r, nrows - 1) * kThetaStride + col];
      v[k] = r < nrows ? x : 0.f;
    }
#pragma unroll
    for (int h = 8; h > 0; h >>= 1)
#pragma unroll
      for (int k = 0; k < h; ++k) v[k] = v[k] + v[k + h];
    acc += v[0];
  }
  s_part[w][lane] = acc;
  __syncthreads();
  if (w != 0 || col > kThetaSize) return;
  float t[kReduceWaves];
#pragma unroll
  for (int k = 0; k < kReduceWaves; ++k) t[k] = s_part[k][lane];
#pragma unroll
  for (int h = kReduceWaves / 2; h > 0; h >>= 1)
#pragma unroll
    for (int k = 0; k < h; ++k) t[k] = t[k] + t[k + h];
  if (col < kThetaSize) grad[col] = t[0];
  else if (loss) *loss = t[0] * err_scale;
}

// ----------------------------------------------------------------------------
// Adam (torch::optim::Adam without weight decay or amsgrad, the optimizer of
// the reference's GaussianModel::trainingSetup, gaussian_model.cpp:485-518):
//   m = b1 m + (1 - b1) g;  v = b2 v + (1 - b2) g^2
//   p -= (lr / (1 - b1^t)) * m / (sqrt(v) / sqrt(1 - b2^t) + eps)
__global__ __launch_bounds__(256) void adam_kernel(AdamGroupTable tab, float beta1, float beta2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= tab.offset[tab.n]) return;
  int gi = 0;
  while (gi + 1 < tab.n && i >= tab.offset[gi + 1]) ++gi;
  const AdamGroupDev& g = tab.g[gi];
  const size_t j = i - tab.offset[gi];
  const float grad = g.grad[j];
  const float m = beta1 * g.exp_avg[j] + tab.one_minus_beta1 * grad;
  const float v = beta2 * g.exp_avg_sq[j] + tab.one_minus_beta2 * grad * grad;
  g.exp_avg[j] = m;
  g.exp_avg_sq[j] = v;
  const float p = g.param[j] - g.step_size * (m / (sqrtf(v) / g.bc2_sqrt + g.eps));
  g.param[j] = p;
  if (g.param_bf16) g.param_bf16[j] = __builtin_bit_cast(uint16_t, (__bf16)p);
}

__global__ __launch_bounds__(256) void to_bf16_kernel(const float* __restrict__ in, uint16_t* __restrict__ out,
                                                      size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = __builtin_bit_cast(uint16_t, (__bf16)in[i]);
}

// ----------------------------------------------------------------------------
// Launchers

int mlp_train_blocks(int P) {
  const int chunks = (P + 64 * kTrainCB - 1) / (64 * kTrainCB);  // 4 waves x 16 kTrainCB Gaussians
  return chunks < kMlpTrainBlocks ? (chunks > 0 ? chunks : 1) : kMlpTrainBlocks;
}

// Workspace: the weight-gradient partial rows (nb rows), then the
// active-Gaussian list (P rounded up to 256 entries) and its per-block counts.
static size_t mlp_theta_floats(int P) { return (size_t)mlp_train_blocks(P) * kThetaStride; }

size_t mlp_workspace_floats(int P) {
  const size_t segs = ((size_t)P + 255) / 256 * 4;  // the preprocess backward's waves
  return align_up(mlp_theta_floats(P) * sizeof(float)) / sizeof(float) + segs * 64 + segs;
}

uint32_t* mlp_act_list(float* ws, int P) {
  return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(ws) + align_up(mlp_theta_floats(P) * sizeof(float)));
}

uint32_t* mlp_act_count(float* ws, int P) { return mlp_act_list(ws, P) + ((size_t)P + 255) / 256 * 256; }

hipError_t launch_mlp_pack(const float* theta, void* packed, hipStream_t s) {
  hipLaunchKernelGGL(mlp_pack_kernel, dim3((1920 + 132 + 255) / 256), dim3(256), 0, s, theta, (char*)packed);
  return hipGetLastError();
}

hipError_t launch_mlp_eval(int N, const uint16_t* latents, const float* dirs, const void* packed, float* rgb,
                           hipStream_t s) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_eval_kernel, dim3((N + 255) / 256), dim3(256), 0, s, N, latents, dirs,
                     (const char*)packed, rgb);
  return hipGetLastError();
}

static hipError_t reduce_theta(int nb, float* ws, float* dL_dtheta, float err_scale, float* loss, hipStream_t s) {
  static_assert(((kThetaSize + 1 + 63) / 64) * 64 <= kThetaStride, "reduction columns stay inside a row");
  hipLaunchKernelGGL(theta_reduce_kernel, dim3((kThetaSize + 1 + 63) / 64), dim3(64 * kReduceWaves), 0, s,
                     (const float*)ws, nb, dL_dtheta, err_scale, loss);
  return hipGetLastError();
}

hipError_t launch_mlp_distill(int P, int M, const float* shs, const float* latents, const void* packed, uint32_t seed,
                              int step, int num_dirs, float* dL_dlatent, float* dL_dtheta, float* loss, float* ws,
                              hipStream_t s, const float* means3D, const float* view_box) {
  MlpTrainArgs a{};
  a.P = P;
  a.means3D = means3D;
  if (means3D) {
    a.view_lo = make_float3(view_box[0], view_box[1], view_box[2]);
    a.view_hi = make_float3(view_box[3], view_box[4], view_box[5]);
  }
  a.packed = (const char*)packed;
  a.latent_f32 = latents;
  a.shs = shs;
  a.M = M;
  a.seed = seed;
  a.step = step;
  a.num_dirs = num_dirs;
  a.V = 1;
  a.dL_dlatent = dL_dlatent;
  a.theta_partial = ws;
  const int nb = mlp_train_blocks(P);
  hipLaunchKernelGGL(mlp_train_kernel<true>, dim3(nb), dim3(256), 0, s, a);
  return reduce_theta(nb, ws, dL_dtheta, 1.f / (3.f * (float)P), loss, s);
}

hipError_t launch_mlp_backward(int P, const float* means3D, const float* campos, const int* radii,
                               const uint16_t* latents, const void* packed, const float* dL_dcolor, float* dL_dmean3D,
                               float* dL_dlatent, float* dL_dtheta, float* ws, hipStream_t s, int V,
                               bool listed) {
  if (P == 0) return hipSuccess;
  MlpTrainArgs a{};
  a.P = P;
  a.V = V;
  if (listed) {
    a.act_list = mlp_act_list(ws, P);
    a.act_count = mlp_act_count(ws, P);
  }
  a.packed = (const char*)packed;
  a.latent_bf16 = latents;
  a.means3D = means3D;
  a.campos = campos;
  a.radii = radii;
  a.dL_dcolor = dL_dcolor;
  a.dL_dmean3D = dL_dmean3D;
  a.dL_dlatent = dL_dlatent;
  a.theta_partial = ws;
  const int nb = mlp_train_blocks(P);
  hipLaunchKernelGGL(mlp_train_kernel<false>, dim3(nb), dim3(256), 0, s, a);
  return reduce_theta(nb, ws, dL_dtheta, 0.f, nullptr, s);
}

hipError_t launch_adam(const AdamGroupTable& tab, float beta1, float beta2, hipStream_t s) {
  const size_t n = tab.offset[tab.n];
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, tab, beta1, beta2);
  return hipGetLastError();
}

hipError_t launch_to_bf16(const float* in, uint16_t* out, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}

}  // namespace omnigs
