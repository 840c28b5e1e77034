// Colour decoder kernels: weight packing, standalone evaluation, the training
// pass (distillation from SH colours, or the backward of a render whose
// colours came from the decoder), the deterministic reduction of the weight
// gradient, Adam, and FP32 -> BF16 conversion.  Device math in mlp.hpp.
//
// Training pass layout: 256-thread blocks (4 waves) walk chunks of 64
// Gaussians.  Wave w runs the forward and backward chain of column block w
// (Gaussians 16w..16w+15 of the chunk), then stages its activations and
// adjoints as BF16 [unit][Gaussian] rows in LDS; after a barrier each wave
// computes a quarter of the weight-gradient tiles with K = the chunk's 64
// Gaussians (two MFMA K steps) and keeps them in registers for the whole
// kernel.  Bias gradients come from MFMAs against an all-ones B operand.
// Each block finally writes its partial gradient row; a two-stage kernel adds
// the rows in a fixed order.
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

// Unit direction of a world-space vector (0, 0, 1 for a zero vector).
__device__ __forceinline__ float3 unit_dir(float3 v, float* len_out) {
  const float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  *len_out = len;
  if (!(len > 0.f)) return make_float3(0.f, 0.f, 1.f);
  return make_float3(v.x / len, v.y / len, v.z / len);
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
  // distillation
  const float* shs;  // [P,M,3] teacher, degree 3
  int M;
  uint32_t seed;
  int step;
  int num_dirs;
  // outputs
  float* dL_dlatent;     // [P,16]
  float* theta_partial;  // [gridDim.x][kThetaStride]; column kThetaSize holds the block's squared error
};

constexpr int kStageRow = 64 + 8;  // bf16 per staged row (64 Gaussians + pad), 144 B
constexpr int kRowsX = 32, kRowsH = 64, kRowsO = 16;
constexpr int kStageRows = kRowsX + 4 * kRowsH + kRowsO;  // X, H1, H2, dH1, dH2, dO

template <bool kDistill>
__global__ __launch_bounds__(256) void mlp_train_kernel(MlpTrainArgs a) {
  __shared__ __attribute__((aligned(16))) char s_w[kMlpPackedBytes];
  __shared__ __attribute__((aligned(16))) __bf16 s_st[kStageRows][kStageRow];
  __shared__ float s_err[4];
  __bf16 (*sX)[kStageRow] = s_st;
  __bf16 (*sH1)[kStageRow] = s_st + kRowsX;
  __bf16 (*sH2)[kStageRow] = sH1 + kRowsH;
  __bf16 (*sD1)[kStageRow] = sH2 + kRowsH;
  __bf16 (*sD2)[kStageRow] = sD1 + kRowsH;
  __bf16 (*sO)[kStageRow] = sD2 + kRowsH;

  stage_weights(s_w, a.packed, kMlpPackedBytes);
  for (int i = threadIdx.x; i < kRowsO * kStageRow; i += blockDim.x) (&sO[0][0])[i] = (__bf16)0.f;
  __syncthreads();

  const int lane = threadIdx.x & 63, c = lane & 15, q = lane >> 4, w = threadIdx.x >> 6;
  const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
  bf16x8 ones;
#pragma unroll
  for (int e = 0; e < 8; ++e) ones[e] = (__bf16)1.f;
  // weight-gradient tiles owned by this wave: dW2 row tile w (4 column tiles),
  // dW1 row tile w (2), db1 / db2 row tile w, dW3 column tile w, db3 (wave 0)
  f32x4 gW2[4] = {z4, z4, z4, z4}, gW1[2] = {z4, z4}, gB1 = z4, gB2 = z4, gW3 = z4, gB3 = z4;
  float sq_err = 0.f;
  const float inv3P = 1.f / (3.f * (float)a.P);

  // Per-lane inputs of one chunk, loaded a chunk ahead so that the HBM
  // latency overlaps the previous chunk's MFMA work.
  struct ChunkIn {
    bf16x8 lat;    // latent inputs 8q..8q+7 (q < 2)
    float3 m;      // mean (render backward)
    float3 dc;     // dL/dcolor (render backward, q = 0)
    float3 dm;     // dL/dmean3D so far (render backward, q = 0)
    int rad;       // radius (render backward, q = 0)
    float sh[12];  // SH coefficients 4q..4q+3 (distillation)
  };
  auto load_chunk = [&](int chunk, ChunkIn& in) {
    const size_t G = (size_t)chunk * 64 + 16 * w + c;
    const bool ok = G < (size_t)a.P;
    if constexpr (kDistill) {
      in.lat = load_latent_half(a.latent_f32, G, q & 1, ok && q < 2);
      const float* sh = a.shs + (G * a.M + 4 * q) * 3;
#pragma unroll
      for (int i = 0; i < 12; ++i) in.sh[i] = (ok && 4 * q + i / 3 < a.M) ? sh[i] : 0.f;
    } else {
      in.lat = load_latent_half(a.latent_bf16, G, q & 1, ok && q < 2);
      in.m = ok ? make_float3(a.means3D[3 * G], a.means3D[3 * G + 1], a.means3D[3 * G + 2])
                : make_float3(0.f, 0.f, 0.f);
      const bool own = ok && q == 0;
      in.dc = own ? make_float3(a.dL_dcolor[3 * G], a.dL_dcolor[3 * G + 1], a.dL_dcolor[3 * G + 2])
                  : make_float3(0.f, 0.f, 0.f);
      in.dm = own ? make_float3(a.dL_dmean3D[3 * G], a.dL_dmean3D[3 * G + 1], a.dL_dmean3D[3 * G + 2])
                  : make_float3(0.f, 0.f, 0.f);
      in.rad = own ? a.radii[G] : 0;
    }
  };

  const int nchunks = (a.P + 63) / 64;
  const int nviews = kDistill ? 1 : a.V;
  ChunkIn cur{}, nxt{};
  if ((int)blockIdx.x < nchunks) load_chunk(blockIdx.x, cur);
  for (int chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    if (chunk + (int)gridDim.x < nchunks) load_chunk(chunk + gridDim.x, nxt);
    const size_t G = (size_t)chunk * 64 + 16 * w + c;
    const bool ok = G < (size_t)a.P;
    float4 dlat = make_float4(0.f, 0.f, 0.f, 0.f);
    float3 dmv = make_float3(0.f, 0.f, 0.f);
    bool any_dir = false;
    for (int vi = 0; vi < nviews; ++vi) {
      // view vi: items vi P + G (one view in distillation)
      const size_t item = (size_t)vi * a.P + G;
      const float3 dcv3 = vi == 0 ? cur.dc
                                  : ((ok && q == 0) ? make_float3(a.dL_dcolor[3 * item], a.dL_dcolor[3 * item + 1],
                                                                  a.dL_dcolor[3 * item + 2])
                                                    : make_float3(0.f, 0.f, 0.f));
      const int rad = vi == 0 ? cur.rad : ((ok && q == 0) ? a.radii[item] : 0);
      // direction (every q lane of Gaussian c holds it)
      float3 d = make_float3(0.f, 0.f, 1.f), v = make_float3(0.f, 0.f, 0.f);
      float len = 1.f;
      if (ok) {
        if (kDistill) {
          d = seeded_direction(a.seed, (uint32_t)G, distill_dir_index(a.seed, a.step, (uint32_t)G, a.num_dirs));
        } else {
          const float3 cp = make_float3(a.campos[3 * vi], a.campos[3 * vi + 1], a.campos[3 * vi + 2]);
          v = make_float3(cur.m.x - cp.x, cur.m.y - cp.y, cur.m.z - cp.z);
          d = unit_dir(v, &len);
        }
      }
      const bf16x8 xin = input_operand(q, cur.lat, d.x, d.y, d.z);
      MlpAct act;
      mlp_forward_block(s_w, lane, xin, act);

      // output adjoint (lanes q = 0, rows 0..2)
      f32x4 dO = z4;
      if (kDistill) {
        // teacher: degree-3 SH colour + 0.5, clamped at 0; lane q adds coefficients 4q..4q+3
        float b[16];
        sh_basis(3, d.x, d.y, d.z, b);
        float3 t = make_float3(0.f, 0.f, 0.f);
  #pragma unroll
        for (int i = 0; i < 4; ++i) {
          t.x += b[4 * q + i] * cur.sh[3 * i];
          t.y += b[4 * q + i] * cur.sh[3 * i + 1];
          t.z += b[4 * q + i] * cur.sh[3 * i + 2];
        }
        t.x += __shfl_xor(t.x, 16);
        t.y += __shfl_xor(t.y, 16);
        t.z += __shfl_xor(t.z, 16);
        t.x += __shfl_xor(t.x, 32);
        t.y += __shfl_xor(t.y, 32);
        t.z += __shfl_xor(t.z, 32);
        t = make_float3(fmaxf(t.x + 0.5f, 0.f), fmaxf(t.y + 0.5f, 0.f), fmaxf(t.z + 0.5f, 0.f));
        if (ok && q == 0) {
          const float tt[3] = {t.x, t.y, t.z};
  #pragma unroll
          for (int e = 0; e < 3; ++e) {
            const float s = sigmoidf(act.o[e]);
            const float err = s - tt[e];
            sq_err += err * err;
            dO[e] = 2.f * err * inv3P * s * (1.f - s);
          }
        }
      } else if (ok && q == 0) {
        const float dcv[3] = {dcv3.x, dcv3.y, dcv3.z};
  #pragma unroll
        for (int e = 0; e < 3; ++e) {
          const float s = sigmoidf(act.o[e]);
          dO[e] = dcv[e] * s * (1.f - s);
        }
      }
      MlpGrad gr;
      mlp_backward_block(s_w, lane, act, dO, gr);

      dlat.x += gr.dx[0][0];
      dlat.y += gr.dx[0][1];
      dlat.z += gr.dx[0][2];
      dlat.w += gr.dx[0][3];
      if (!kDistill) {
        // basis adjoint 4q..4q+3 held here -> partial direction gradient, summed over q
        float gb[9];
  #pragma unroll
        for (int i = 0; i < 9; ++i) gb[i] = 0.f;
  #pragma unroll
        for (int e = 0; e < 4; ++e)
          if (4 * q + e < 9) gb[4 * q + e] = gr.dx[1][e];
        float3 gd = sh2_basis_vjp(d.x, d.y, d.z, gb);
        gd.x += __shfl_xor(gd.x, 16);
        gd.y += __shfl_xor(gd.y, 16);
        gd.z += __shfl_xor(gd.z, 16);
        gd.x += __shfl_xor(gd.x, 32);
        gd.y += __shfl_xor(gd.y, 32);
        gd.z += __shfl_xor(gd.z, 32);
        if (ok && q == 0 && rad > 0 && len > 0.f) {
          // d(v/|v|)/dv applied to gd (dnormvdv, auxiliary.h:134-144)
          const float dot = d.x * gd.x + d.y * gd.y + d.z * gd.z;
          const float il = 1.f / len;
          dmv.x += (gd.x - d.x * dot) * il;
          dmv.y += (gd.y - d.y * dot) * il;
          dmv.z += (gd.z - d.z * dot) * il;
          any_dir = true;
        }
      }

      // stage this column block (columns 16w + c) for the weight gradients
      const int col = 16 * w + c;
  #pragma unroll
      for (int e = 0; e < 8; ++e) sX[8 * q + e][col] = xin[e];
  #pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
  #pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = 16 * rt + 4 * q + e;
          sH1[r][col] = (__bf16)act.h1[rt][e];
          sH2[r][col] = (__bf16)act.h2[rt][e];
          sD1[r][col] = (__bf16)gr.dh1[rt][e];
          sD2[r][col] = (__bf16)gr.dh2[rt][e];
        }
      }
      if (q == 0) {
  #pragma unroll
        for (int e = 0; e < 3; ++e) sO[e][col] = (__bf16)dO[e];
      }
      __syncthreads();
  #pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        const int k0 = 32 * ks + 8 * q;
        const bf16x8 aD2 = *reinterpret_cast<const bf16x8*>(&sD2[16 * w + c][k0]);
        const bf16x8 aD1 = *reinterpret_cast<const bf16x8*>(&sD1[16 * w + c][k0]);
  #pragma unroll
        for (int ct = 0; ct < 4; ++ct)
          gW2[ct] = mfma16(aD2, *reinterpret_cast<const bf16x8*>(&sH1[16 * ct + c][k0]), gW2[ct]);
  #pragma unroll
        for (int ct = 0; ct < 2; ++ct)
          gW1[ct] = mfma16(aD1, *reinterpret_cast<const bf16x8*>(&sX[16 * ct + c][k0]), gW1[ct]);
        gB1 = mfma16(aD1, ones, gB1);
        gB2 = mfma16(aD2, ones, gB2);
        const bf16x8 aO = *reinterpret_cast<const bf16x8*>(&sO[c][k0]);
        gW3 = mfma16(aO, *reinterpret_cast<const bf16x8*>(&sH2[16 * w + c][k0]), gW3);
        if (w == 0) gB3 = mfma16(aO, ones, gB3);
      }
      __syncthreads();
    }
    if (ok) *reinterpret_cast<float4*>(a.dL_dlatent + G * kLatent + 4 * q) = dlat;
    if (!kDistill && ok && q == 0 && any_dir) {
      a.dL_dmean3D[3 * G] = cur.dm.x + dmv.x;
      a.dL_dmean3D[3 * G + 1] = cur.dm.y + dmv.y;
      a.dL_dmean3D[3 * G + 2] = cur.dm.z + dmv.z;
    }
    cur = nxt;
  }

  // block partial row (each wave writes its own tiles)
  float* row = a.theta_partial + (size_t)blockIdx.x * kThetaStride;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int n = 16 * w + 4 * q + e;  // hidden unit of this lane's D rows
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) row[kThW2 + n * kHidden + 16 * ct + c] = gW2[ct][e];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
      if (16 * ct + c < kMlpIn) row[kThW1 + n * kMlpIn + 16 * ct + c] = gW1[ct][e];
    if (c == 0) {
      row[kThB1 + n] = gB1[e];
      row[kThB2 + n] = gB2[e];
    }
    if (q == 0 && e < kMlpOut) {
      row[kThW3 + e * kHidden + 16 * w + c] = gW3[e];
      if (w == 0 && c == 0) row[kThB3 + e] = gB3[e];
    }
  }
  for (int o = 32; o > 0; o >>= 1) sq_err += __shfl_xor(sq_err, o);
  if (lane == 0) s_err[w] = sq_err;
  __syncthreads();
  if (threadIdx.x == 0) row[kThetaSize] = (s_err[0] + s_err[1]) + (s_err[2] + s_err[3]);
}

// Stage 1: sums rows [r0, r0 + rows) of each column.  Stage 2 then sums the
// stage-1 rows into out (columns < kThetaSize) and scales the error column.
__global__ __launch_bounds__(256) void theta_reduce1_kernel(const float* __restrict__ part, int nrows, int rows,
                                                            float* __restrict__ out1) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col > kThetaSize) return;
  const int r0 = blockIdx.y * rows, r1 = min(nrows, r0 + rows);
  float s = 0.f;
  for (int r = r0; r < r1; ++r) s += part[(size_t)r * kThetaStride + col];
  out1[(size_t)blockIdx.y * kThetaStride + col] = s;
}

__global__ __launch_bounds__(256) void theta_reduce2_kernel(const float* __restrict__ in1, int nrows,
                                                            float* __restrict__ grad, float err_scale,
                                                            float* __restrict__ loss) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col > kThetaSize) return;
  float s = 0.f;
  for (int r = 0; r < nrows; ++r) s += in1[(size_t)r * kThetaStride + col];
  if (col < kThetaSize) grad[col] = s;
  else if (loss) *loss = s * err_scale;
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
  const int chunks = (P + 63) / 64;
  return chunks < kMlpTrainBlocks ? (chunks > 0 ? chunks : 1) : kMlpTrainBlocks;
}

size_t mlp_workspace_floats(int P) {
  const int nb = mlp_train_blocks(P);
  const int n1 = (nb + kThetaReduceRows - 1) / kThetaReduceRows;
  return (size_t)(nb + n1) * kThetaStride;
}

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
  const int n1 = (nb + kThetaReduceRows - 1) / kThetaReduceRows;
  float* stage1 = ws + (size_t)nb * kThetaStride;
  const dim3 g1((kThetaSize + 1 + 255) / 256, n1);
  hipLaunchKernelGGL(theta_reduce1_kernel, g1, dim3(256), 0, s, (const float*)ws, nb, kThetaReduceRows, stage1);
  hipLaunchKernelGGL(theta_reduce2_kernel, dim3((kThetaSize + 1 + 255) / 256), dim3(256), 0, s,
                     (const float*)stage1, n1, dL_dtheta, err_scale, loss);
  return hipGetLastError();
}

hipError_t launch_mlp_distill(int P, int M, const float* shs, const float* latents, const void* packed, uint32_t seed,
                              int step, int num_dirs, float* dL_dlatent, float* dL_dtheta, float* loss, float* ws,
                              hipStream_t s) {
  MlpTrainArgs a{};
  a.P = P;
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
                               float* dL_dlatent, float* dL_dtheta, float* ws, hipStream_t s, int V) {
  if (P == 0) return hipSuccess;
  MlpTrainArgs a{};
  a.P = P;
  a.V = V;
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
