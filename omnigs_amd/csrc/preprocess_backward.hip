// Backward preprocess, one thread per Gaussian, fusing what the reference
// splits into computeCov2DLonLatCUDA (backward.cu:297-485) and
// preprocessLonLatCUDA (backward.cu:613-669):
//   dL/dconic -> dL/d(a,b,c) -> dL/dcov3D and dL/dJ -> dL/dt via the second
//     derivatives of the lon/lat projection (supp.pdf Eq. 43-55),
//   + the 2D-mean term (dL_dmean2D is in NDC units, scaled by 2/W, 2/H),
//   + the SH view-direction term, dL/dsh, and dL/dscale, dL/drot
//     (w.r.t. the quaternion as given; normalisation is the caller's).
// The reference's dpx_dt / dpy_dt round trip through HBM disappears: the
// Jacobian stays in registers (it is written only when asked for).
// HBM-bound: ~290 B read and ~250 B written per visible Gaussian.
#include "kernels.hpp"
#include "lonlat_math.hpp"
#include "grad_gather.hpp"

namespace omnigs {

// d(basis_i)/d(x,y,z) of the degree <= 3 real SH basis (unused tail = 0).
__device__ __forceinline__ void sh_basis_grad(int deg, float x, float y, float z, float bx[16], float by[16],
                                              float bz[16]) {
#pragma unroll
  for (int i = 0; i < 16; ++i) bx[i] = by[i] = bz[i] = 0.f;
  if (deg > 0) {
    by[1] = -kSH1;
    bz[2] = kSH1;
    bx[3] = -kSH1;
    if (deg > 1) {
      const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
      bx[4] = kSH2_0 * y;          by[4] = kSH2_0 * x;
      by[5] = kSH2_1 * z;          bz[5] = kSH2_1 * y;
      bx[6] = kSH2_2 * -2.f * x;   by[6] = kSH2_2 * -2.f * y;   bz[6] = kSH2_2 * 4.f * z;
      bx[7] = kSH2_3 * z;          bz[7] = kSH2_3 * x;
      bx[8] = kSH2_4 * 2.f * x;    by[8] = kSH2_4 * -2.f * y;
      if (deg > 2) {
        bx[9] = kSH3_0 * 6.f * xy;                  by[9] = kSH3_0 * 3.f * (xx - yy);
        bx[10] = kSH3_1 * yz;                       by[10] = kSH3_1 * xz;          bz[10] = kSH3_1 * xy;
        bx[11] = kSH3_2 * -2.f * xy;                by[11] = kSH3_2 * (4.f * zz - xx - 3.f * yy);
        bz[11] = kSH3_2 * 8.f * yz;
        bx[12] = kSH3_3 * -6.f * xz;                by[12] = kSH3_3 * -6.f * yz;
        bz[12] = kSH3_3 * 3.f * (2.f * zz - xx - yy);
        bx[13] = kSH3_4 * (4.f * zz - 3.f * xx - yy); by[13] = kSH3_4 * -2.f * xy; bz[13] = kSH3_4 * 8.f * xz;
        bx[14] = kSH3_5 * 2.f * xz;                 by[14] = kSH3_5 * -2.f * yz;   bz[14] = kSH3_5 * (xx - yy);
        bx[15] = kSH3_6 * 3.f * (xx - yy);          by[15] = kSH3_6 * -6.f * xy;
      }
    }
  }
}

// Covariance and 2D-mean terms of one view (computeCov2DLonLatCUDA, and the
// 2D-mean part of preprocessLonLatCUDA): dL/dconic -> dcov[6] and the world
// mean gradient; j receives the lon/lat Jacobian.
__device__ __forceinline__ void cov_mean_backward(const View& v, float3 p, const Sym3& S, int W, int H, float4 dcon4,
                                                  float dmx, float dmy, float dcov[6], float3& dmean, Jac& j) {
  const float3 t = to_camera(v, p);
  j = lonlat_jacobian(t, W, H);
  const Proj2 A = project_rows(j, v);
  const float ca = quad_form(S, A.ax, A.ax) + 0.3f;
  const float cb = quad_form(S, A.ax, A.ay);
  const float cc = quad_form(S, A.ay, A.ay) + 0.3f;

  const float dcx = dcon4.x, dcy = dcon4.y, dcz = dcon4.w;
  const float denom = ca * cc - cb * cb;
  const float d2inv = 1.0f / ((denom * denom) + 1e-7f);
  float dL_da = 0.f, dL_db = 0.f, dL_dc = 0.f;
#pragma unroll
  for (int k = 0; k < 6; ++k) dcov[k] = 0.f;
  const float* ax = A.ax;
  const float* ay = A.ay;
  if (d2inv != 0.f) {
    dL_da = d2inv * (-cc * cc * dcx + 2.f * cb * cc * dcy + (denom - ca * cc) * dcz);
    dL_dc = d2inv * (-ca * ca * dcz + 2.f * ca * cb * dcy + (denom - ca * cc) * dcx);
    dL_db = d2inv * 2.f * (cb * cc * dcx - (denom + 2.f * cb * cb) * dcy + ca * cb * dcz);
    dcov[0] = ax[0] * ax[0] * dL_da + ax[0] * ay[0] * dL_db + ay[0] * ay[0] * dL_dc;
    dcov[3] = ax[1] * ax[1] * dL_da + ax[1] * ay[1] * dL_db + ay[1] * ay[1] * dL_dc;
    dcov[5] = ax[2] * ax[2] * dL_da + ax[2] * ay[2] * dL_db + ay[2] * ay[2] * dL_dc;
    dcov[1] = 2.f * ax[0] * ax[1] * dL_da + (ax[0] * ay[1] + ax[1] * ay[0]) * dL_db + 2.f * ay[0] * ay[1] * dL_dc;
    dcov[2] = 2.f * ax[0] * ax[2] * dL_da + (ax[0] * ay[2] + ax[2] * ay[0]) * dL_db + 2.f * ay[0] * ay[2] * dL_dc;
    dcov[4] = 2.f * ax[2] * ax[1] * dL_da + (ax[1] * ay[2] + ax[2] * ay[1]) * dL_db + 2.f * ay[1] * ay[2] * dL_dc;
  }

  // dL/dA (rows of J R), then dL/dJ = dL/dA R^T
  float dAx[3], dAy[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float sx = S.at(k, 0) * ax[0] + S.at(k, 1) * ax[1] + S.at(k, 2) * ax[2];
    const float sy = S.at(k, 0) * ay[0] + S.at(k, 1) * ay[1] + S.at(k, 2) * ay[2];
    dAx[k] = 2.f * sx * dL_da + sy * dL_db;
    dAy[k] = 2.f * sy * dL_dc + sx * dL_db;
  }
  const float dJ00 = v.r[0][0] * dAx[0] + v.r[0][1] * dAx[1] + v.r[0][2] * dAx[2];
  const float dJ02 = v.r[2][0] * dAx[0] + v.r[2][1] * dAx[1] + v.r[2][2] * dAx[2];
  const float dJ10 = v.r[0][0] * dAy[0] + v.r[0][1] * dAy[1] + v.r[0][2] * dAy[2];
  const float dJ11 = v.r[1][0] * dAy[0] + v.r[1][1] * dAy[1] + v.r[1][2] * dAy[2];
  const float dJ12 = v.r[2][0] * dAy[0] + v.r[2][1] * dAy[1] + v.r[2][2] * dAy[2];

  // second derivatives of (px, py) w.r.t. t
  const float txtx = t.x * t.x, tyty = t.y * t.y, tztz = t.z * t.z;
  const float rxz2 = txtx + tztz;
  const float inv_rxz2 = 1.0f / (rxz2 + 1e-7f);
  const float inv_rxz4 = inv_rxz2 * inv_rxz2;
  const float rxz = sqrtf(rxz2);
  const float inv_rxz = 1.0f / (rxz + 1e-7f);
  const float r2 = rxz2 + tyty;
  const float inv_r2 = 1.0f / (r2 + 1e-7f);
  const float inv_r4 = inv_r2 * inv_r2;
  const float inv_rxz_r4 = inv_rxz * inv_r4;
  const float inv_rxz3_r4 = inv_rxz2 * inv_rxz_r4;
  const float Hp = j.hpi, Wp = j.w2pi;
  const float temp1 = Hp * (tyty - rxz2) * inv_rxz_r4;
  const float temp2 = Hp * (t.x * t.y * t.z) * (r2 + 2.f * rxz2) * inv_rxz3_r4;
  const float temp3 = Wp * (txtx - tztz) * inv_rxz4;
  const float temp4 = Wp * 2.f * t.x * t.z * inv_rxz4;
  const float temp5 = Hp * t.y * inv_rxz3_r4;
  const float dtx = -dJ00 * temp4 + dJ02 * temp3 + dJ10 * temp5 * (2.f * txtx * rxz2 - tztz * r2) +
                    dJ11 * t.x * temp1 + dJ12 * temp2;
  const float dty = dJ10 * t.x * temp1 - dJ11 * Hp * 2.f * rxz * t.y * inv_r4 + dJ12 * t.z * temp1;
  const float dtz = dJ00 * temp3 + dJ02 * temp4 + dJ10 * temp2 + dJ11 * t.z * temp1 +
                    dJ12 * temp5 * (2.f * tztz * rxz2 - txtx * r2);
  dmean = to_world_dir(v, make_float3(dtx, dty, dtz));

  // ---- 2D-mean term (NDC -> pixel: x 2/W, 2/H) ----
  const float dL_dpx = dmx * (2.0f / (float)W);
  const float dL_dpy = dmy * (2.0f / (float)H);
  const float3 dt2 = make_float3(dL_dpx * j.px_x + dL_dpy * j.py_x, dL_dpy * j.py_y, dL_dpx * j.px_z + dL_dpy * j.py_z);
  const float3 dm2 = to_world_dir(v, dt2);
  dmean.x += dm2.x;
  dmean.y += dm2.y;
  dmean.z += dm2.z;
}

// SH term of one view (computeColorFromSH backward, backward.cu:30-151):
// f[3i + ch] = basis_i dL/dcolor_ch (clamped channels give 0); returns the
// view-direction gradient on the mean.  c: the nc coefficients in registers.
__device__ __forceinline__ float3 sh_backward(int D, int nc, const float3 c[16], float3 p, float3 campos, uint8_t cm,
                                              float3 dcol, float f[48]) {
  const float3 d0 = make_float3(p.x - campos.x, p.y - campos.y, p.z - campos.z);
  const float len = sqrtf(d0.x * d0.x + d0.y * d0.y + d0.z * d0.z);
  const float x = d0.x / len, y = d0.y / len, z = d0.z / len;
  const float gx = (cm & 1) ? 0.f : dcol.x;
  const float gy = (cm & 2) ? 0.f : dcol.y;
  const float gz = (cm & 4) ? 0.f : dcol.z;
  float b[16];
  sh_basis(D, x, y, z, b);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const bool on = i < nc;
    f[3 * i] = on ? b[i] * gx : 0.f;
    f[3 * i + 1] = on ? b[i] * gy : 0.f;
    f[3 * i + 2] = on ? b[i] * gz : 0.f;
  }
  float3 dm = make_float3(0.f, 0.f, 0.f);
  if (D > 0) {
    float bx[16], by[16], bz[16];
    sh_basis_grad(D, x, y, z, bx, by, bz);
    float ddx = 0.f, ddy = 0.f, ddz = 0.f;
#pragma unroll
    for (int i = 1; i < 16; ++i) {
      if (i < nc) {
        const float sg = c[i].x * gx + c[i].y * gy + c[i].z * gz;
        ddx += bx[i] * sg;
        ddy += by[i] * sg;
        ddz += bz[i] * sg;
      }
    }
    // derivative of the normalisation d0/|d0| (dnormvdv, auxiliary.h:134-144)
    const float sum2 = d0.x * d0.x + d0.y * d0.y + d0.z * d0.z;
    const float invsum32 = 1.0f / sqrtf(sum2 * sum2 * sum2);
    dm.x = ((sum2 - d0.x * d0.x) * ddx - d0.y * d0.x * ddy - d0.z * d0.x * ddz) * invsum32;
    dm.y = (-d0.x * d0.y * ddx + (sum2 - d0.y * d0.y) * ddy - d0.z * d0.y * ddz) * invsum32;
    dm.z = (-d0.x * d0.z * ddx - d0.y * d0.z * ddy + (sum2 - d0.z * d0.z) * ddz) * invsum32;
  }
  return dm;
}

// computeCov3D backward (backward.cu:489-552): dL/dscale (w.r.t. s, ignoring
// the modifier) and dL/dq (w.r.t. the quaternion as given).
__device__ __forceinline__ void scale_rot_backward(const float dcov[6], float3 s, float4 q, float mod, float dsc[3],
                                                   float4& dq) {
  float R[3][3];
  quat_to_rot(q, R);
  const float sv[3] = {mod * s.x, mod * s.y, mod * s.z};
  // dL/dSigma with halved off-diagonals
  float D[3][3];
  D[0][0] = dcov[0];
  D[1][1] = dcov[3];
  D[2][2] = dcov[5];
  D[0][1] = D[1][0] = 0.5f * dcov[1];
  D[0][2] = D[2][0] = 0.5f * dcov[2];
  D[1][2] = D[2][1] = 0.5f * dcov[4];
  // M[k][i] = s_k R[i][k];  dM = 2 M D;  dscale_k = sum_i R[i][k] dM[k][i]
  float E[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float acc = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float dMki = 2.f * (sv[k] * R[0][k] * D[0][i] + sv[k] * R[1][k] * D[1][i] + sv[k] * R[2][k] * D[2][i]);
      acc += R[i][k] * dMki;
      E[k][i] = sv[k] * dMki;
    }
    dsc[k] = acc;
  }
  const float w = q.x, x = q.y, y = q.z, z = q.w;
  dq.x = 2.f * z * (E[0][1] - E[1][0]) + 2.f * y * (E[2][0] - E[0][2]) + 2.f * x * (E[1][2] - E[2][1]);
  dq.y = 2.f * y * (E[1][0] + E[0][1]) + 2.f * z * (E[2][0] + E[0][2]) + 2.f * w * (E[1][2] - E[2][1]) -
         4.f * x * (E[2][2] + E[1][1]);
  dq.z = 2.f * x * (E[1][0] + E[0][1]) + 2.f * w * (E[2][0] - E[0][2]) + 2.f * z * (E[1][2] + E[2][1]) -
         4.f * y * (E[2][2] + E[0][0]);
  dq.w = 2.f * w * (E[0][1] - E[1][0]) + 2.f * x * (E[2][0] + E[0][2]) + 2.f * y * (E[1][2] + E[2][1]) -
         4.f * z * (E[1][1] + E[0][0]);
}

// Writes one Gaussian's SH gradient row (M coefficients, zero past 16).
__device__ __forceinline__ void store_dsh(float* dsh, int M, const float f[48]) {
  if (M == 16) {
    float4* d4 = reinterpret_cast<float4*>(dsh);
#pragma unroll
    for (int i = 0; i < 12; ++i) d4[i] = make_float4(f[4 * i], f[4 * i + 1], f[4 * i + 2], f[4 * i + 3]);
  } else {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if (i < M) {
        dsh[3 * i] = f[3 * i];
        dsh[3 * i + 1] = f[3 * i + 1];
        dsh[3 * i + 2] = f[3 * i + 2];
      }
    }
    for (int k = 48; k < 3 * M; ++k) dsh[k] = 0.f;
  }
}

// Wave-cooperative SH input and output for M = 16 (48 floats per Gaussian):
// the coefficients of half a wave's Gaussians at a time pass through an LDS
// tile, so that the global reads and writes are contiguous float4 runs (a
// lane's own 192-byte row would put each float4 of one instruction in a
// different cache line).  Every lane of the wave must call these; rows past
// P are neither read nor written.
constexpr int kShRow = 52;  // floats per LDS row: 48 + 4, so 8 lanes' 16-byte accesses hit distinct banks
constexpr int kShTile = 32 * kShRow;

__device__ __forceinline__ void sh_load_wave(const float* __restrict__ shs, int P, int wave_base, int lane,
                                             float* tile, float3 c[16]) {
  // both halves' loads are issued before the first LDS round (one memory latency)
  float4 ld[2][6];
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int g0 = wave_base + 32 * half;
    const float4* src = reinterpret_cast<const float4*>(shs) + (size_t)g0 * 12;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int ch = k * kWave + lane, r = ch / 12;
      ld[half][k] = g0 + r < P ? src[ch] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int ch = k * kWave + lane, r = ch / 12, col = ch % 12;
      *reinterpret_cast<float4*>(tile + r * kShRow + 4 * col) = ld[half][k];
    }
    wave_lds_sync();
    if ((lane >> 5) == half) {
      const float* row = tile + (lane & 31) * kShRow;
      float f[48];
#pragma unroll
      for (int i = 0; i < 12; ++i) {
        const float4 v = *reinterpret_cast<const float4*>(row + 4 * i);
        f[4 * i] = v.x; f[4 * i + 1] = v.y; f[4 * i + 2] = v.z; f[4 * i + 3] = v.w;
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) c[i] = make_float3(f[3 * i], f[3 * i + 1], f[3 * i + 2]);
    }
  }
}

__device__ __forceinline__ void sh_store_wave(float* __restrict__ dsh, int P, int wave_base, int lane, float* tile,
                                              const float f[48]) {
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int g0 = wave_base + 32 * half;
    float4* dst = reinterpret_cast<float4*>(dsh) + (size_t)g0 * 12;
    wave_lds_sync();
    if ((lane >> 5) == half) {
      float* row = tile + (lane & 31) * kShRow;
#pragma unroll
      for (int i = 0; i < 12; ++i)
        *reinterpret_cast<float4*>(row + 4 * i) = make_float4(f[4 * i], f[4 * i + 1], f[4 * i + 2], f[4 * i + 3]);
    }
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int ch = k * kWave + lane, r = ch / 12, col = ch % 12;
      if (g0 + r < P) dst[ch] = *reinterpret_cast<const float4*>(tile + r * kShRow + 4 * col);
    }
  }
}

// kGather: the 2D gradients are first summed from the render backward's
// records (every lane of a wave takes part) and written out, then used here
// from registers.  With M = 16 SH coefficients the SH input and output are
// wave-cooperative (sh_load_wave / sh_store_wave), so every lane runs to the
// end, with `live` predicating the per-Gaussian work.
template <bool kGather>
__global__ __launch_bounds__(256) void preprocess_backward_kernel(PreprocessBackwardArgs a) {
  __shared__ __attribute__((aligned(16))) float s_sh[256 / kWave][kShTile];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  float gs[kGradTerms];
  if constexpr (kGather) {
    gather_sum(a.gather, idx, idx < a.P, gs);
    if (idx < a.P) store_gathered((size_t)idx, gs, a.out_dmean2D, a.out_dconic, a.out_dopacity, a.out_dcolor);
  }
  if (a.act_list) {
    // The decoder's work list: this wave's Gaussians with a colour gradient,
    // in index order, from one ballot (no block barrier).
    float3 dc = make_float3(0.f, 0.f, 0.f);
    if (idx < a.P)
      dc = kGather ? make_float3(gs[0], gs[1], gs[2])
                   : make_float3(a.dL_dcolor[3 * idx], a.dL_dcolor[3 * idx + 1], a.dL_dcolor[3 * idx + 2]);
    const bool act = idx < a.P && (dc.x != 0.f || dc.y != 0.f || dc.z != 0.f);
    const uint64_t bal = __ballot(act);
    const size_t wave_base = (size_t)idx - (size_t)lane;
    if (act) a.act_list[wave_base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = (uint32_t)idx;
    if (lane == 0) a.act_count[wave_base / kWave] = (uint32_t)__popcll(bal);
    if (idx < a.P && !act) {
      float4* dl = reinterpret_cast<float4*>(a.dL_dlatent + (size_t)idx * 16);
      dl[0] = dl[1] = dl[2] = dl[3] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  const int M = a.M;
  const bool in = idx < a.P;
  const bool staged = a.shs && a.dL_dsh && M == 16;  // uniform over the launch
  if (!staged && !in) return;
  const bool live = in && a.radii[idx] > 0;

  if (!live && in) {
    a.dL_dmean3D[3 * idx] = 0.f;
    a.dL_dmean3D[3 * idx + 1] = 0.f;
    a.dL_dmean3D[3 * idx + 2] = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) a.dL_dcov3D[6 * idx + k] = 0.f;
    if (a.dL_dsh && !staged)
      for (int k = 0; k < 3 * M; ++k) a.dL_dsh[(size_t)3 * M * idx + k] = 0.f;
    if (a.dL_dscale) {
#pragma unroll
      for (int k = 0; k < 3; ++k) a.dL_dscale[3 * idx + k] = 0.f;
    }
    if (a.dL_drot) reinterpret_cast<float4*>(a.dL_drot)[idx] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (a.dpx_dt) {
#pragma unroll
      for (int k = 0; k < 3; ++k) a.dpx_dt[3 * idx + k] = 0.f;
    }
    if (a.dpy_dt) {
#pragma unroll
      for (int k = 0; k < 3; ++k) a.dpy_dt[3 * idx + k] = 0.f;
    }
  }
  if (!staged && !live) return;

  const int wave_base = idx - lane;
  float3 c[16];
  if (staged) sh_load_wave(a.shs, a.P, wave_base, lane, s_sh[threadIdx.x / kWave], c);
  float f[48];
#pragma unroll
  for (int k = 0; k < 48; ++k) f[k] = 0.f;

  if (live) {
    const View v = load_view(a.viewmatrix);
    const float3 p = make_float3(a.means3D[3 * idx], a.means3D[3 * idx + 1], a.means3D[3 * idx + 2]);

    // ---- 3D covariance (recomputed; identical to the forward's) ----
    Sym3 S;
    float3 s = make_float3(0.f, 0.f, 0.f);
    float4 q = make_float4(1.f, 0.f, 0.f, 0.f);
    if (a.cov3D_precomp) {
#pragma unroll
      for (int k = 0; k < 6; ++k) S.m[k] = a.cov3D_precomp[6 * idx + k];
    } else {
      s = make_float3(a.scales[3 * idx], a.scales[3 * idx + 1], a.scales[3 * idx + 2]);
      q = reinterpret_cast<const float4*>(a.rotations)[idx];
      S = cov3d_from(s, a.scale_modifier, q);
    }

    float dcov[6];
    float3 dmean;
    Jac j;
    if constexpr (kGather)
      cov_mean_backward(v, p, S, a.W, a.H, make_float4(gs[5], gs[6], 0.f, gs[7]), gs[3], gs[4], dcov, dmean, j);
    else
      cov_mean_backward(v, p, S, a.W, a.H, reinterpret_cast<const float4*>(a.dL_dconic)[idx], a.dL_dmean2D[3 * idx],
                        a.dL_dmean2D[3 * idx + 1], dcov, dmean, j);
    if (a.dpx_dt) {
      a.dpx_dt[3 * idx] = j.px_x;
      a.dpx_dt[3 * idx + 1] = 0.f;
      a.dpx_dt[3 * idx + 2] = j.px_z;
    }
    if (a.dpy_dt) {
      a.dpy_dt[3 * idx] = j.py_x;
      a.dpy_dt[3 * idx + 1] = j.py_y;
      a.dpy_dt[3 * idx + 2] = j.py_z;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) a.dL_dcov3D[6 * idx + k] = dcov[k];

    // ---- SH term ----
    if (a.shs) {
      const int nc = min(M, sh_coeffs(a.D));
      if (!staged) load_sh(a.shs, M, idx, nc, c);
      const float3 dcol = kGather ? make_float3(gs[0], gs[1], gs[2])
                                  : make_float3(a.dL_dcolor[3 * idx], a.dL_dcolor[3 * idx + 1],
                                                a.dL_dcolor[3 * idx + 2]);
      const float3 dm = sh_backward(a.D, nc, c, p, make_float3(a.campos[0], a.campos[1], a.campos[2]),
                                    a.clamped[idx], dcol, f);
      if (!staged) store_dsh(a.dL_dsh + (size_t)3 * M * idx, M, f);
      if (a.D > 0) {
        dmean.x += dm.x;
        dmean.y += dm.y;
        dmean.z += dm.z;
      }
    }
    a.dL_dmean3D[3 * idx] = dmean.x;
    a.dL_dmean3D[3 * idx + 1] = dmean.y;
    a.dL_dmean3D[3 * idx + 2] = dmean.z;

    // ---- scale / rotation term ----
    if (a.scales && a.dL_dscale && a.dL_drot) {
      float dsc[3];
      float4 dq;
      scale_rot_backward(dcov, s, q, a.scale_modifier, dsc, dq);
      a.dL_dscale[3 * idx] = dsc[0];
      a.dL_dscale[3 * idx + 1] = dsc[1];
      a.dL_dscale[3 * idx + 2] = dsc[2];
      reinterpret_cast<float4*>(a.dL_drot)[idx] = dq;
    }
  }
  // zeros for Gaussians that are not live
  if (staged) sh_store_wave(a.dL_dsh, a.P, wave_base, lane, s_sh[threadIdx.x / kWave], f);
}

// Batched backward: one thread per Gaussian sums every view's terms (items
// v P + g) in registers and writes each parameter gradient once, including
// dL/dopacity (the render backward leaves it per item).  With the colour
// decoder (shs NULL) the colour term is the decoder backward's.
__global__ __launch_bounds__(256) void preprocess_backward_views_kernel(PreprocessBackwardArgs a,
                                                                        const float* __restrict__ viewmats,
                                                                        const float* __restrict__ camposv,
                                                                        ViewTable vt,
                                                                        const float* __restrict__ dopacity_items,
                                                                        float* __restrict__ dL_dopacity) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.P) return;
  const int M = a.M;
  const float3 p = make_float3(a.means3D[3 * idx], a.means3D[3 * idx + 1], a.means3D[3 * idx + 2]);
  Sym3 S;
  float3 s = make_float3(0.f, 0.f, 0.f);
  float4 q = make_float4(1.f, 0.f, 0.f, 0.f);
  if (a.cov3D_precomp) {
#pragma unroll
    for (int k = 0; k < 6; ++k) S.m[k] = a.cov3D_precomp[6 * idx + k];
  } else {
    s = make_float3(a.scales[3 * idx], a.scales[3 * idx + 1], a.scales[3 * idx + 2]);
    q = reinterpret_cast<const float4*>(a.rotations)[idx];
    S = cov3d_from(s, a.scale_modifier, q);
  }
  const int nc = a.shs ? min(M, sh_coeffs(a.D)) : 0;
  float3 c[16];
  if (a.shs) load_sh(a.shs, M, idx, nc, c);
  float dcov[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float3 dmean = make_float3(0.f, 0.f, 0.f);
  float fsum[48];
#pragma unroll
  for (int k = 0; k < 48; ++k) fsum[k] = 0.f;
  float dop = 0.f;
  bool act = false;  // a colour gradient in any view: the decoder trains this Gaussian
  for (int vi = 0; vi < vt.V; ++vi) {
    const size_t item = (size_t)vi * a.P + idx;
    dop += dopacity_items[item];
    // read with the radius, before any branch, so the loads share one wait
    // (the item's dL/dcolor is written, zero when it was not rendered)
    const float3 dci = make_float3(a.dL_dcolor[3 * item], a.dL_dcolor[3 * item + 1], a.dL_dcolor[3 * item + 2]);
    act |= (dci.x != 0.f) | (dci.y != 0.f) | (dci.z != 0.f);
    if (!(a.radii[item] > 0)) continue;
    const View v = load_view(viewmats + 16 * vi);
    float dc[6];
    float3 dm;
    Jac j;
    cov_mean_backward(v, p, S, vt.W[vi], vt.H[vi], reinterpret_cast<const float4*>(a.dL_dconic)[item],
                      a.dL_dmean2D[3 * item], a.dL_dmean2D[3 * item + 1], dc, dm, j);
#pragma unroll
    for (int k = 0; k < 6; ++k) dcov[k] += dc[k];
    dmean.x += dm.x;
    dmean.y += dm.y;
    dmean.z += dm.z;
    if (a.shs) {
      float f[48];
      const float3 dsv = sh_backward(
          a.D, nc, c, p, make_float3(camposv[3 * vi], camposv[3 * vi + 1], camposv[3 * vi + 2]), a.clamped[item],
          dci, f);
#pragma unroll
      for (int k = 0; k < 48; ++k) fsum[k] += f[k];
      if (a.D > 0) {
        dmean.x += dsv.x;
        dmean.y += dsv.y;
        dmean.z += dsv.z;
      }
    }
  }
  if (a.act_list) {
    // The decoder's work list: the wave's Gaussians with a colour gradient in
    // any view, in index order (one ballot); the others' latent gradients are
    // zero.  Lanes past P left above, so lane 0 holds a valid Gaussian.
    const int lane = threadIdx.x & 63;
    const uint64_t bal = __ballot(act);
    const size_t wave_base = (size_t)idx - (size_t)lane;
    if (act) a.act_list[wave_base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = (uint32_t)idx;
    if (lane == 0) a.act_count[wave_base / kWave] = (uint32_t)__popcll(bal);
    if (!act) {
      float4* dl = reinterpret_cast<float4*>(a.dL_dlatent + (size_t)idx * 16);
      dl[0] = dl[1] = dl[2] = dl[3] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  a.dL_dmean3D[3 * idx] = dmean.x;
  a.dL_dmean3D[3 * idx + 1] = dmean.y;
  a.dL_dmean3D[3 * idx + 2] = dmean.z;
#pragma unroll
  for (int k = 0; k < 6; ++k) a.dL_dcov3D[6 * idx + k] = dcov[k];
  if (a.shs && a.dL_dsh) store_dsh(a.dL_dsh + (size_t)3 * M * idx, M, fsum);
  dL_dopacity[idx] = dop;
  if (a.scales && a.dL_dscale && a.dL_drot) {
    float dsc[3];
    float4 dq;
    scale_rot_backward(dcov, s, q, a.scale_modifier, dsc, dq);
    a.dL_dscale[3 * idx] = dsc[0];
    a.dL_dscale[3 * idx + 1] = dsc[1];
    a.dL_dscale[3 * idx + 2] = dsc[2];
    reinterpret_cast<float4*>(a.dL_drot)[idx] = dq;
  }
}

hipError_t launch_preprocess_backward_views(const PreprocessBackwardArgs& a, const float* viewmats,
                                            const float* campos, const ViewTable& vt, const float* dopacity_items,
                                            float* dL_dopacity, hipStream_t s) {
  if (a.P == 0) return hipSuccess;
  hipLaunchKernelGGL(preprocess_backward_views_kernel, dim3((a.P + 255) / 256), dim3(256), 0, s, a, viewmats, campos,
                     vt, dopacity_items, dL_dopacity);
  return hipGetLastError();
}

hipError_t launch_preprocess_backward(const PreprocessBackwardArgs& a, hipStream_t s) {
  if (a.P == 0) return hipSuccess;
  if (a.gather.tiles_touched)
    hipLaunchKernelGGL(preprocess_backward_kernel<true>, dim3((a.P + 255) / 256), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(preprocess_backward_kernel<false>, dim3((a.P + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace omnigs
