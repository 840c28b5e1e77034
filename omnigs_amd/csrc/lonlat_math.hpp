// Per-Gaussian device math shared by the forward and backward preprocess:
// view transform, lon/lat projection Jacobian, 3D covariance and SH colour.
//
// The equirectangular camera maps the camera-frame point t to
//   lon = atan2(t.x, t.z),  lat = asin(t.y / |t|),
//   px  = ((lon/pi + 1) W - 1) / 2,  py = ((2 lat/pi + 1) H - 1) / 2
// (reference: auxiliary.h:236-248 and ndc2Pix auxiliary.h:51-54), whose
// Jacobian rows are
//   dpx/dt = W/(2pi) * ( t.z, 0, -t.x) / (t.x^2 + t.z^2)
//   dpy/dt = H/pi * (-t.x t.y, rxz^2, -t.z t.y) / (rxz |t|^2)
// (supp.pdf Eq. 17-22; forward.cu:135-155 keeps 1e-7 in each denominator).
#pragma once

#include "common.hpp"

namespace omnigs {

struct View {
  // rotation rows r[row][col] and translation of Tcw; viewmatrix is Tcw
  // column-major, so r[row][col] = m[row + 4 col], t[row] = m[12 + row].
  float r[3][3];
  float t[3];
};

__device__ __forceinline__ View load_view(const float* __restrict__ m) {
  View v;
#pragma unroll
  for (int row = 0; row < 3; ++row) {
#pragma unroll
    for (int col = 0; col < 3; ++col) v.r[row][col] = m[row + 4 * col];
    v.t[row] = m[12 + row];
  }
  return v;
}

__device__ __forceinline__ float3 to_camera(const View& v, float3 p) {
  return make_float3(v.r[0][0] * p.x + v.r[0][1] * p.y + v.r[0][2] * p.z + v.t[0],
                     v.r[1][0] * p.x + v.r[1][1] * p.y + v.r[1][2] * p.z + v.t[1],
                     v.r[2][0] * p.x + v.r[2][1] * p.y + v.r[2][2] * p.z + v.t[2]);
}

// R^T d : a camera-frame gradient back to world frame.
__device__ __forceinline__ float3 to_world_dir(const View& v, float3 d) {
  return make_float3(v.r[0][0] * d.x + v.r[1][0] * d.y + v.r[2][0] * d.z,
                     v.r[0][1] * d.x + v.r[1][1] * d.y + v.r[2][1] * d.z,
                     v.r[0][2] * d.x + v.r[1][2] * d.y + v.r[2][2] * d.z);
}

struct Jac {
  float px_x, px_z;        // dpx/dt.x, dpx/dt.z (dpx/dt.y = 0)
  float py_x, py_y, py_z;  // dpy/dt
  float w2pi, hpi;         // W/(2 pi), H/pi
};

__device__ __forceinline__ Jac lonlat_jacobian(float3 t, int W, int H) {
  const float eps = 1e-7f;
  const float rxz2 = t.x * t.x + t.z * t.z;
  const float inv_rxz2 = 1.0f / (rxz2 + eps);
  const float rxz = sqrtf(rxz2);
  const float inv_rxz = 1.0f / (rxz + eps);
  const float r2 = rxz2 + t.y * t.y;
  const float inv_r2 = 1.0f / (r2 + eps);
  Jac j;
  j.w2pi = (float)W * 0.5f * kInvPi;
  j.hpi = (float)H * kInvPi;
  j.px_x = j.w2pi * t.z * inv_rxz2;
  j.px_z = -j.w2pi * t.x * inv_rxz2;
  j.py_x = -j.hpi * t.x * t.y * inv_rxz * inv_r2;
  j.py_y = j.hpi * rxz * inv_r2;
  j.py_z = -j.hpi * t.z * t.y * inv_rxz * inv_r2;
  return j;
}

// Rows of A = J R (2x3): screen-space directions of the world axes.
struct Proj2 {
  float ax[3], ay[3];
};
__device__ __forceinline__ Proj2 project_rows(const Jac& j, const View& v) {
  Proj2 p;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    p.ax[c] = j.px_x * v.r[0][c] + j.px_z * v.r[2][c];
    p.ay[c] = j.py_x * v.r[0][c] + j.py_y * v.r[1][c] + j.py_z * v.r[2][c];
  }
  return p;
}

// Symmetric 3x3 stored as its upper triangle [xx, xy, xz, yy, yz, zz].
struct Sym3 {
  float m[6];
  __device__ __forceinline__ float at(int r, int c) const {
    const int lo = r < c ? r : c, hi = r < c ? c : r;
    return m[lo == 0 ? hi : (lo == 1 ? 2 + hi : 5)];
  }
};

// Rotation matrix of the (unnormalised, as given) quaternion q = (w, x, y, z).
__device__ __forceinline__ void quat_to_rot(float4 q, float R[3][3]) {
  const float w = q.x, x = q.y, y = q.z, z = q.w;
  R[0][0] = 1.f - 2.f * (y * y + z * z);
  R[0][1] = 2.f * (x * y - w * z);
  R[0][2] = 2.f * (x * z + w * y);
  R[1][0] = 2.f * (x * y + w * z);
  R[1][1] = 1.f - 2.f * (x * x + z * z);
  R[1][2] = 2.f * (y * z - w * x);
  R[2][0] = 2.f * (x * z - w * y);
  R[2][1] = 2.f * (y * z + w * x);
  R[2][2] = 1.f - 2.f * (x * x + y * y);
}

// Sigma = R S^2 R^T with S = diag(mod * s)  (reference computeCov3D,
// forward.cu:194-228, whose glm M = S R^T-layout gives the same Sigma).
__device__ __forceinline__ Sym3 cov3d_from(float3 s, float mod, float4 q) {
  float R[3][3];
  quat_to_rot(q, R);
  const float sx = mod * s.x, sy = mod * s.y, sz = mod * s.z;
  // M[k][i] = s_k * R[i][k]: row k of M is s_k times column k of R.
  float Mk[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Mk[0][i] = sx * R[i][0];
    Mk[1][i] = sy * R[i][1];
    Mk[2][i] = sz * R[i][2];
  }
  Sym3 S;
  int o = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) S.m[o++] = Mk[0][i] * Mk[0][j] + Mk[1][i] * Mk[1][j] + Mk[2][i] * Mk[2][j];
  return S;
}

// v^T Sigma w
__device__ __forceinline__ float quad_form(const Sym3& S, const float* v, const float* w) {
  float acc = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float sw = S.at(i, 0) * w[0] + S.at(i, 1) * w[1] + S.at(i, 2) * w[2];
    acc += v[i] * sw;
  }
  return acc;
}

// Degree <= 3 real SH basis of a unit direction, 16 values (unused tail = 0).
__device__ __forceinline__ void sh_basis(int deg, float x, float y, float z, float b[16]) {
  b[0] = kSH0;
#pragma unroll
  for (int i = 1; i < 16; ++i) b[i] = 0.f;
  if (deg > 0) {
    b[1] = -kSH1 * y;
    b[2] = kSH1 * z;
    b[3] = -kSH1 * x;
    if (deg > 1) {
      const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
      b[4] = kSH2_0 * xy;
      b[5] = kSH2_1 * yz;
      b[6] = kSH2_2 * (2.f * zz - xx - yy);
      b[7] = kSH2_3 * xz;
      b[8] = kSH2_4 * (xx - yy);
      if (deg > 2) {
        b[9] = kSH3_0 * y * (3.f * xx - yy);
        b[10] = kSH3_1 * xy * z;
        b[11] = kSH3_2 * y * (4.f * zz - xx - yy);
        b[12] = kSH3_3 * z * (2.f * zz - 3.f * xx - 3.f * yy);
        b[13] = kSH3_4 * x * (4.f * zz - xx - yy);
        b[14] = kSH3_5 * z * (xx - yy);
        b[15] = kSH3_6 * x * (xx - 3.f * yy);
      }
    }
  }
}

// Loads the first min(M,16) SH coefficients (float3 each) of Gaussian `g`.
__device__ __forceinline__ void load_sh(const float* __restrict__ shs, int M, uint32_t g, int ncoef, float3 c[16]) {
  const float* base = shs + (size_t)g * M * 3;
  if (M == 16) {
    const float4* b4 = reinterpret_cast<const float4*>(base);
    float f[48];
#pragma unroll
    for (int i = 0; i < 12; ++i) {
      const float4 v = b4[i];
      f[4 * i] = v.x; f[4 * i + 1] = v.y; f[4 * i + 2] = v.z; f[4 * i + 3] = v.w;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) c[i] = make_float3(f[3 * i], f[3 * i + 1], f[3 * i + 2]);
  } else {
#pragma unroll
    for (int i = 0; i < 16; ++i)
      c[i] = i < ncoef ? make_float3(base[3 * i], base[3 * i + 1], base[3 * i + 2]) : make_float3(0.f, 0.f, 0.f);
  }
}

__device__ __forceinline__ int sh_coeffs(int deg) { return (deg + 1) * (deg + 1); }

}  // namespace omnigs
