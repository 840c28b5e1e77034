// Forward preprocess: one thread per Gaussian.
//
// Follows the semantics of the reference's preprocessLonlatCUDA
// (cuda_rasterizer/forward.cu:593-703): distance cull at |t|^2 <= 0.04,
// lon/lat projection, 3D covariance, EWA 2D covariance through the lon/lat
// Jacobian (+0.3 I, no clamping), conic, 3-sigma radius, clamped tile rect
// (no longitude-seam wrap), SH colour with clamp-at-zero flags.  The sort key
// written here is the float bit pattern of the camera distance |t|.
//
// HBM-bound: ~236 B read (means, scales, rotations, opacity, 48 SH floats)
// and ~45 B written per Gaussian.  With the colour decoder the 192 B of SH
// become a 32 B BF16 latent and the colour comes from a wave-wide MFMA
// evaluation (mlp.hpp).
#include "kernels.hpp"
#include "lonlat_math.hpp"
#include "mlp.hpp"

namespace omnigs {

// Projected geometry of one Gaussian.
struct GeomOut {
  float3 p;  // world mean
  float r;   // camera distance |t|
  float2 pix;
  float4 conic;
  int radius, area;
  Rect rc;  // tile rect
};

// Per-Gaussian inputs of the projection (read once per Gaussian).
struct GaussGeo {
  float3 p;
  Sym3 S;
  float opacity;
};

__device__ __forceinline__ GaussGeo load_gauss(const PreprocessArgs& a, int idx) {
  GaussGeo gg;
  gg.p = make_float3(a.means3D[3 * idx], a.means3D[3 * idx + 1], a.means3D[3 * idx + 2]);
  if (a.cov3D_precomp) {
#pragma unroll
    for (int k = 0; k < 6; ++k) gg.S.m[k] = a.cov3D_precomp[6 * idx + k];
  } else {
    const float3 s = make_float3(a.scales[3 * idx], a.scales[3 * idx + 1], a.scales[3 * idx + 2]);
    const float4 q = reinterpret_cast<const float4*>(a.rotations)[idx];
    gg.S = cov3d_from(s, a.scale_modifier, q);
  }
  gg.opacity = a.opacities[idx];
  return gg;
}

// Geometry part of preprocessLonlatCUDA for one view; false when the
// Gaussian is culled (too close, singular 2D covariance or empty tile rect).
__device__ __forceinline__ bool project_view(const GaussGeo& gg, const float* viewmatrix, int W, int H, TileGrid grid,
                                             GeomOut& o) {
  const View v = load_view(viewmatrix);
  const float3 p = gg.p;
  o.p = p;
  const float3 t = to_camera(v, p);
  const float rr = t.x * t.x + t.y * t.y + t.z * t.z;
  if (rr <= 0.04f) return false;
  const float r = sqrtf(rr);
  const float inv_r = 1.0f / (r + 1e-7f);
  const float lon = atan2f(t.x, t.z);
  const float lat = asinf(t.y * inv_r);
  const float sx = lon * kInvPi;
  const float sy = lat * kTwoInvPi;

  const Sym3& S = gg.S;
  const Jac j = lonlat_jacobian(t, W, H);
  const Proj2 A = project_rows(j, v);
  const float cxx = quad_form(S, A.ax, A.ax) + 0.3f;
  const float cxy = quad_form(S, A.ax, A.ay);
  const float cyy = quad_form(S, A.ay, A.ay) + 0.3f;

  const float det = cxx * cyy - cxy * cxy;
  if (det == 0.0f) return false;
  const float det_inv = 1.f / det;
  o.conic = make_float4(cyy * det_inv, -cxy * det_inv, cxx * det_inv, gg.opacity);

  const float mid = 0.5f * (cxx + cyy);
  const float disc = sqrtf(fmaxf(0.1f, mid * mid - det));
  const float lambda = fmaxf(mid + disc, mid - disc);
  o.radius = (int)ceilf(3.f * sqrtf(lambda));

  // ndc -> pixel in FP64, rounded to float (reference ndc2Pix evaluates in double).
  const float px = (float)((((double)sx + 1.0) * (double)W - 1.0) * 0.5);
  const float py = (float)((((double)sy + 1.0) * (double)H - 1.0) * 0.5);
  o.pix = make_float2(px, py);
  const Rect rc = tile_rect(o.pix, o.radius, grid);
  o.rc = rc;
  o.area = (rc.x1 - rc.x0) * (rc.y1 - rc.y0);
  o.r = r;
  return o.area != 0;
}

__device__ __forceinline__ bool preprocess_geometry(const PreprocessArgs& a, int idx, TileGrid grid, GeomOut& o) {
  return project_view(load_gauss(a, idx), a.viewmatrix, a.W, a.H, grid, o);
}

// SH colour (computeColorFromSH, forward.cu:30-83) from coefficients already
// in registers; returns the clamp-at-zero flags.
__device__ __forceinline__ float3 sh_colour(int D, int nc, const float3 c[16], float3 p, float3 campos,
                                            uint8_t* clamped) {
  float3 d = make_float3(p.x - campos.x, p.y - campos.y, p.z - campos.z);
  const float len = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
  d = make_float3(d.x / len, d.y / len, d.z / len);
  float b[16];
  sh_basis(D, d.x, d.y, d.z, b);
  float3 rgb = make_float3(0.f, 0.f, 0.f);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    if (i < nc) {
      rgb.x += b[i] * c[i].x;
      rgb.y += b[i] * c[i].y;
      rgb.z += b[i] * c[i].z;
    }
  }
  rgb.x += 0.5f;
  rgb.y += 0.5f;
  rgb.z += 0.5f;
  *clamped = (uint8_t)((rgb.x < 0.f) | ((rgb.y < 0.f) << 1) | ((rgb.z < 0.f) << 2));
  return make_float3(fmaxf(rgb.x, 0.f), fmaxf(rgb.y, 0.f), fmaxf(rgb.z, 0.f));
}

// kMlp: colours from the decoder (every lane of a wave that holds at least one
// Gaussian runs the wave-wide MFMA evaluation, culled or not).
template <bool kMlp>
__global__ __launch_bounds__(256) void preprocess_kernel(PreprocessArgs a, GeomState g, int* __restrict__ radii_out,
                                                         TileGrid grid) {
  __shared__ __attribute__((aligned(16))) char s_w[kMlp ? kMlpFwdBytes : 16];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in_range = idx < a.P;
  // MLP: the decoder weights (for LDS) and the wave's latents are requested
  // before the geometry, so their latencies overlap it; the weights reach
  // LDS before the block scan's barriers, which publish them.
  // (four named registers: an indexed array here would live in scratch)
  constexpr int kWn = kMlpFwdBytes / 16;
  static_assert(kWn > 768 && kWn <= 1024, "four 16-byte weight loads per thread of 256");
  uint4 w0{}, w1{}, w2{}, w3{};
  bf16x8 lat[kMlp ? 4 : 1];
  if constexpr (kMlp) {
    const uint4* src = reinterpret_cast<const uint4*>(a.mlp_packed);
    const int t = threadIdx.x;
    w0 = src[t];
    w1 = src[t + 256];
    w2 = src[t + 512];
    w3 = src[min(t + 768, kWn - 1)];
    mlp_wave_latents(a.latents, a.P, (size_t)(idx & ~63), lat);
  }

  if (in_range) {
    radii_out[idx] = 0;
    g.tiles_touched[idx] = 0;
    g.depth_key[idx] = 0xFFFFFFFFu;  // culled Gaussians sort last and emit nothing
  }
  GeomOut o;
  const bool live = in_range && preprocess_geometry(a, idx, grid, o);
  if constexpr (kMlp) {
    uint4* dst = reinterpret_cast<uint4*>(s_w);
    const int t = threadIdx.x;
    dst[t] = w0;
    dst[t + 256] = w1;
    dst[t + 512] = w2;
    if (t + 768 < kWn) dst[t + 768] = w3;
  }
  // gradient slots: the block's Gaussians in index order, one slot per tile
  // of the rect (every thread of the block takes part in the scan)
  uint32_t block_total;
  const uint32_t slot = block_exclusive_scan(live ? (uint32_t)o.area : 0u, &block_total);
  if (threadIdx.x == 0) g.block_slots[blockIdx.x] = block_total;
  if (in_range) {
    g.slot_local[idx] = slot;
    g.dup_rect[idx] = pack_dup_rect(live, o.area, o.rc);
  }
  if ((kMlp ? (idx & ~63) : idx) >= a.P) return;  // MLP: whole waves evaluate the decoder

  float3 col;
  if constexpr (kMlp) {
    float3 dir = make_float3(0.f, 0.f, 1.f);
    if (in_range) {
      const float3 d = make_float3(o.p.x - a.campos[0], o.p.y - a.campos[1], o.p.z - a.campos[2]);
      const float len = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
      if (len > 0.f) dir = make_float3(d.x / len, d.y / len, d.z / len);
    }
    col = mlp_colour_wave(s_w, lat, dir);
    if (!live) return;
    g.clamped[idx] = 0;
  } else {
    if (!live) return;
    if (a.colors_precomp) {
      col = make_float3(a.colors_precomp[3 * idx], a.colors_precomp[3 * idx + 1], a.colors_precomp[3 * idx + 2]);
    } else {
      const int nc = min(a.M, sh_coeffs(a.D));
      float3 c[16];
      load_sh(a.shs, a.M, idx, nc, c);
      col = sh_colour(a.D, nc, c, o.p, make_float3(a.campos[0], a.campos[1], a.campos[2]), &g.clamped[idx]);
    }
  }

  g.depth_key[idx] = __float_as_uint(o.r);
  radii_out[idx] = o.radius;
  store_rec(g.rec, idx, o.pix, o.conic, col, o.radius, slot);
  g.tiles_touched[idx] = (uint32_t)o.area;
}

// Batched preprocess: one thread per Gaussian reads its parameters (and its
// SH coefficients or decoder latents) once and writes item v P + g of every
// view v.  viewmats [V,16] and campos [V,3] are device arrays.
template <bool kMlp>
__global__ __launch_bounds__(256) void preprocess_views_kernel(PreprocessArgs a, const float* __restrict__ viewmats,
                                                               const float* __restrict__ camposv, ViewTable vt,
                                                               GeomState g, int* __restrict__ radii_out) {
  __shared__ __attribute__((aligned(16))) char s_w[kMlp ? kMlpFwdBytes : 16];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in_range = idx < a.P;
  const bool wave_has = (idx & ~63) < a.P;  // MLP: whole waves evaluate the decoder
  // MLP: weights and the wave's latents (the same for every view) requested
  // first, as in preprocess_kernel; the first view's block scan publishes
  // the weights in LDS before any evaluation
  constexpr int kWn = kMlpFwdBytes / 16;
  uint4 w0{}, w1{}, w2{}, w3{};
  bf16x8 lat[kMlp ? 4 : 1];
  if constexpr (kMlp) {
    const uint4* src = reinterpret_cast<const uint4*>(a.mlp_packed);
    const int t = threadIdx.x;
    w0 = src[t];
    w1 = src[t + 256];
    w2 = src[t + 512];
    w3 = src[min(t + 768, kWn - 1)];
    mlp_wave_latents(a.latents, a.P, (size_t)(idx & ~63), lat);
  }
  GaussGeo gg{};
  float3 c[16];
  const int nc = (kMlp || a.colors_precomp) ? 0 : min(a.M, sh_coeffs(a.D));
  if (in_range) {
    gg = load_gauss(a, idx);
    if (!kMlp && !a.colors_precomp) load_sh(a.shs, a.M, idx, nc, c);
  }
  if constexpr (kMlp) {
    uint4* dst = reinterpret_cast<uint4*>(s_w);
    const int t = threadIdx.x;
    dst[t] = w0;
    dst[t + 256] = w1;
    dst[t + 512] = w2;
    if (t + 768 < kWn) dst[t + 768] = w3;
  }
  // gradient slots of the block: view by view, the block's Gaussians in index
  // order within each view (every thread of the block takes part in the scans)
  uint32_t view_prefix = 0;
  for (int v = 0; v < vt.V; ++v) {
    const size_t item = (size_t)v * a.P + idx;
    if (in_range) {
      radii_out[item] = 0;
      g.tiles_touched[item] = 0;
      g.depth_key[item] = 0xFFFFFFFFu;
    }
    GeomOut o;
    const bool live = in_range && project_view(gg, viewmats + 16 * v, vt.W[v], vt.H[v], vt.grid(v), o);
    uint32_t view_total;
    const uint32_t slot = view_prefix + block_exclusive_scan(live ? (uint32_t)o.area : 0u, &view_total);
    view_prefix += view_total;
    if (in_range) {
      g.slot_local[item] = slot;
      g.dup_rect[item] = pack_dup_rect(live, o.area, o.rc);
    }
    const float3 cp = make_float3(camposv[3 * v], camposv[3 * v + 1], camposv[3 * v + 2]);
    float3 col;
    if constexpr (kMlp) {
      if (!wave_has) continue;
      float3 dir = make_float3(0.f, 0.f, 1.f);
      if (in_range) {
        const float3 d = make_float3(gg.p.x - cp.x, gg.p.y - cp.y, gg.p.z - cp.z);
        const float len = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
        if (len > 0.f) dir = make_float3(d.x / len, d.y / len, d.z / len);
      }
      col = mlp_colour_wave(s_w, lat, dir);
      if (!live) continue;
      g.clamped[item] = 0;
    } else {
      if (!live) continue;
      if (a.colors_precomp) {
        col = make_float3(a.colors_precomp[3 * idx], a.colors_precomp[3 * idx + 1], a.colors_precomp[3 * idx + 2]);
      } else {
        col = sh_colour(a.D, nc, c, gg.p, cp, &g.clamped[item]);
      }
    }
    g.depth_key[item] = __float_as_uint(o.r);
    radii_out[item] = o.radius;
    store_rec(g.rec, item, o.pix, o.conic, col, o.radius, slot);
    g.tiles_touched[item] = (uint32_t)o.area;
  }
  if (threadIdx.x == 0) g.block_slots[blockIdx.x] = view_prefix;
}

hipError_t launch_preprocess_views(const PreprocessArgs& a, const float* viewmats, const float* campos,
                                   const ViewTable& vt, const GeomState& g, int* radii, hipStream_t s) {
  if (a.P == 0 || vt.V == 0) return hipSuccess;
  if (a.mlp_packed && a.latents)
    hipLaunchKernelGGL(preprocess_views_kernel<true>, dim3((a.P + 255) / 256), dim3(256), 0, s, a, viewmats, campos,
                       vt, g, radii);
  else
    hipLaunchKernelGGL(preprocess_views_kernel<false>, dim3((a.P + 255) / 256), dim3(256), 0, s, a, viewmats,
                       campos, vt, g, radii);
  return hipGetLastError();
}

__global__ void mark_visible_kernel(int P, uint8_t* __restrict__ present) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < P) present[idx] = 1;
}

hipError_t launch_preprocess(const PreprocessArgs& a, const GeomState& g, int* radii, hipStream_t s) {
  if (a.P == 0) return hipSuccess;
  const TileGrid grid = tile_grid(a.W, a.H, a.seam_wrap, a.tile_x, a.tile_y);
  if (a.mlp_packed && a.latents)
    hipLaunchKernelGGL(preprocess_kernel<true>, dim3((a.P + 255) / 256), dim3(256), 0, s, a, g, radii, grid);
  else
    hipLaunchKernelGGL(preprocess_kernel<false>, dim3((a.P + 255) / 256), dim3(256), 0, s, a, g, radii, grid);
  return hipGetLastError();
}

hipError_t launch_mark_visible(int P, uint8_t* present, hipStream_t s) {
  if (P == 0) return hipSuccess;
  hipLaunchKernelGGL(mark_visible_kernel, dim3((P + 255) / 256), dim3(256), 0, s, P, present);
  return hipGetLastError();
}

}  // namespace omnigs
