// Tile render, forward and backward.  One 256-thread workgroup per tile of
// 256 pixels (16x16 as in the reference; 32x8 and 8x32 as extensions); each
// 64-lane wavefront owns an 8x8 quadrant of the tile so that a wave's pixels
// are spatially compact (more waves skip a small Gaussian).  Batched launches
// (kMulti) cover every tile of every view: a block finds its view in the
// ViewTable and renders into that view's slice of the pixel arrays.
//
// Per-pixel semantics follow the reference's renderCUDA forward
// (forward.cu:346-467) and backward (backward.cu:672-843):
//   power = -0.5 (A dx^2 + C dy^2) - B dx dy   (skip if > 0)
//   alpha = min(0.99, o exp(power))            (skip if < 1/255)
//   stop before blending once T (1 - alpha) < 1e-4
//   out = sum c alpha T + T_final bg;  n_contrib = 1-based index of the last blended instance.
// The backward walks each tile's list from the back, recovers T by division,
// and reduces every Gaussian's nine gradient terms across the wave with DPP
// adds, then across the four waves in LDS, before one global atomic per
// (tile, Gaussian, term).
#include "kernels.hpp"

namespace omnigs {

constexpr int kBatch = kTilePixels;  // instances staged per LDS round
constexpr int kWavesPerTile = kTilePixels / kWave;

// Quadrant layout of a TX x TY tile: QX x QY quadrants of 8x8 pixels.
template <int TX, int TY>
struct Quads {
  static constexpr int QX = TX / 8, QY = TY / 8;
  static_assert(TX * TY == kTilePixels && QX * QY == kWavesPerTile, "256-pixel tiles of 8x8 quadrants");
};

template <int TX, int TY>
__device__ __forceinline__ void pixel_of(int tile, int gx, int tid, int& px, int& py) {
  const int lane = tid & (kWave - 1), w = tid / kWave;
  const int tx = tile % gx, ty = tile / gx;
  px = tx * TX + (w % Quads<TX, TY>::QX) * 8 + (lane & 7);
  py = ty * TY + (w / Quads<TX, TY>::QX) * 8 + (lane >> 3);
}

// Where this block renders: view size, tile grid width, tile within the view
// and the view's first pixel in the (batched) pixel arrays.
struct TileLoc {
  int W, H, gx, tile;
  size_t pix;
};
template <bool kMulti>
__device__ __forceinline__ TileLoc locate(int W, int H, int gx, const ViewTable& vt) {
  if constexpr (kMulti) {
    const int v = vt.view_of_tile(blockIdx.x);
    return {vt.W[v], vt.H[v], vt.gx[v], (int)(blockIdx.x - vt.tile_base[v]), (size_t)vt.pix_base[v]};
  } else {
    return {W, H, gx, (int)blockIdx.x, 0};
  }
}

// exp(power) is evaluated as exp2 of a power prescaled by log2(e) at staging:
// power2 = a dx^2 + c dy^2 + b dx dy with (a, b, c) = -log2(e) (A/2, B, C/2),
// G = 2^power2 on the hardware exp2.  The forward and the backward use the
// same staged values, so every alpha decision agrees between them.
constexpr float kLog2e = 1.4426950408889634f;
__device__ __forceinline__ float3 prescale_conic(float4 co) {
  return make_float3(-0.5f * kLog2e * co.x, -kLog2e * co.y, -0.5f * kLog2e * co.z);
}
__device__ __forceinline__ float power2_of(float3 q, float dx, float dy) {
  return dx * (q.x * dx + q.y * dy) + q.z * dy * dy;
}
__device__ __forceinline__ float exp2_hw(float x) { return __builtin_amdgcn_exp2f(x); }

// Horizontal offset to the nearest copy of the Gaussian across the longitude
// seam (seam-wrap mode only).
template <bool kWrap>
__device__ __forceinline__ float seam_dx(float dx, float W, float invW) {
  if (kWrap) dx -= W * rintf(dx * invW);
  return dx;
}

// Conservative 4-bit mask of the tile's 8x8 wave quadrants that Gaussian
// (m, co) can reach with alpha >= 1/255: outside the ellipse
// 0.5 d^T Q d <= ln(255 o) the alpha test fails for every pixel, so skipping
// those (wave, instance) pairs changes no result.  The box is padded by
// 1% + 0.05 px against rounding; degenerate conics get the full mask.
template <int TX, int TY>
__device__ __forceinline__ uint32_t quadrant_mask(float2 m, float4 co, int tile, int gx, float wrapW) {
  constexpr int QX = Quads<TX, TY>::QX;
  const float det = co.x * co.z - co.y * co.y;
  const float o255 = 255.0f * co.w;
  if (!(o255 > 0.9999f)) return 0u;  // alpha < 1/255 everywhere
  if (!(det > 0.f) || !(co.x > 0.f) || !(co.z > 0.f)) return 0xFu;
  const float tau2 = fmaxf(2.0f * __logf(o255) * 1.0001f, 0.f) + 1e-3f;
  const float ex = sqrtf(tau2 * co.z / det) * 1.01f + 0.05f;
  const float ey = sqrtf(tau2 * co.x / det) * 1.01f + 0.05f;
  const float x0 = (float)((tile % gx) * TX), y0 = (float)((tile / gx) * TY);
  if (wrapW > 0.f) m.x -= wrapW * rintf((m.x - (x0 + 0.5f * TX)) / wrapW);  // nearest copy across the seam
  uint32_t mk = 0;
#pragma unroll
  for (int q = 0; q < kWavesPerTile; ++q) {
    const float qx = x0 + (float)((q % QX) * 8), qy = y0 + (float)((q / QX) * 8);
    const bool hit = (m.x + ex >= qx) && (m.x - ex <= qx + 7.f) && (m.y + ey >= qy) && (m.y - ey <= qy + 7.f);
    mk |= hit ? (1u << q) : 0u;
  }
  return mk;
}

// Each wave compacts the batch positions whose mask has its bit set into its
// own ascending list; returns the list length (wave-uniform).  Slots below
// `first_slot` are dropped (backward: positions past every lane's last
// contributor).
__device__ __forceinline__ int build_wave_list(const uint8_t* s_mask, uint8_t* list, int cnt, int w, int lane,
                                               int first_slot = 0) {
  int count = 0;
  const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
  for (int c = 0; c < kBatch / kWave; ++c) {
    const int j = c * kWave + lane;
    const bool on = j >= first_slot && j < cnt && ((s_mask[j] >> w) & 1u);
    const uint64_t bal = __ballot(on);
    if (on) list[count + __popcll(bal & below)] = (uint8_t)j;
    count += __popcll(bal);
  }
  return count;
}

template <int TX, int TY, bool kWrap, bool kMulti>
__global__ __launch_bounds__(kTilePixels) void render_forward_kernel(
    int gx_, const uint2* __restrict__ ranges, const uint32_t* __restrict__ point_list, int W_, int H_,
    ViewTable vt, const float4* __restrict__ rec, float* __restrict__ final_T, uint32_t* __restrict__ n_contrib,
    const float* __restrict__ bg, float* __restrict__ out_color) {
  const TileLoc loc = locate<kMulti>(W_, H_, gx_, vt);
  const int W = loc.W, H = loc.H, gx = loc.gx;
  const float Wf = (float)W, invW = 1.0f / (float)W;
  __shared__ float2 s_xy[kBatch];
  __shared__ float4 s_q[kBatch];  // prescaled conic (a, b, c) and opacity
  __shared__ float s_rgb[kChannels][kBatch];
  __shared__ uint8_t s_mask[kBatch];
  __shared__ uint8_t s_list[kWavesPerTile][kBatch];

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), w = tid / kWave;
  const int tile = loc.tile;
  int px, py;
  pixel_of<TX, TY>(tile, gx, tid, px, py);
  const bool inside = px < W && py < H;
  const float fx = (float)px, fy = (float)py;
  const uint2 range = ranges[blockIdx.x];
  const int n = (int)(range.y - range.x);

  bool done = !inside;
  float T = 1.0f;
  float C0 = 0.f, C1 = 0.f, C2 = 0.f;
  uint32_t last = 0;

  for (int base = 0; base < n; base += kBatch) {
    if (__syncthreads_count(done) == kTilePixels) break;
    const int k = base + tid;
    if (k < n) {
      const uint32_t g = point_list[range.x + k];
      const RenderRec rr = load_rec(rec, g);
      const float2 m = rr.xy;
      const float4 co = rr.co;
      const float3 q = prescale_conic(co);
      s_xy[tid] = m;
      s_q[tid] = make_float4(q.x, q.y, q.z, co.w);
      s_rgb[0][tid] = rr.rgb.x;
      s_rgb[1][tid] = rr.rgb.y;
      s_rgb[2][tid] = rr.rgb.z;
      s_mask[tid] = (uint8_t)quadrant_mask<TX, TY>(m, co, tile, gx, kWrap ? Wf : 0.f);
    }
    __syncthreads();
    const int cnt = min(kBatch, n - base);
    const int len = build_wave_list(s_mask, s_list[w], cnt, w, lane);
    for (int ii = 0; ii < len && !done; ++ii) {
      const int j = s_list[w][ii];
      const float2 m = s_xy[j];
      const float4 q = s_q[j];
      const float dx = seam_dx<kWrap>(m.x - fx, Wf, invW), dy = m.y - fy;
      const float power2 = power2_of(make_float3(q.x, q.y, q.z), dx, dy);
      if (power2 > 0.0f) continue;
      const float alpha = fminf(0.99f, q.w * exp2_hw(power2));
      if (alpha < 1.0f / 255.0f) continue;
      const float test_T = T * (1.0f - alpha);
      if (test_T < 0.0001f) {
        done = true;
        continue;
      }
      const float wgt = alpha * T;
      C0 += s_rgb[0][j] * wgt;
      C1 += s_rgb[1][j] * wgt;
      C2 += s_rgb[2][j] * wgt;
      T = test_T;
      last = (uint32_t)(base + j + 1);  // 1-based position in the tile's list
    }
  }

  if (inside) {
    const size_t pid = (size_t)py * W + px;
    const size_t N = (size_t)W * H;
    final_T[loc.pix + pid] = T;
    n_contrib[loc.pix + pid] = last;
    float* out = out_color + 3 * loc.pix;
    out[pid] = C0 + T * bg[0];
    out[N + pid] = C1 + T * bg[1];
    out[2 * N + pid] = C2 + T * bg[2];
  }
}

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

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) v = max(v, __shfl_xor(v, o, kWave));
  return v;
}

constexpr int kGradTerms = 9;  // dcolor x3, dmean2D x2, dconic x3, dopacity

template <int TX, int TY, bool kWrap, bool kMulti>
__global__ __launch_bounds__(kTilePixels) void render_backward_kernel(
    int gx_, const uint2* __restrict__ ranges, const uint32_t* __restrict__ point_list, int W_, int H_,
    ViewTable vt, const float* __restrict__ bg, const float4* __restrict__ rec, const float* __restrict__ final_Ts,
    const uint32_t* __restrict__ n_contrib,
    const float* __restrict__ dL_dpix, float* __restrict__ dL_dmean2D, float* __restrict__ dL_dconic,
    float* __restrict__ dL_dopacity, float* __restrict__ dL_dcolors) {
  const TileLoc loc = locate<kMulti>(W_, H_, gx_, vt);
  const int W = loc.W, H = loc.H, gx = loc.gx;
  const float Wf = (float)W, invW = 1.0f / (float)W;
  __shared__ uint32_t s_gid[kBatch];
  __shared__ float2 s_xy[kBatch];
  __shared__ float4 s_co[kBatch];
  __shared__ float4 s_q[kBatch];  // prescaled conic, as in the forward
  __shared__ float s_rgb[kChannels][kBatch];
  __shared__ float s_acc[kGradTerms][kBatch];
  __shared__ uint8_t s_mask[kBatch];
  __shared__ uint8_t s_list[kWavesPerTile][kBatch];
  __shared__ int s_nlive;

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), w = tid / kWave;
  const int tile = loc.tile;
  int px, py;
  pixel_of<TX, TY>(tile, gx, tid, px, py);
  const bool inside = px < W && py < H;
  const float fx = (float)px, fy = (float)py;
  const uint2 range = ranges[blockIdx.x];
  const size_t N = (size_t)W * H;
  const size_t pid = (size_t)py * W + px;
  final_Ts += loc.pix;
  n_contrib += loc.pix;
  dL_dpix += 3 * loc.pix;

  const float T_final = inside ? final_Ts[pid] : 0.f;
  float T = T_final;
  const int last = inside ? (int)n_contrib[pid] : 0;
  // Positions >= max(last) over the tile blend into no pixel: the traversal
  // starts there instead of at the end of the list.  Per wave, positions
  // >= the wave's own max(last) are dropped from its lists.
  if (tid == 0) s_nlive = 0;
  __syncthreads();
  const int wlast = wave_max_int(last);
  if (lane == 0) atomicMax(&s_nlive, wlast);
  __syncthreads();
  const int n = s_nlive;
  float dp0 = 0.f, dp1 = 0.f, dp2 = 0.f;
  if (inside) {
    dp0 = dL_dpix[pid];
    dp1 = dL_dpix[N + pid];
    dp2 = dL_dpix[2 * N + pid];
  }
  const float ddelx_dx = 0.5f * (float)W, ddely_dy = 0.5f * (float)H;  // NDC scale of dL/dmean2D
  // -T_final (bg . dL/dpix): the background term of dL/dalpha is this times 1/(1 - alpha)
  const float neg_tb = -T_final * (bg[0] * dp0 + bg[1] * dp1 + bg[2] * dp2);
  float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;
  float last_alpha = 0.f;
  float lc0 = 0.f, lc1 = 0.f, lc2 = 0.f;

  for (int base = 0; base < n; base += kBatch) {
    __syncthreads();
    const int k = base + tid;
    if (k < n) {
      const uint32_t g = point_list[range.x + (n - 1 - k)];
      const RenderRec rr = load_rec(rec, g);
      const float2 m = rr.xy;
      const float4 co = rr.co;
      const float3 q = prescale_conic(co);
      s_gid[tid] = g;
      s_xy[tid] = m;
      s_co[tid] = co;
      s_q[tid] = make_float4(q.x, q.y, q.z, co.w);
      s_rgb[0][tid] = rr.rgb.x;
      s_rgb[1][tid] = rr.rgb.y;
      s_rgb[2][tid] = rr.rgb.z;
      s_mask[tid] = (uint8_t)quadrant_mask<TX, TY>(m, co, tile, gx, kWrap ? Wf : 0.f);
    }
#pragma unroll
    for (int qq = 0; qq < kGradTerms; ++qq) s_acc[qq][tid] = 0.f;
    __syncthreads();

    const int cnt = min(kBatch, n - base);
    // slot j holds position n - 1 - (base + j); keep positions < wlast
    const int first_slot = n - base - wlast;
    const int len = build_wave_list(s_mask, s_list[w], cnt, w, lane, first_slot);
    for (int ii = 0; ii < len; ++ii) {
      const int j = s_list[w][ii];
      // 0-based position of this instance in the tile's front-to-back list
      const int pos = n - 1 - (base + j);
      const float2 m = s_xy[j];
      const float4 q = s_q[j];
      const float dx = seam_dx<kWrap>(m.x - fx, Wf, invW), dy = m.y - fy;
      const float power2 = power2_of(make_float3(q.x, q.y, q.z), dx, dy);
      const float G = exp2_hw(power2);
      const float alpha = fminf(0.99f, q.w * G);
      const bool ok = inside && pos < last && power2 <= 0.0f && alpha >= 1.0f / 255.0f;
      // per-pixel moments: w = G dL/dalpha, v = alpha T
      float m_wdx = 0.f, m_wdy = 0.f, m_wdx2 = 0.f, m_wdxdy = 0.f, m_wdy2 = 0.f, m_w = 0.f;
      float m_v0 = 0.f, m_v1 = 0.f, m_v2 = 0.f;
      if (ok) {
        // T and 1/(1 - alpha) only scale gradients here (no decision depends
        // on them), so the hardware reciprocal replaces two IEEE divisions.
        const float r = __builtin_amdgcn_rcpf(1.f - alpha);
        T = T * r;
        const float v = alpha * T;
        const float c0 = s_rgb[0][j], c1 = s_rgb[1][j], c2 = s_rgb[2][j];
        const float one_minus_la = 1.f - last_alpha;
        acc0 = last_alpha * lc0 + one_minus_la * acc0;
        acc1 = last_alpha * lc1 + one_minus_la * acc1;
        acc2 = last_alpha * lc2 + one_minus_la * acc2;
        const float dL_dalpha = ((c0 - acc0) * dp0 + (c1 - acc1) * dp1 + (c2 - acc2) * dp2) * T + neg_tb * r;
        last_alpha = alpha;
        lc0 = c0;
        lc1 = c1;
        lc2 = c2;
        const float wgt = G * dL_dalpha;
        m_wdx = wgt * dx;
        m_wdy = wgt * dy;
        m_wdx2 = m_wdx * dx;
        m_wdxdy = m_wdx * dy;
        m_wdy2 = m_wdy * dy;
        m_w = wgt;
        m_v0 = v * dp0;
        m_v1 = v * dp1;
        m_v2 = v * dp2;
      }
      if (__ballot(ok) == 0ull) continue;  // wave-uniform
      const float t[8] = {m_v0, m_v1, m_v2, m_wdx, m_wdy, m_wdx2, m_wdxdy, m_wdy2};
      const float grp = wave_reduce8(t, lane);
      const float sw = wave_reduce1(m_w);
      if ((lane & 7) == 0) atomicAdd(&s_acc[lane >> 3][j], grp);
      if (lane == 0) atomicAdd(&s_acc[8][j], sw);
    }
    __syncthreads();
    if (tid < cnt) {
      // moments -> the reference's per-pair terms, once per (tile, Gaussian):
      //   dL/dmean2D = -o (0.5W (A Swdx + B Swdy), 0.5H (C Swdy + B Swdx))
      //   dL/dconic  = -0.5 o (Swdx2, Swdxdy, Swdy2),  dL/dopacity = Sw,  dL/dcolor = Sv dL/dpix
      const float4 co = s_co[tid];
      const float sv0 = s_acc[0][tid], sv1 = s_acc[1][tid], sv2 = s_acc[2][tid];
      const float swdx = s_acc[3][tid], swdy = s_acc[4][tid];
      const float swdx2 = s_acc[5][tid], swdxdy = s_acc[6][tid], swdy2 = s_acc[7][tid];
      const float sw = s_acc[8][tid];
      const float v[kGradTerms] = {sv0,
                                   sv1,
                                   sv2,
                                   -co.w * ddelx_dx * (co.x * swdx + co.y * swdy),
                                   -co.w * ddely_dy * (co.z * swdy + co.y * swdx),
                                   -0.5f * co.w * swdx2,
                                   -0.5f * co.w * swdxdy,
                                   -0.5f * co.w * swdy2,
                                   sw};
      const uint32_t g = s_gid[tid];
      float* dst[kGradTerms] = {&dL_dcolors[3 * g],     &dL_dcolors[3 * g + 1], &dL_dcolors[3 * g + 2],
                                &dL_dmean2D[3 * g],     &dL_dmean2D[3 * g + 1], &dL_dconic[4 * g],
                                &dL_dconic[4 * g + 1],  &dL_dconic[4 * g + 3],  &dL_dopacity[g]};
#pragma unroll
      for (int qq = 0; qq < kGradTerms; ++qq)
        if (v[qq] != 0.f) unsafeAtomicAdd(dst[qq], v[qq]);
    }
  }
}

namespace {

// Dispatches a kernel template over (tile shape, wrap).
template <bool kMulti, template <int, int, bool, bool> class K, class... Args>
hipError_t dispatch_shape(int bx, int by, bool wrap, unsigned blocks, hipStream_t s, Args... args) {
#define OMNIGS_SHAPE(TX, TY)                                                                                \
  if (bx == TX && by == TY) {                                                                               \
    if (wrap)                                                                                               \
      hipLaunchKernelGGL((K<TX, TY, true, kMulti>::fn), dim3(blocks), dim3(kTilePixels), 0, s, args...);   \
    else                                                                                                    \
      hipLaunchKernelGGL((K<TX, TY, false, kMulti>::fn), dim3(blocks), dim3(kTilePixels), 0, s, args...);  \
    return hipGetLastError();                                                                               \
  }
  OMNIGS_SHAPE(16, 16)
  OMNIGS_SHAPE(32, 8)
  OMNIGS_SHAPE(8, 32)
#undef OMNIGS_SHAPE
  return hipErrorInvalidValue;
}

template <int TX, int TY, bool kWrap, bool kMulti>
struct FwdK {
  static constexpr auto fn = render_forward_kernel<TX, TY, kWrap, kMulti>;
};
template <int TX, int TY, bool kWrap, bool kMulti>
struct BwdK {
  static constexpr auto fn = render_backward_kernel<TX, TY, kWrap, kMulti>;
};

}  // namespace

bool tile_shape_supported(int bx, int by) {
  return (bx == 16 && by == 16) || (bx == 32 && by == 8) || (bx == 8 && by == 32);
}

hipError_t launch_render_forward(TileGrid grid, const uint2* ranges, const uint32_t* point_list, int W, int H,
                                 const float4* rec, float* final_T, uint32_t* n_contrib, const float* bg,
                                 float* out_color, hipStream_t s) {
  const int tiles = grid.tiles();
  if (tiles == 0) return hipSuccess;
  ViewTable none{};
  return dispatch_shape<false, FwdK>(grid.bx, grid.by, grid.wrap != 0, (unsigned)tiles, s, grid.gx, ranges,
                                     point_list, W, H, none, rec, final_T, n_contrib, bg, out_color);
}

hipError_t launch_render_backward(TileGrid grid, const uint2* ranges, const uint32_t* point_list, int W, int H,
                                  const float* bg, const float4* rec, const float* final_T, const uint32_t* n_contrib,
                                  const float* dL_dpix, float* dL_dmean2D, float* dL_dconic, float* dL_dopacity,
                                  float* dL_dcolors, hipStream_t s) {
  const int tiles = grid.tiles();
  if (tiles == 0) return hipSuccess;
  ViewTable none{};
  return dispatch_shape<false, BwdK>(grid.bx, grid.by, grid.wrap != 0, (unsigned)tiles, s, grid.gx, ranges,
                                     point_list, W, H, none, bg, rec, final_T, n_contrib, dL_dpix, dL_dmean2D,
                                     dL_dconic, dL_dopacity, dL_dcolors);
}

hipError_t launch_render_forward_views(const ViewTable& vt, const uint2* ranges, const uint32_t* point_list,
                                       const float4* rec, float* final_T, uint32_t* n_contrib, const float* bg,
                                       float* out_color, hipStream_t s) {
  const uint32_t tiles = vt.tile_base[vt.V];
  if (tiles == 0) return hipSuccess;
  return dispatch_shape<true, FwdK>(vt.bx, vt.by, vt.wrap != 0, tiles, s, 0, ranges, point_list, 0, 0, vt, rec,
                                    final_T, n_contrib, bg, out_color);
}

hipError_t launch_render_backward_views(const ViewTable& vt, const uint2* ranges, const uint32_t* point_list,
                                        const float* bg, const float4* rec, const float* final_T,
                                        const uint32_t* n_contrib, const float* dL_dpix, float* dL_dmean2D,
                                        float* dL_dconic, float* dL_dopacity, float* dL_dcolors, hipStream_t s) {
  const uint32_t tiles = vt.tile_base[vt.V];
  if (tiles == 0) return hipSuccess;
  return dispatch_shape<true, BwdK>(vt.bx, vt.by, vt.wrap != 0, tiles, s, 0, ranges, point_list, 0, 0, vt, bg, rec,
                                    final_T, n_contrib, dL_dpix, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors);
}

}  // namespace omnigs
