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
// and reduces every Gaussian's nine gradient terms across the wave with
// permlane/DPP adds (two instances per reduce-scatter), then across the four
// waves in LDS.  Each (tile, Gaussian) instance's terms go to its own record
// with plain stores, and a per-Gaussian pass (gather_grads) sums the records,
// in place of device-scope atomics on the per-Gaussian arrays.
#include "kernels.hpp"
#include "wave_ops.hpp"
#include "grad_gather.hpp"

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
  uint32_t key;  // the tile's entry in the ranges
};
// tile0 (single view): the tile of block 0 (a grid over a band of tile rows).
template <bool kMulti>
__device__ __forceinline__ TileLoc locate(int W, int H, int gx, const ViewTable& vt, int tile0 = 0) {
  if constexpr (kMulti) {
    const int v = vt.view_of_tile(blockIdx.x);
    const int t = (int)(blockIdx.x - vt.tile_base[v]);
    return {vt.W[v], vt.H[v], vt.gx[v], t, (size_t)vt.pix_base[v], vt.key_base(v) + (uint32_t)t};
  } else {
    return {W, H, gx, tile0 + (int)blockIdx.x, 0, (uint32_t)(tile0 + (int)blockIdx.x)};
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
// d^T Q d <= tau2 = 2 ln(255 o) the alpha test fails for every pixel, so
// skipping those (wave, instance) pairs changes no result.  A quadrant (the
// square of its 64 pixel centres) is dropped when one of four axes separates
// it from the ellipse: x and y (the ellipse's bounding box) and the ellipse's
// two principal axes u_i, along which it reaches sqrt(tau2 / lambda_i) --
// anisotropic Gaussians are much narrower than their box.  Extents are padded
// by 1% + 0.05 px against rounding; degenerate conics get the full mask.
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
  // principal axes: lambda_1,2 = mid +- disc; u1 along the larger eigenvalue
  const float mid = 0.5f * (co.x + co.z);
  const float disc = sqrtf(0.25f * (co.x - co.z) * (co.x - co.z) + co.y * co.y);
  const float l1 = mid + disc, l2 = mid - disc;
  float ux = co.y, uy = l1 - co.x;                     // (Q - l1 I) u = 0, first row
  const float vx = l1 - co.z, vy = co.y;               // second row: better conditioned when A > C
  if (vx * vx + vy * vy > ux * ux + uy * uy) { ux = vx; uy = vy; }
  const float un = ux * ux + uy * uy;
  const bool axes = un > 0.f && l2 > 0.f;              // else a circle or a degenerate axis: box test only
  const float ui = axes ? rsqrtf(un) : 0.f;
  ux *= ui;
  uy *= ui;
  // half extents of the ellipse and of the 7x7-pixel-centre square along u1, u2 = (-uy, ux)
  const float sq = 3.5f * (fabsf(ux) + fabsf(uy));
  const float e1 = axes ? sqrtf(tau2 / l1) * 1.01f + 0.05f + sq : 0.f;
  const float e2 = axes ? sqrtf(tau2 / l2) * 1.01f + 0.05f + sq : 0.f;
  const float x0 = (float)((tile % gx) * TX), y0 = (float)((tile / gx) * TY);
  if (wrapW > 0.f) m.x -= wrapW * rintf((m.x - (x0 + 0.5f * TX)) / wrapW);  // nearest copy across the seam
  uint32_t mk = 0;
#pragma unroll
  for (int q = 0; q < kWavesPerTile; ++q) {
    const float qx = x0 + (float)((q % QX) * 8), qy = y0 + (float)((q / QX) * 8);
    bool hit = (m.x + ex >= qx) && (m.x - ex <= qx + 7.f) && (m.y + ey >= qy) && (m.y - ey <= qy + 7.f);
    if (axes) {
      const float cx = qx + 3.5f - m.x, cy = qy + 3.5f - m.y;  // square centre relative to the mean
      hit = hit && fabsf(cx * ux + cy * uy) <= e1 && fabsf(cy * ux - cx * uy) <= e2;
    }
    mk |= hit ? (1u << q) : 0u;
  }
  return mk;
}

// A staged instance: {x, y, a, b}, {c, opacity, r, g}, {blue, ...} with
// (a, b, c) the prescaled conic, 48 bytes, so that one LDS address register
// (the list entry, a byte offset) reaches every operand of a pair.
constexpr int kStageBytes = 48;
struct alignas(16) Staged {
  float4 a, b, c;
};
// An element of an LDS array addressed by its byte offset.
template <typename T, typename A>
__device__ __forceinline__ T& lds_at(A* base, int off) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off);
}
template <typename T, typename A>
__device__ __forceinline__ const T& lds_at(const A* base, int off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
__device__ __forceinline__ const Staged& staged_at(const Staged* base, int off) {
  return *reinterpret_cast<const Staged*>(reinterpret_cast<const char*>(base) + off);
}

// Each wave compacts the batch slots whose mask has its bit set into its own
// ascending list, entry j * kStride (the forward stores staging byte offsets,
// the backward 4-byte-element offsets); returns the list length
// (wave-uniform).
template <int kStride, typename Entry>
__device__ __forceinline__ int build_wave_list(const uint8_t* s_mask, Entry* list, int cnt, int w, int lane) {
  int count = 0;
  const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
  for (int c = 0; c < kBatch / kWave; ++c) {
    const int j = c * kWave + lane;
    const bool on = j < cnt && ((s_mask[j] >> w) & 1u);
    const uint64_t bal = __ballot(on);
    if (on) list[count + __popcll(bal & below)] = (Entry)(j * kStride);
    count += __popcll(bal);
  }
  return count;
}

template <int TX, int TY, bool kWrap, bool kMulti>
__global__ __launch_bounds__(kTilePixels) void render_forward_kernel(
    int gx_, const uint2* __restrict__ ranges, const uint32_t* __restrict__ point_list, int W_, int H_,
    ViewTable vt, const float4* __restrict__ rec, float* __restrict__ final_T, uint32_t* __restrict__ n_contrib,
    const float* __restrict__ bg, float* __restrict__ out_color, uint8_t* __restrict__ blend_out,
    uint32_t* __restrict__ blend_ok) {
  const TileLoc loc = locate<kMulti>(W_, H_, gx_, vt);
  const int W = loc.W, H = loc.H, gx = loc.gx;
  const float Wf = (float)W, invW = 1.0f / (float)W;
  __shared__ Staged s_op[kBatch];
  __shared__ uint8_t s_mask[kBatch];
  __shared__ uint16_t s_list[kWavesPerTile][kBatch];

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), w = tid / kWave;
  const int tile = loc.tile;
  int px, py;
  pixel_of<TX, TY>(tile, gx, tid, px, py);
  const bool inside = px < W && py < H;
  const float fx = (float)px, fy = (float)py;
  const uint2 range = ranges[loc.key];
  const int n = (int)(range.y - range.x);

  // Lanes that are done (outside the image, or stopped), as a wave-uniform
  // lane mask: the per-pair decisions below are lane masks combined on the
  // scalar unit rather than per-lane booleans.
  uint64_t done_mask = __builtin_amdgcn_ballot_w64(!inside);
  float T = 1.0f;
  float C0 = 0.f, C1 = 0.f, C2 = 0.f;
  uint32_t last_b = 0;  // n_contrib times kStageBytes

  if (blend_out && blockIdx.x == 0 && tid == 0) *blend_ok = 1u;
  for (int base = 0;; base += kBatch) {
    const int all_done = __syncthreads_count(__builtin_amdgcn_inverse_ballot_w64(done_mask));
    // the previous batch's blend bytes (every wave has finished it): byte w of
    // the staged entry is 1 when some lane of wave w blended it
    if (blend_out && base > 0 && base - kBatch + tid < n) {
      const uint32_t f = __float_as_uint(s_op[tid].c.y);
      blend_out[range.x + (base - kBatch + tid)] =
          (uint8_t)((f & 1u) | ((f >> 7) & 2u) | ((f >> 14) & 4u) | ((f >> 21) & 8u));
    }
    if (base >= n || all_done == kTilePixels) break;
    const int k = base + tid;
    if (k < n) {
      const uint32_t g = point_list[range.x + k];
      const RenderRec rr = load_rec(rec, g);
      const float2 m = rr.xy;
      const float4 co = rr.co;
      const float3 q = prescale_conic(co);
      s_op[tid].a = make_float4(m.x, m.y, q.x, q.y);
      s_op[tid].b = make_float4(q.z, co.w, rr.rgb.x, rr.rgb.y);
      s_op[tid].c = make_float4(rr.rgb.z, 0.f, 0.f, 0.f);  // y: one blend byte per wave
      s_mask[tid] = (uint8_t)quadrant_mask<TX, TY>(m, co, tile, gx, kWrap ? Wf : 0.f);
    }
    __syncthreads();
    const int cnt = min(kBatch, n - base);
    const int len = build_wave_list<kStageBytes>(s_mask, s_list[w], cnt, w, lane);
    // a blended pair's 1-based position base + j + 1, in staging bytes
    const uint32_t pos_bias = (uint32_t)(base + 1) * kStageBytes;
    // Branch-free over the wave's list: every lane evaluates every pair and
    // the reference's skip / stop decisions become selects, so the loop stays
    // wave-uniform; it ends once every lane of the wave is done.  The next
    // pair's operands are fetched from LDS while the current one computes.
    int jo = s_list[w][0];  // a stale entry when len == 0: never used
    float4 pa = staged_at(s_op, jo).a, pb = staged_at(s_op, jo).b;
    float c2 = staged_at(s_op, jo).c.x;
    for (int ii = 0; ii < len; ++ii) {
      const int jn = s_list[w][min(ii + 1, len - 1)];
      const float dx = seam_dx<kWrap>(pa.x - fx, Wf, invW), dy = pa.y - fy;
      const float power2 = power2_of(make_float3(pa.z, pa.w, pb.x), dx, dy);
      const float alpha = fminf(0.99f, pb.y * exp2_hw(power2));
      const float test_T = T * (1.0f - alpha);
      // power2 > 0 or alpha < 1/255: skipped; test_T < 1e-4: stop before blending
      const uint64_t ok = __builtin_amdgcn_ballot_w64(power2 <= 0.0f) &
                          __builtin_amdgcn_ballot_w64(alpha >= 1.0f / 255.0f) & ~done_mask;
      const uint64_t stop = ok & __builtin_amdgcn_ballot_w64(test_T < 0.0001f);
      const bool blend = __builtin_amdgcn_inverse_ballot_w64(ok & ~stop);
      done_mask |= stop;
      // whether some lane of this wave blended the instance: every lane
      // stores the same byte at the same address (the wave's own byte)
      reinterpret_cast<uint8_t*>(s_op)[jo + 36 + w] = (ok & ~stop) != 0ull ? 1 : 0;
      const float wgt = blend ? alpha * T : 0.f;
      C0 += pb.z * wgt;
      C1 += pb.w * wgt;
      C2 += c2 * wgt;
      T = blend ? test_T : T;
      last_b = blend ? pos_bias + (uint32_t)jo : last_b;  // 1-based position in the tile's list, in bytes
      jo = jn;
      pa = staged_at(s_op, jo).a;
      pb = staged_at(s_op, jo).b;
      c2 = staged_at(s_op, jo).c.x;
      if (done_mask == ~0ull) break;  // all 64 lanes done
    }
  }

  if (inside) {
    const size_t pid = (size_t)py * W + px;
    const size_t N = (size_t)W * H;
    final_T[loc.pix + pid] = T;
    n_contrib[loc.pix + pid] = last_b / (uint32_t)kStageBytes;
    float* out = out_color + 3 * loc.pix;
    out[pid] = C0 + T * bg[0];
    out[N + pid] = C1 + T * bg[1];
    out[2 * N + pid] = C2 + T * bg[2];
  }
}

// Reduce-scatter of two instances' eight terms (a[0..7] of instance A, b[0..7]
// of instance B) in one pass: on return each 4-lane group holds one complete
// sum, instance (lane >> 5) and term ((lane >> 2) & 7).  35 VALU operations
// for 16 sums, against 18 for eight with wave_reduce8.
__device__ __forceinline__ float wave_reduce16(const float a[8], const float b[8], int lane) {
  // lanes 0-31 carry A's terms, lanes 32-63 B's
  float s1[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) s1[k] = fold32(a[k], b[k]);
  // 16-lane rows: 0 -> A t[k], 1 -> A t[k+4], 2 -> B t[k], 3 -> B t[k+4]
  float s2[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) s2[k] = fold16(s1[k], s1[k + 4]);
  // half rows: lanes 0-7 keep t[k], lanes 8-15 keep t[k+2]
  const bool h8 = (lane >> 3) & 1;
  float s3[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const float keep = h8 ? s2[k + 2] : s2[k], give = h8 ? s2[k] : s2[k + 2];
    s3[k] = keep + dpp<0x128>(give);  // row_ror:8
  }
  // quarter rows: lanes 0-3 of each 8 keep t[0 or 2 (+4)], lanes 4-7 t[1 or 3 (+4)]
  const bool h4 = (lane >> 2) & 1;
  const float keep = h4 ? s3[1] : s3[0], give = h4 ? s3[0] : s3[1];
  float c = keep + dpp<0x141>(give);  // row_half_mirror: lane i <-> 7 - i
  c += dpp<0x4e>(c);                  // quad_perm [2,3,0,1]
  c += dpp<0xb1>(c);                  // quad_perm [1,0,3,2]
  return c;
}

// Full wave sums of two terms: lanes 0-31 receive sum(x), lanes 32-63 sum(y).
__device__ __forceinline__ float wave_reduce2(float x, float y) {
  float c = fold32(x, y);
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

constexpr int kAccStride = kBatch + 1;  // padded so one instance's terms sit in distinct banks

template <int TX, int TY, bool kWrap, bool kMulti>
// Six blocks per CU fit the kernel's LDS; the register budget is held to
// match (80 VGPRs), which the compiler otherwise exceeds by hoisting staging
// loads.
__global__ __launch_bounds__(kTilePixels) __attribute__((amdgpu_waves_per_eu(6, 8))) void render_backward_kernel(
    int gx_, const uint2* __restrict__ ranges, const uint32_t* __restrict__ point_list, int W_, int H_,
    ViewTable vt, const float* __restrict__ bg, const float4* __restrict__ rec, const float* __restrict__ final_Ts,
    const uint32_t* __restrict__ n_contrib, const float* __restrict__ dL_dpix, const uint32_t* __restrict__ slot_base,
    uint32_t nslots, float4* __restrict__ grad_rec, float* __restrict__ grad_opacity,
    const uint8_t* __restrict__ blend_in, const uint32_t* __restrict__ blend_ok, int band_y0,
    const uint32_t* __restrict__ slot_words, const uint2* __restrict__ blend_ranges) {
  // band_y0 (single view): the grid covers the tile rows from band_y0 on, and
  // the lists and gradient slots hold only the rect rows from band_y0 on;
  // slot_words (a band's): each Gaussian's first slot, in place of slot_base
  // plus the record's slot word; blend_ranges: the ranges of the list the
  // blend bytes belong to (position p of a tile's list is entry
  // blend_ranges[tile].x + p there)
  const TileLoc loc = locate<kMulti>(W_, H_, gx_, vt, band_y0 * gx_);
  const int W = loc.W, H = loc.H, gx = loc.gx;
  const float Wf = (float)W, invW = 1.0f / (float)W;
  const TileGrid tg{gx, (H + TY - 1) / TY, kWrap ? 1 : 0, TX, TY};
  __shared__ uint32_t s_inst[kBatch];  // gradient slot of the staged instance
  // the forward's layout: {x, y, a, b}, {c, opacity, r, g}, blue
  __shared__ float4 s_pa[kBatch];
  __shared__ float4 s_pb[kBatch];
  __shared__ float s_blue[kBatch];
  __shared__ float s_acc[kGradTerms][kAccStride];
  __shared__ uint8_t s_mask[kBatch];
  // per wave, the slots it evaluates as 4 j (the byte offset of slot j in the
  // 4-byte arrays; 16-byte ones take 4 times that)
  __shared__ uint16_t s_list[kWavesPerTile][kBatch];
  // positions still to stage, back to front, as (position << 4 | wave bits),
  // in a ring of 4 kBatch entries from q_head
  __shared__ uint32_t s_queue[4 * kBatch];
  __shared__ int s_wlast[kWavesPerTile];
  __shared__ int s_nlive;

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1), w = tid / kWave;
  const int tile = loc.tile;
  int px, py;
  pixel_of<TX, TY>(tile, gx, tid, px, py);
  const bool inside = px < W && py < H;
  const float fx = (float)px, fy = (float)py;
  const uint2 range = ranges[loc.key];
  const uint32_t blend_x = blend_ranges[loc.key].x;
  const size_t N = (size_t)W * H;
  const size_t pid = (size_t)py * W + px;
  final_Ts += loc.pix;
  n_contrib += loc.pix;
  dL_dpix += 3 * loc.pix;

  const bool use_blend = blend_in && *blend_ok == 1u;
  const float T_final = inside ? final_Ts[pid] : 0.f;
  float T = T_final;
  const int last = inside ? (int)n_contrib[pid] : 0;
  // Positions >= max(last) over the tile blend into no pixel: the traversal
  // starts there instead of at the end of the list.  Per wave, positions
  // >= the wave's own max(last) are dropped from its lists.
  if (tid == 0) s_nlive = 0;
  __syncthreads();
  const int wlast = wave_max_int(last);
  if (lane == 0) {
    atomicMax(&s_nlive, wlast);
    s_wlast[w] = wlast;
  }
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
  // The reference carries the colour accumulated behind each instance per
  // channel, acc_c <- la lc_c + (1 - la) acc_c, and only ever uses
  // sum_c (c_c - acc_c) dL/dpix_c; the same recurrence on the dot product
  // s = acc . dL/dpix (with lcd = lc . dL/dpix) gives that sum directly.  It
  // is advanced right after each blended instance rather than at the next
  // one, which reuses c . dL/dpix - s and carries no last alpha or colour.
  float acc_dot = 0.f;

  // One (wave, instance) pair: updates the lanes' blending state and returns,
  // wave-uniformly, whether any lane blends it.  If one does, (wgt, v) are
  // G dL/dalpha and alpha T on blending lanes (zero elsewhere) and (dx, dy)
  // the offsets; the nine moments follow from these four (moments()).
  // The instance is slot j, passed as j4 = 4 j (a list entry).  A lane blends
  // it only before its last contributor (its position < last <=> j4 >= lo4,
  // positions falling with j).  `search`: the pair may have no blending lane
  // (quadrant-mask lists); with the forward's blend bits it always has one
  // and the test is skipped.
  auto eval = [&](int j4, int lo4, float& wgt, float& v, float& dx, float& dy, bool search) -> bool {
    const float4 pa = lds_at<float4>(s_pa, 4 * j4), pb = lds_at<float4>(s_pb, 4 * j4);
    dx = seam_dx<kWrap>(pa.x - fx, Wf, invW);
    dy = pa.y - fy;
    const float power2 = power2_of(make_float3(pa.z, pa.w, pb.x), dx, dy);
    const float G = exp2_hw(power2);
    const float alpha = fminf(0.99f, pb.y * G);
    // three compares whose lane masks combine on the scalar unit (a ballot of
    // the combined bool would rematerialise it as a vector value first)
    const bool c_pos = j4 >= lo4, c_pow = power2 <= 0.0f, c_alpha = alpha >= 1.0f / 255.0f;
    const bool ok = c_pos && c_pow && c_alpha;
    wgt = 0.f;
    v = 0.f;
    if (search && (__builtin_amdgcn_ballot_w64(c_pos) & __builtin_amdgcn_ballot_w64(c_pow) &
                   __builtin_amdgcn_ballot_w64(c_alpha)) == 0ull)
      return false;
    // evaluated by the whole wave (some lane always blends), which keeps the
    // operand loads out of the branch
    const float cd = pb.z * dp0 + pb.w * dp1 + lds_at<float>(s_blue, j4) * dp2;
    if (ok) {
      // T and 1/(1 - alpha) only scale gradients here (no decision depends
      // on them), so the hardware reciprocal replaces two IEEE divisions.
      const float r = __builtin_amdgcn_rcpf(1.f - alpha);
      T = T * r;
      v = alpha * T;
      const float diff = cd - acc_dot;
      const float dL_dalpha = diff * T + neg_tb * r;
      // the colour behind the next blended instance: alpha cd + (1 - alpha) acc
      acc_dot += alpha * diff;
      wgt = G * dL_dalpha;
    }
    return true;
  };
  // per-pixel moments: dL/dpix-weighted v for the colours, w dx, w dy, w dx^2, w dx dy, w dy^2
  auto moments = [&](float wgt, float v, float dx, float dy, float t[8]) {
    const float wdx = wgt * dx, wdy = wgt * dy;
    t[0] = v * dp0;
    t[1] = v * dp1;
    t[2] = v * dp2;
    t[3] = wdx;
    t[4] = wdy;
    t[5] = wdx * dx;
    t[6] = wdx * dy;
    t[7] = wdy * dy;
  };

  // The list is walked from position n - 1 down.  Positions that some wave
  // must evaluate -- with the forward's blend bits, those some wave blended
  // (often under half of them) -- enter a queue in LDS, 512 positions (two
  // per thread) per block scan, and batches of up to 256 are staged from its
  // front, so batches hold no instance that no wave evaluates.  The queue
  // holds under kBatch entries before a fill, so at most 3 kBatch - 1 after.
  constexpr uint32_t kRing = 4 * kBatch - 1;
  int q_head = 0, q_len = 0, scanned = 0;  // block-uniform
  for (;;) {
    __syncthreads();
    while (q_len < kBatch && scanned < n) {
      const int k = scanned + 2 * tid;  // positions n - 1 - k and n - 2 - k
      uint32_t b0 = 0u, b1 = 0u;
      if (k < n) b0 = use_blend ? blend_in[blend_x + (n - 1 - k)] & 0xFu : 0xFu;
      if (k + 1 < n) b1 = use_blend ? blend_in[blend_x + (n - 2 - k)] & 0xFu : 0xFu;
      uint32_t added;
      uint32_t at = block_exclusive_scan((b0 ? 1u : 0u) + (b1 ? 1u : 0u), &added);
      if (b0) s_queue[(q_head + q_len + (int)at++) & kRing] = ((uint32_t)(n - 1 - k) << 4) | b0;
      if (b1) s_queue[(q_head + q_len + (int)at) & kRing] = ((uint32_t)(n - 2 - k) << 4) | b1;
      q_len += __builtin_amdgcn_readfirstlane((int)added);  // uniform: kept in a scalar register
      scanned += 2 * kBatch;
    }
    if (q_len == 0) break;
    __syncthreads();
    const int cnt = min(kBatch, q_len);
    s_mask[tid] = 0;
    s_inst[tid] = kNoRecord;  // >= nslots: no record written
    if (tid < cnt) {
      const uint32_t e = s_queue[(q_head + tid) & kRing], p = e >> 4, bits = e & 0xFu;
      const uint32_t g = point_list[range.x + p];
      const RenderRec rr = load_rec(rec, g);
      const float2 m = rr.xy;
      const float4 co = rr.co;
      const float3 q = prescale_conic(co);
      // this instance's gradient slot: the Gaussian's first slot (its block's
      // base plus its offset in the block) plus the tile's position in its rect
      const float4 r2 = rec[3 * g + 2];
      Rect rc = tile_rect(m, __float_as_int(r2.y), tg);
      rc.y0 = max(rc.y0, band_y0);
      const uint32_t gauss = kMulti ? g % (uint32_t)vt.P : g;
      const uint32_t first =
          slot_words ? slot_words[gauss] : slot_base[gauss / kSlotBlock] + __float_as_uint(r2.z);
      s_inst[tid] = first + rect_local(rc, max(1, rc.x1 - rc.x0), tile % gx, tile / gx, tg);
      s_pa[tid] = make_float4(m.x, m.y, q.x, q.y);
      s_pb[tid] = make_float4(q.z, co.w, rr.rgb.x, rr.rgb.y);
      s_blue[tid] = rr.rgb.z;
      // the waves that blended the instance in the forward (exactly the
      // pairs whose evaluation here finds a blending lane); else the
      // quadrants the ellipse reaches, for the waves with a lane whose last
      // contributor lies behind it
      uint32_t mk = bits;
      if (!use_blend) {
        mk = quadrant_mask<TX, TY>(m, co, tile, gx, kWrap ? Wf : 0.f);
#pragma unroll
        for (int qw = 0; qw < kWavesPerTile; ++qw)
          if ((int)p >= s_wlast[qw]) mk &= ~(1u << qw);
      }
      s_mask[tid] = (uint8_t)mk;
    }
#pragma unroll
    for (int qq = 0; qq < kGradTerms; ++qq) s_acc[qq][tid] = 0.f;
    __syncthreads();
    // lo = the number of staged entries at or past this lane's last
    // contributor (a prefix: positions fall along the batch); the batch's
    // queue entries stay in place until the next fill
    int lo = 0;
#pragma unroll
    for (int step = kBatch; step > 0; step >>= 1)
      if (lo + step <= cnt && (int)(s_queue[(q_head + lo + step - 1) & kRing] >> 4) >= last) lo += step;
    const int lo4 = 4 * lo;
    q_head = (q_head + cnt) & kRing;
    q_len -= cnt;

    const int len = build_wave_list<4>(s_mask, s_list[w], cnt, w, lane);
    // The list order (back to front) is kept.  Lanes outside the image have
    // last = 0, so no position passes their limit.
    // sums of one instance (wave_reduce8 + 1) or of two (wave_reduce16 + 2),
    // added to the block's accumulators
    auto add_one = [&](int ja, float wa, float va, float xa, float ya) {
      float ta[8];
      moments(wa, va, xa, ya, ta);
      float grp = wave_reduce8(ta, lane);
      float sw = wave_reduce1(wa);
      asm volatile("" : "+v"(grp), "+v"(sw));
      if ((lane & 7) == 0) atomicAdd(&lds_at<float>(s_acc[lane >> 3], ja), grp);
      if (lane == 0) atomicAdd(&lds_at<float>(s_acc[8], ja), sw);
    };
    auto add_two = [&](int ja, float wa, float va, float xa, float ya, int jb, float wb, float vb, float xb,
                       float yb) {
      float ta[8], tb[8];
      moments(wa, va, xa, ya, ta);
      moments(wb, vb, xb, yb, tb);
      float grp = wave_reduce16(ta, tb, lane);
      float sw = wave_reduce2(wa, wb);
      // materialise both sums here, so that their last DPP add stays fused
      // instead of being sunk into the atomics' lane branches
      asm volatile("" : "+v"(grp), "+v"(sw));
      const int jj = (lane & 32) ? jb : ja;
      if ((lane & 3) == 0) atomicAdd(&lds_at<float>(s_acc[(lane >> 2) & 7], jj), grp);
      if ((lane & 31) == 0) atomicAdd(&lds_at<float>(s_acc[8], jj), sw);
    };
    int ii = 0;
    if (use_blend) {
      // Every listed pair has a blending lane (the forward's bits), so the
      // instances are reduced two consecutive entries at a time, with no
      // search between them.
      for (; ii + 1 < len; ii += 2) {
        // ii is even: both entries in one 4-byte read
        const uint32_t e = *reinterpret_cast<const uint32_t*>(&s_list[w][ii]);
        const int ja = (int)(e & 0xFFFFu), jb = (int)(e >> 16);
        float wa, va, xa, ya, wb, vb, xb, yb;
        eval(ja, lo4, wa, va, xa, ya, false);
        eval(jb, lo4, wb, vb, xb, yb, false);
        add_two(ja, wa, va, xa, ya, jb, wb, vb, xb, yb);
      }
      if (ii < len) {
        const int ja = s_list[w][ii];
        float wa, va, xa, ya;
        eval(ja, lo4, wa, va, xa, ya, false);
        add_one(ja, wa, va, xa, ya);
      }
    } else {
      // Pairs that some lane blends are found by evaluating the list in order
      // and reduced two at a time; a last odd one alone.
      for (;;) {
        float wa, va, xa, ya, wb, vb, xb, yb;
        int ja = 0, jb = 0;
        bool fa = false, fb = false;
        while (ii < len && !fa) {
          ja = s_list[w][ii++];
          fa = eval(ja, lo4, wa, va, xa, ya, true);
        }
        if (!fa) break;
        while (ii < len && !fb) {
          jb = s_list[w][ii++];
          fb = eval(jb, lo4, wb, vb, xb, yb, true);
        }
        if (!fb) {
          add_one(ja, wa, va, xa, ya);
          break;
        }
        add_two(ja, wa, va, xa, ya, jb, wb, vb, xb, yb);
      }
    }
    __syncthreads();
    if (tid < cnt) {
      // moments -> the reference's per-pair terms, once per (tile, Gaussian):
      //   dL/dmean2D = -o (0.5W (A Swdx + B Swdy), 0.5H (C Swdy + B Swdx))
      //   dL/dconic  = -0.5 o (Swdx2, Swdxdy, Swdy2),  dL/dopacity = Sw,  dL/dcolor = Sv dL/dpix
      // (A, B, C) from the staged prescaled conic
      const float4 pa = s_pa[tid], pb = s_pb[tid];
      const float4 co = make_float4(pa.z * (-2.0f / kLog2e), pa.w * (-1.0f / kLog2e), pb.x * (-2.0f / kLog2e), pb.y);
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
      // plain stores, zeros included: every staged instance owns its record
      // (slots past the capacity exist only when a batch overflowed)
      const uint32_t j = s_inst[tid];
      if (j < nslots) {
        grad_rec[2 * (size_t)j] = make_float4(v[0], v[1], v[2], v[3]);
        grad_rec[2 * (size_t)j + 1] = make_float4(v[4], v[5], v[6], v[7]);
        grad_opacity[j] = v[8];
      }
    }
  }
}

// Per-Gaussian (per-item) sums of the gradient records (gather_sum), written
// in full, zeros included, so the outputs need no clearing.
__global__ __launch_bounds__(256) void gather_grads_kernel(GatherSrc src, int items, float* __restrict__ dL_dmean2D,
                                                           float* __restrict__ dL_dconic,
                                                           float* __restrict__ dL_dopacity,
                                                           float* __restrict__ dL_dcolors) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < items;  // every lane stays to the end: the wave-wide path needs them all
  float acc[kGradTerms];
  gather_sum(src, i, valid, acc);
  if (valid) store_gathered((size_t)i, acc, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors);
}

// One band's records added to the per-Gaussian sums, one Gaussian per lane in
// index order (as the whole-image gather): Gaussian i owns the slots
// [slot_words[i], + its band area) (written by the band's owner pass).  Waves
// with no Gaussian in the band leave at once.
__global__ __launch_bounds__(256) void band_gather_kernel(const uint2* __restrict__ dup_rect, uint32_t by0,
                                                          uint32_t by1, int P, const uint32_t* __restrict__ slot_words,
                                                          uint32_t nslots, const float4* __restrict__ grad_rec,
                                                          const float* __restrict__ grad_opacity,
                                                          float* __restrict__ dL_dmean2D, float* __restrict__ dL_dconic,
                                                          float* __restrict__ dL_dopacity,
                                                          float* __restrict__ dL_dcolors, bool zero_empty) {
  // The band holding the first row of a Gaussian's rect writes its sums;
  // later bands add theirs to them.  zero_empty (the first band launched):
  // Gaussians with no instance at all get their zeros here, so no band
  // reads an output that was not written.
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint2 full = i < P ? dup_rect[i] : make_uint2(0u, 1u << 16);
  const uint32_t cnt = dup_rect_area(band_dup_rect(full, by0, by1));
  const bool zero = zero_empty && i < P && dup_rect_area(full) == 0u;
  if (__ballot(cnt != 0u || zero) == 0ull) return;  // wave-uniform
  const uint32_t base = cnt ? slot_words[i] : 0u;
  float acc[kGradTerms];
#pragma unroll
  for (int q = 0; q < kGradTerms; ++q) acc[q] = 0.f;
  if (cnt && (full.y & 0xFFFFu) < by0) load_gathered((size_t)i, acc, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors);
  gather_range<false>(grad_rec, grad_opacity, nslots, base, cnt, acc);
  if (cnt || zero) store_gathered((size_t)i, acc, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors);
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
                                 float* out_color, uint8_t* blend, uint32_t* blend_ok, hipStream_t s) {
  const int tiles = grid.tiles();
  if (tiles == 0) return hipSuccess;
  ViewTable none{};
  return dispatch_shape<false, FwdK>(grid.bx, grid.by, grid.wrap != 0, (unsigned)tiles, s, grid.gx, ranges,
                                     point_list, W, H, none, rec, final_T, n_contrib, bg, out_color, blend, blend_ok);
}

hipError_t launch_render_backward(TileGrid grid, const uint2* ranges, const uint32_t* point_list, int W, int H,
                                  const float* bg, const float4* rec, const float* final_T, const uint32_t* n_contrib,
                                  const float* dL_dpix, const uint32_t* slot_base, uint32_t nslots, float4* grad_rec,
                                  float* grad_opacity, const uint8_t* blend, const uint32_t* blend_ok, hipStream_t s,
                                  int band_y0, int band_y1, const uint32_t* slot_words, const uint2* blend_ranges) {
  if (band_y1 < 0 || band_y1 > grid.gy) band_y1 = grid.gy;
  if (band_y0 < 0 || band_y0 > band_y1) return hipErrorInvalidValue;
  const int tiles = (band_y1 - band_y0) * grid.gx;
  if (tiles == 0) return hipSuccess;
  ViewTable none{};
  return dispatch_shape<false, BwdK>(grid.bx, grid.by, grid.wrap != 0, (unsigned)tiles, s, grid.gx, ranges,
                                     point_list, W, H, none, bg, rec, final_T, n_contrib, dL_dpix, slot_base, nslots,
                                     grad_rec, grad_opacity, blend, blend_ok, band_y0, slot_words,
                                     blend_ranges ? blend_ranges : ranges);
}

hipError_t launch_gather_grads(int items, int P, uint32_t nslots, const uint32_t* R_dev,
                               const uint32_t* tiles_touched, const uint32_t* slot_local, const uint32_t* slot_base,
                               const float4* grad_rec, const float* grad_opacity, float* dL_dmean2D,
                               float* dL_dconic, float* dL_dopacity, float* dL_dcolors, hipStream_t s) {
  if (items == 0) return hipSuccess;
  const GatherSrc src{P, nslots, R_dev, tiles_touched, slot_local, slot_base, grad_rec, grad_opacity};
  hipLaunchKernelGGL(gather_grads_kernel, dim3((items + 255) / 256), dim3(256), 0, s, src, items, dL_dmean2D,
                     dL_dconic, dL_dopacity, dL_dcolors);
  return hipGetLastError();
}

hipError_t launch_band_gather(int P, const uint2* dup_rect, uint32_t by0, uint32_t by1, const uint32_t* slot_words,
                              uint32_t nslots, const float4* grad_rec, const float* grad_opacity, float* dL_dmean2D,
                              float* dL_dconic, float* dL_dopacity, float* dL_dcolors, bool zero_empty, hipStream_t s) {
  if (P == 0 || nslots == 0) return hipSuccess;
  hipLaunchKernelGGL(band_gather_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, dup_rect, by0, by1, P,
                     slot_words, nslots, grad_rec, grad_opacity, dL_dmean2D, dL_dconic, dL_dopacity, dL_dcolors,
                     zero_empty);
  return hipGetLastError();
}

hipError_t launch_render_forward_views(const ViewTable& vt, const uint2* ranges, const uint32_t* point_list,
                                       const float4* rec, float* final_T, uint32_t* n_contrib, const float* bg,
                                       float* out_color, uint8_t* blend, uint32_t* blend_ok, hipStream_t s) {
  const uint32_t tiles = vt.tile_base[vt.V];
  if (tiles == 0) return hipSuccess;
  return dispatch_shape<true, FwdK>(vt.bx, vt.by, vt.wrap != 0, tiles, s, 0, ranges, point_list, 0, 0, vt, rec,
                                    final_T, n_contrib, bg, out_color, blend, blend_ok);
}

hipError_t launch_render_backward_views(const ViewTable& vt, const uint2* ranges, const uint32_t* point_list,
                                        const float* bg, const float4* rec, const float* final_T,
                                        const uint32_t* n_contrib, const float* dL_dpix, const uint32_t* slot_base,
                                        uint32_t nslots, float4* grad_rec, float* grad_opacity, const uint8_t* blend,
                                        const uint32_t* blend_ok, hipStream_t s) {
  const uint32_t tiles = vt.tile_base[vt.V];
  if (tiles == 0) return hipSuccess;
  return dispatch_shape<true, BwdK>(vt.bx, vt.by, vt.wrap != 0, tiles, s, 0, ranges, point_list, 0, 0, vt, bg, rec,
                                    final_T, n_contrib, dL_dpix, slot_base, nslots, grad_rec, grad_opacity, blend,
                                    blend_ok, 0, (const uint32_t*)nullptr, ranges);
}

}  // namespace omnigs
