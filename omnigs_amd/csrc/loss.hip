// Fused photometric loss and its image gradient:
//   loss = (1 - lambda) mean|x - y| + lambda (1 - mean SSIM(x, y))
// over the image with its bottom `skip` rows dropped, as the reference's
// trainer computes it with torch ops (gaussian_mapper.cpp:384-407,
// loss_utils.h:31-129: 11x11 Gaussian window with sigma 1.5, zero padding,
// C1 = 0.01^2, C2 = 0.03^2, means over every element).  x is the rendered
// image times the optional undistortion mask (gaussian_mapper.cpp:385).
//
// With E1 = w*x, E2 = w*y, E11 = w*x^2, E22 = w*y^2, E12 = w*xy (separable
// 11-tap filtering, zero padded) the SSIM map is
//   m = (2 mu1 mu2 + C1)(2 s12 + C2) / ((mu1^2 + mu2^2 + C1)(s11 + s22 + C2)),
// mu = E, s11 = E11 - mu1^2, s12 = E12 - mu1 mu2, and
//   dSSIM/dx = (1/n) [w*g1 + 2 x (w*g11) + y (w*g12)]
// with g1 = dm/dmu1 - 2 mu1 dm/ds11 - mu2 dm/ds12, g11 = dm/ds11,
// g12 = dm/ds12 (the window is symmetric, so its transpose is itself).
//
// lambda != 0: a stats pass (the five filtered maps of a 32x16 tile staged in
// LDS; writes g1, g11, g12 and the tile's sums of |x - y| and m) and a
// gradient pass (filters the three adjoint maps).  lambda = 0: one streaming
// pass (float4 when W % 4 == 0) writes the gradient and the L1 block sums.
// A one-block kernel adds the block sums in a fixed order (deterministic).
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "kernels.hpp"

namespace omnigs {

constexpr int kWin = 11, kHalf = 5;
constexpr int kLTX = 32, kLTY = 16;  // output tile per 256-thread block (2 pixels per thread)
constexpr int kLInX = kLTX + 2 * kHalf, kLInY = kLTY + 2 * kHalf;
constexpr int kL1Blocks = 2048;      // block count cap of the streaming L1 pass
constexpr int kFinalThreads = 1024;

struct LossArgs {
  const float* x;     // rendered image [C,H,W]
  const float* y;     // target image [C,H,W]
  const float* mask;  // optional [H,W] multiplier of x, NULL = ones
  int C, H, W, Hc;    // Hc = rows kept (H - skip)
  float lambda;
  float win[kWin];
  float* g;        // [3, C, Hc, W] SSIM adjoint maps (g1, g11, g12)
  float* partial;  // [blocks, 2] (sum |x - y|, sum m)
  float* dL_dx;    // [C,H,W]
};

// masked rendered value x * mask (zero outside the kept rows / image)
__device__ __forceinline__ float fetch_x(const LossArgs& a, int c, int yy, int xx) {
  if (!(yy >= 0 && yy < a.Hc && xx >= 0 && xx < a.W)) return 0.f;
  const float v = a.x[((size_t)c * a.H + yy) * a.W + xx];
  return a.mask ? v * a.mask[(size_t)yy * a.W + xx] : v;
}

__device__ __forceinline__ float fetch_y(const LossArgs& a, int c, int yy, int xx) {
  return (yy >= 0 && yy < a.Hc && xx >= 0 && xx < a.W) ? a.y[((size_t)c * a.H + yy) * a.W + xx] : 0.f;
}

__device__ __forceinline__ float sgn(float d) { return d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f); }

// Horizontal 11-tap pass of K maps over the (kLInY x kLInX) input tile.
template <int K>
__device__ __forceinline__ void filter_rows(const float (*in)[kLInY][kLInX], float (*mid)[kLInY][kLTX],
                                            const float* win) {
  for (int i = threadIdx.x; i < K * kLInY * kLTX; i += blockDim.x) {
    const int k = i / (kLInY * kLTX), r = (i / kLTX) % kLInY, cx = i % kLTX;
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < kWin; ++t) s += win[t] * in[k][r][cx + t];
    mid[k][r][cx] = s;
  }
}

// Vertical 11-tap pass of map k at output (r, cx).
__device__ __forceinline__ float filter_col(const float (*mid)[kLInY][kLTX], int k, int r, int cx, const float* win) {
  float s = 0.f;
#pragma unroll
  for (int t = 0; t < kWin; ++t) s += win[t] * mid[k][r + t][cx];
  return s;
}

__device__ __forceinline__ void block_sum2(float& a0, float& a1, float (*s_red)[256 / kWave]) {
  for (int o = kWave / 2; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o, kWave);
    a1 += __shfl_xor(a1, o, kWave);
  }
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) {
    s_red[0][w] = a0;
    s_red[1][w] = a1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a0 = a1 = 0.f;
    for (int i = 0; i < 256 / kWave; ++i) {
      a0 += s_red[0][i];
      a1 += s_red[1][i];
    }
  }
}

__global__ __launch_bounds__(256) void ssim_stats_kernel(LossArgs a) {
  __shared__ float s_in[5][kLInY][kLInX];
  __shared__ float s_mid[5][kLInY][kLTX];
  __shared__ float s_red[2][256 / kWave];
  const int c = blockIdx.z;
  const int x0 = blockIdx.x * kLTX, y0 = blockIdx.y * kLTY;
  const float C1 = 0.01f * 0.01f, C2 = 0.03f * 0.03f;
  for (int i = threadIdx.x; i < kLInY * kLInX; i += blockDim.x) {
    const int r = i / kLInX, cx = i % kLInX;
    const float xv = fetch_x(a, c, y0 - kHalf + r, x0 - kHalf + cx);
    const float yv = fetch_y(a, c, y0 - kHalf + r, x0 - kHalf + cx);
    s_in[0][r][cx] = xv;
    s_in[1][r][cx] = yv;
    s_in[2][r][cx] = xv * xv;
    s_in[3][r][cx] = yv * yv;
    s_in[4][r][cx] = xv * yv;
  }
  __syncthreads();
  filter_rows<5>(s_in, s_mid, a.win);
  __syncthreads();
  const size_t plane = (size_t)a.Hc * a.W;
  float l1 = 0.f, msum = 0.f;
  const int cx = threadIdx.x % kLTX;
  for (int r = threadIdx.x / kLTX; r < kLTY; r += 256 / kLTX) {
    const int yy = y0 + r, xx = x0 + cx;
    if (yy >= a.Hc || xx >= a.W) continue;
    l1 += fabsf(s_in[0][r + kHalf][cx + kHalf] - s_in[1][r + kHalf][cx + kHalf]);
    const float mu1 = filter_col(s_mid, 0, r, cx, a.win), mu2 = filter_col(s_mid, 1, r, cx, a.win);
    const float s11 = filter_col(s_mid, 2, r, cx, a.win) - mu1 * mu1;
    const float s22 = filter_col(s_mid, 3, r, cx, a.win) - mu2 * mu2;
    const float s12 = filter_col(s_mid, 4, r, cx, a.win) - mu1 * mu2;
    const float A1 = 2.f * mu1 * mu2 + C1, A2 = 2.f * s12 + C2;
    const float B1 = mu1 * mu1 + mu2 * mu2 + C1, B2 = s11 + s22 + C2;
    const float inv = 1.f / (B1 * B2);
    const float m = A1 * A2 * inv;
    msum += m;
    const float dm_dmu1 = 2.f * A2 * (mu2 * B1 - mu1 * A1) * inv / B1;
    const float dm_ds12 = 2.f * A1 * inv;
    const float dm_ds11 = -m / B2;
    const size_t gi = (size_t)c * plane + (size_t)yy * a.W + xx;
    a.g[gi] = dm_dmu1 - 2.f * mu1 * dm_ds11 - mu2 * dm_ds12;
    a.g[(size_t)a.C * plane + gi] = dm_ds11;
    a.g[(size_t)2 * a.C * plane + gi] = dm_ds12;
  }
  block_sum2(l1, msum, s_red);
  if (threadIdx.x == 0) {
    const size_t b = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    a.partial[2 * b] = l1;
    a.partial[2 * b + 1] = msum;
  }
}

__global__ __launch_bounds__(256) void ssim_grad_kernel(LossArgs a) {
  __shared__ float s_in[3][kLInY][kLInX];
  __shared__ float s_mid[3][kLInY][kLTX];
  const int c = blockIdx.z;
  const int x0 = blockIdx.x * kLTX, y0 = blockIdx.y * kLTY;
  const size_t plane = (size_t)a.Hc * a.W;
  const float n = (float)((size_t)a.C * plane);
  if (y0 < a.Hc) {  // tiles wholly inside the dropped rows only write zeros
    for (int i = threadIdx.x; i < kLInY * kLInX; i += blockDim.x) {
      const int r = i / kLInX, cx = i % kLInX;
      const int yy = y0 - kHalf + r, xx = x0 - kHalf + cx;
      const bool in = yy >= 0 && yy < a.Hc && xx >= 0 && xx < a.W;
      const size_t gi = (size_t)c * plane + (size_t)(in ? yy : 0) * a.W + (in ? xx : 0);
#pragma unroll
      for (int k = 0; k < 3; ++k) s_in[k][r][cx] = in ? a.g[(size_t)k * a.C * plane + gi] : 0.f;
    }
    __syncthreads();
    filter_rows<3>(s_in, s_mid, a.win);
    __syncthreads();
  }
  const float l1w = (1.f - a.lambda) / n, sw = -a.lambda / n;
  const int cx = threadIdx.x % kLTX;
  for (int r = threadIdx.x / kLTX; r < kLTY; r += 256 / kLTX) {
    const int yy = y0 + r, xx = x0 + cx;
    if (yy >= a.H || xx >= a.W) continue;
    const size_t idx = ((size_t)c * a.H + yy) * a.W + xx;
    if (yy >= a.Hc) {  // dropped rows carry no gradient
      a.dL_dx[idx] = 0.f;
      continue;
    }
    const float xv = fetch_x(a, c, yy, xx), yv = a.y[idx];
    float gv = l1w * sgn(xv - yv);
    gv += sw * (filter_col(s_mid, 0, r, cx, a.win) + 2.f * xv * filter_col(s_mid, 1, r, cx, a.win) +
                yv * filter_col(s_mid, 2, r, cx, a.win));
    a.dL_dx[idx] = a.mask ? gv * a.mask[(size_t)yy * a.W + xx] : gv;
  }
}

// L1 alone: dL/dx = sign(x*mask - y) * mask / n over the kept rows, zero below.
// Elements are taken in groups of kV (float4 when the row length allows it,
// so a group never straddles a row or the crop boundary).
template <int kV>
__global__ __launch_bounds__(256) void l1_grad_kernel(LossArgs a, float inv_n) {
  using vec = typename std::conditional<kV == 4, float4, float>::type;
  __shared__ float s_red[2][256 / kWave];
  const size_t plane = (size_t)a.H * a.W, kept = (size_t)a.Hc * a.W;
  const size_t groups = (size_t)a.C * plane / kV;
  const size_t gplane = plane / kV, gkept = kept / kV;
  const vec* x = reinterpret_cast<const vec*>(a.x);
  const vec* y = reinterpret_cast<const vec*>(a.y);
  const vec* mk = reinterpret_cast<const vec*>(a.mask);
  vec* out = reinterpret_cast<vec*>(a.dL_dx);
  float l1 = 0.f, unused = 0.f;
  for (size_t gidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; gidx < groups;
       gidx += (size_t)gridDim.x * blockDim.x) {
    const size_t o = gidx % gplane;
    if (o >= gkept) {
      out[gidx] = vec{};
      continue;
    }
    vec xl = x[gidx], yl = y[gidx], r{};
    vec ml{};
    if (a.mask) ml = mk[o];
    float* xs = reinterpret_cast<float*>(&xl);
    const float* ys = reinterpret_cast<const float*>(&yl);
    const float* ms = reinterpret_cast<const float*>(&ml);
    float* rs = reinterpret_cast<float*>(&r);
#pragma unroll
    for (int k = 0; k < kV; ++k) {
      const float xm = a.mask ? xs[k] * ms[k] : xs[k];
      const float d = xm - ys[k];
      l1 += fabsf(d);
      rs[k] = a.mask ? inv_n * sgn(d) * ms[k] : inv_n * sgn(d);
    }
    out[gidx] = r;
  }
  block_sum2(l1, unused, s_red);
  if (threadIdx.x == 0) {
    a.partial[2 * blockIdx.x] = l1;
    a.partial[2 * blockIdx.x + 1] = 0.f;
  }
}

// Sums nb (l1, m) pairs in a fixed order and writes {loss, L1, SSIM}.
__global__ __launch_bounds__(kFinalThreads) void loss_final_kernel(const float2* __restrict__ partial, int nb, float n,
                                                                   float lambda, float* __restrict__ out) {
  __shared__ float s[2][kFinalThreads / kWave];
  float a0 = 0.f, a1 = 0.f;
  int i = threadIdx.x;
  for (; i + 3 * kFinalThreads < nb; i += 4 * kFinalThreads) {  // four independent loads in flight
    const float2 p0 = partial[i], p1 = partial[i + kFinalThreads];
    const float2 p2 = partial[i + 2 * kFinalThreads], p3 = partial[i + 3 * kFinalThreads];
    a0 += (p0.x + p1.x) + (p2.x + p3.x);
    a1 += (p0.y + p1.y) + (p2.y + p3.y);
  }
  for (; i < nb; i += kFinalThreads) {
    a0 += partial[i].x;
    a1 += partial[i].y;
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o, kWave);
    a1 += __shfl_xor(a1, o, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s[0][threadIdx.x / kWave] = a0;
    s[1][threadIdx.x / kWave] = a1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float t0 = 0.f, t1 = 0.f;
    for (int w = 0; w < kFinalThreads / kWave; ++w) {
      t0 += s[0][w];
      t1 += s[1][w];
    }
    const float l1 = t0 / n;
    const float ssim = lambda != 0.f ? t1 / n : 0.f;
    out[0] = lambda != 0.f ? (1.f - lambda) * l1 + lambda * (1.f - ssim) : l1;
    out[1] = l1;
    out[2] = ssim;
  }
}

// The three SSIM gradient maps, rounded up to 64 floats so that the block
// partials placed after them (read as float2) stay 256-byte aligned.
static size_t ssim_map_floats(int Hc, int W, int C) { return ((size_t)3 * C * Hc * W + 63) & ~(size_t)63; }

static size_t ssim_blocks(int C, int Hc, int W) {
  return (size_t)((W + kLTX - 1) / kLTX) * ((Hc + kLTY - 1) / kLTY) * C;
}

size_t loss_workspace_floats(int C, int H, int W, int skip) {
  const int Hc = H - skip;
  if (Hc <= 0) return 0;
  const size_t nb = std::max(ssim_blocks(C, Hc, W), (size_t)kL1Blocks);
  return ssim_map_floats(Hc, W, C) + 2 * nb + 64;
}

hipError_t launch_loss(const float* x, const float* y, const float* mask, int C, int H, int W, int skip, float lambda,
                       float* workspace, float* dL_dx, float* out, hipStream_t s) {
  LossArgs a{};
  a.x = x;
  a.y = y;
  a.mask = mask;
  a.C = C;
  a.H = H;
  a.W = W;
  a.Hc = H - skip;
  a.lambda = lambda;
  // 1D Gaussian window of the reference (loss_utils.h:53-67): exp(-t^2 / (2 sigma^2)), t = -5..5, normalised
  float sum = 0.f;
  for (int k = 0; k < kWin; ++k) {
    const int t = k - kHalf;
    a.win[k] = std::exp(-(float)(t * t) / (2.0f * 1.5f * 1.5f));
    sum += a.win[k];
  }
  for (int k = 0; k < kWin; ++k) a.win[k] /= sum;
  a.g = workspace;
  a.partial = workspace + ssim_map_floats(a.Hc, W, C);
  a.dL_dx = dL_dx;
  const float n = (float)((size_t)C * a.Hc * W);
  const float2* partial = reinterpret_cast<const float2*>(a.partial);
  if (lambda != 0.f) {
    const dim3 grid((W + kLTX - 1) / kLTX, (a.Hc + kLTY - 1) / kLTY, C);
    const dim3 ggrid((W + kLTX - 1) / kLTX, (H + kLTY - 1) / kLTY, C);  // every row: dropped rows get zeros
    hipLaunchKernelGGL(ssim_stats_kernel, grid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(ssim_grad_kernel, ggrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(loss_final_kernel, dim3(1), dim3(kFinalThreads), 0, s, partial,
                       (int)ssim_blocks(C, a.Hc, W), n, lambda, out);
  } else {
    const bool vec4 = (W % 4) == 0 && (reinterpret_cast<uintptr_t>(x) % 16) == 0 &&
                      (reinterpret_cast<uintptr_t>(y) % 16) == 0 && (reinterpret_cast<uintptr_t>(dL_dx) % 16) == 0 &&
                      (reinterpret_cast<uintptr_t>(mask) % 16) == 0;
    const size_t groups = (size_t)C * H * W / (vec4 ? 4 : 1);
    const int blocks = (int)std::min<size_t>((groups + 255) / 256, (size_t)kL1Blocks);
    if (vec4)
      hipLaunchKernelGGL(l1_grad_kernel<4>, dim3(blocks), dim3(256), 0, s, a, 1.f / n);
    else
      hipLaunchKernelGGL(l1_grad_kernel<1>, dim3(blocks), dim3(256), 0, s, a, 1.f / n);
    hipLaunchKernelGGL(loss_final_kernel, dim3(1), dim3(kFinalThreads), 0, s, partial, blocks, n, lambda, out);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free form: the workspace is this file-scope device array instead
// of caller memory.  Every GPU the library runs on holds its own copy (80 MiB)
// until the process ends; a call writes every workspace word it later reads
// (the SSIM maps and the block partials), so nothing needs clearing, and
// calls that overlap on one GPU race on it.  In host code `scratch` names a
// host-side stand-in, so its device address is looked up once per device
// with hipGetSymbolAddress and only that address is passed to the kernels.
static __device__ __align__(256) float scratch[kLossScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t loss_scratch(float** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace omnigs
