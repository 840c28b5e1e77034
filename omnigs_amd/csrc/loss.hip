// Fused photometric loss and its image gradient:
//   loss = (1 - lambda) mean|x - y| + lambda (1 - mean SSIM(x, y))
// over the image with its bottom `skip` rows dropped, as the reference's
// trainer computes it with torch ops (gaussian_mapper.cpp:384-407,
// loss_utils.h:31-129: 11x11 Gaussian window with sigma 1.5, zero padding,
// C1 = 0.01^2, C2 = 0.03^2, means over every element).
//
// With E1 = w*x, E2 = w*y, E11 = w*x^2, E22 = w*y^2, E12 = w*xy (separable
// 11-tap filtering, zero padded) the SSIM map is
//   m = (2 mu1 mu2 + C1)(2 s12 + C2) / ((mu1^2 + mu2^2 + C1)(s11 + s22 + C2)),
// mu = E, s11 = E11 - mu1^2, s12 = E12 - mu1 mu2, and
//   dSSIM/dx = (1/n) [w*g1 + 2 x (w*g11) + y (w*g12)]
// with g1 = dm/dmu1 - 2 mu1 dm/ds11 - mu2 dm/ds12, g11 = dm/ds11,
// g12 = dm/ds12 (the window is symmetric, so its transpose is itself).
// Two passes: stats (writes g1, g11, g12 and per-block partial sums) and
// gradient; a one-block kernel reduces the partials deterministically.  With
// lambda = 0 only the gradient pass runs and it writes the L1 partial sums.
#include "kernels.hpp"

namespace omnigs {

constexpr int kWin = 11, kHalf = 5;
constexpr int kLTX = 32, kLTY = 8;  // output tile per 256-thread block
constexpr int kLInX = kLTX + 2 * kHalf, kLInY = kLTY + 2 * kHalf;

struct LossArgs {
  const float* x;     // rendered image [C,H,W]
  const float* y;     // target image [C,H,W]
  const float* mask;  // optional [H,W] multiplier of x (undistortion mask), NULL = ones
  int C, H, W, Hc;  // Hc = rows kept (H - skip)
  float lambda;
  float win[kWin];
  float* g;         // [3, C, Hc, W] SSIM adjoint maps (g1, g11, g12)
  float* partial;   // [blocks, 2] (sum |x - y|, sum m)
  float* dL_dx;     // [C,H,W]
  float* out;       // [3]: loss, L1, SSIM
};

__device__ __forceinline__ float fetch(const float* __restrict__ p, int c, int yy, int xx, int Hc, int W, int H) {
  return (yy >= 0 && yy < Hc && xx >= 0 && xx < W) ? p[((size_t)c * H + yy) * W + xx] : 0.f;
}

// masked rendered value x * mask (zero outside the kept rows)
__device__ __forceinline__ float fetch_x(const LossArgs& a, int c, int yy, int xx) {
  if (!(yy >= 0 && yy < a.Hc && xx >= 0 && xx < a.W)) return 0.f;
  const float v = a.x[((size_t)c * a.H + yy) * a.W + xx];
  return a.mask ? v * a.mask[(size_t)yy * a.W + xx] : v;
}

template <int K>
__device__ __forceinline__ void filter_tile(const float (*in)[kLInY][kLInX], float (*mid)[kLInY][kLTX],
                                            float (*out)[kLTY][kLTX], const float* win) {
  // horizontal then vertical 11-tap pass over K maps held in LDS
  for (int i = threadIdx.x; i < K * kLInY * kLTX; i += blockDim.x) {
    const int k = i / (kLInY * kLTX), r = (i / kLTX) % kLInY, cx = i % kLTX;
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < kWin; ++t) s += win[t] * in[k][r][cx + t];
    mid[k][r][cx] = s;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < K * kLTY * kLTX; i += blockDim.x) {
    const int k = i / (kLTY * kLTX), r = (i / kLTX) % kLTY, cx = i % kLTX;
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < kWin; ++t) s += win[t] * mid[k][r + t][cx];
    out[k][r][cx] = s;
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void loss_stats_kernel(LossArgs a) {
  __shared__ float s_in[5][kLInY][kLInX];
  __shared__ float s_mid[5][kLInY][kLTX];
  __shared__ float s_out[5][kLTY][kLTX];
  __shared__ float s_red[2][256 / kWave];
  const int c = blockIdx.z;
  const int x0 = blockIdx.x * kLTX, y0 = blockIdx.y * kLTY;
  const bool want_ssim = a.lambda != 0.f;
  const float C1 = 0.01f * 0.01f, C2 = 0.03f * 0.03f;
  float l1 = 0.f, msum = 0.f;
  if (want_ssim) {
    for (int i = threadIdx.x; i < kLInY * kLInX; i += blockDim.x) {
      const int r = i / kLInX, cx = i % kLInX;
      const float xv = fetch_x(a, c, y0 - kHalf + r, x0 - kHalf + cx);
      const float yv = fetch(a.y, c, y0 - kHalf + r, x0 - kHalf + cx, a.Hc, a.W, a.H);
      s_in[0][r][cx] = xv;
      s_in[1][r][cx] = yv;
      s_in[2][r][cx] = xv * xv;
      s_in[3][r][cx] = yv * yv;
      s_in[4][r][cx] = xv * yv;
    }
    __syncthreads();
    filter_tile<5>(s_in, s_mid, s_out, a.win);
  }
  const size_t plane = (size_t)a.Hc * a.W;
  for (int i = threadIdx.x; i < kLTY * kLTX; i += blockDim.x) {
    const int r = i / kLTX, cx = i % kLTX;
    const int yy = y0 + r, xx = x0 + cx;
    if (yy >= a.Hc || xx >= a.W) continue;
    const size_t idx = ((size_t)c * a.H + yy) * a.W + xx;
    l1 += fabsf(fetch_x(a, c, yy, xx) - a.y[idx]);
    if (want_ssim) {
      const float mu1 = s_out[0][r][cx], mu2 = s_out[1][r][cx];
      const float s11 = s_out[2][r][cx] - mu1 * mu1, s22 = s_out[3][r][cx] - mu2 * mu2;
      const float s12 = s_out[4][r][cx] - mu1 * mu2;
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
  }
  // block sums of |x - y| and m
  for (int o = kWave / 2; o > 0; o >>= 1) {
    l1 += __shfl_xor(l1, o, kWave);
    msum += __shfl_xor(msum, o, kWave);
  }
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  if (lane == 0) {
    s_red[0][w] = l1;
    s_red[1][w] = msum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float a0 = 0.f, a1 = 0.f;
    for (int i = 0; i < 256 / kWave; ++i) {
      a0 += s_red[0][i];
      a1 += s_red[1][i];
    }
    const size_t b = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    a.partial[2 * b] = a0;
    a.partial[2 * b + 1] = a1;
  }
}

__global__ __launch_bounds__(256) void loss_grad_kernel(LossArgs a) {
  __shared__ float s_in[3][kLInY][kLInX];
  __shared__ float s_mid[3][kLInY][kLTX];
  __shared__ float s_out[3][kLTY][kLTX];
  const int c = blockIdx.z;
  const int x0 = blockIdx.x * kLTX, y0 = blockIdx.y * kLTY;
  const bool want_ssim = a.lambda != 0.f;
  const size_t plane = (size_t)a.Hc * a.W;
  const float n = (float)((size_t)a.C * plane);
  if (want_ssim) {
    for (int i = threadIdx.x; i < kLInY * kLInX; i += blockDim.x) {
      const int r = i / kLInX, cx = i % kLInX;
      const int yy = y0 - kHalf + r, xx = x0 - kHalf + cx;
      const bool in = yy >= 0 && yy < a.Hc && xx >= 0 && xx < a.W;
      const size_t gi = (size_t)c * plane + (size_t)(in ? yy : 0) * a.W + (in ? xx : 0);
#pragma unroll
      for (int k = 0; k < 3; ++k) s_in[k][r][cx] = in ? a.g[(size_t)k * a.C * plane + gi] : 0.f;
    }
    __syncthreads();
    filter_tile<3>(s_in, s_mid, s_out, a.win);
  }
  const float l1w = (1.f - a.lambda) / n, sw = -a.lambda / n;
  float l1 = 0.f;
  for (int i = threadIdx.x; i < kLTY * kLTX; i += blockDim.x) {
    const int r = i / kLTX, cx = i % kLTX;
    const int yy = y0 + r, xx = x0 + cx;
    if (yy >= a.H || xx >= a.W) continue;
    const size_t idx = ((size_t)c * a.H + yy) * a.W + xx;
    if (yy >= a.Hc) {  // cropped rows carry no gradient
      a.dL_dx[idx] = 0.f;
      continue;
    }
    const float xv = fetch_x(a, c, yy, xx), yv = a.y[idx];
    const float d = xv - yv;
    l1 += fabsf(d);
    float gv = l1w * (d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f));
    if (want_ssim) gv += sw * (s_out[0][r][cx] + 2.f * xv * s_out[1][r][cx] + yv * s_out[2][r][cx]);
    a.dL_dx[idx] = a.mask ? gv * a.mask[(size_t)yy * a.W + xx] : gv;
  }
  if (want_ssim) return;  // the stats pass already wrote the partial sums
  // L1 alone: this pass is the only one, so it writes the block's |x - y| sum
  __shared__ float s_red[256 / kWave];
  for (int o = kWave / 2; o > 0; o >>= 1) l1 += __shfl_xor(l1, o, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) s_red[threadIdx.x / kWave] = l1;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int i = 0; i < 256 / kWave; ++i) t += s_red[i];
    const size_t b = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    a.partial[2 * b] = t;
    a.partial[2 * b + 1] = 0.f;
  }
}

__global__ __launch_bounds__(256) void loss_final_kernel(const float* __restrict__ partial, size_t nb, float n,
                                                         float lambda, float* __restrict__ out) {
  __shared__ float s[2][256];
  float a0 = 0.f, a1 = 0.f;
  for (size_t i = threadIdx.x; i < nb; i += 256) {
    a0 += partial[2 * i];
    a1 += partial[2 * i + 1];
  }
  s[0][threadIdx.x] = a0;
  s[1][threadIdx.x] = a1;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s[0][threadIdx.x] += s[0][threadIdx.x + o];
      s[1][threadIdx.x] += s[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float l1 = s[0][0] / n;
    const float ssim = lambda != 0.f ? s[1][0] / n : 0.f;
    out[0] = (1.f - lambda) * l1 + (lambda != 0.f ? lambda * (1.f - ssim) : 0.f);
    out[1] = l1;
    out[2] = ssim;
  }
}

static size_t ssim_map_floats(int Hc, int W, int C) { return (size_t)3 * C * Hc * W; }

size_t loss_workspace_floats(int C, int H, int W, int skip) {
  const int Hc = H - skip;
  if (Hc <= 0) return 0;
  // partial sums for the larger of the two grids (stats over Hc rows, gradient over H rows)
  const size_t nb = (size_t)((W + kLTX - 1) / kLTX) * ((H + kLTY - 1) / kLTY) * C;
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
  const dim3 grid((W + kLTX - 1) / kLTX, (a.Hc + kLTY - 1) / kLTY, C);
  // gradient over every row (the cropped rows get zeros)
  const dim3 ggrid((W + kLTX - 1) / kLTX, (H + kLTY - 1) / kLTY, C);
  a.g = workspace;
  a.partial = workspace + ssim_map_floats(a.Hc, W, C);
  a.dL_dx = dL_dx;
  a.out = out;
  const float n = (float)((size_t)C * a.Hc * W);
  if (lambda != 0.f) {
    hipLaunchKernelGGL(loss_stats_kernel, grid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(loss_grad_kernel, ggrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(loss_final_kernel, dim3(1), dim3(256), 0, s, (const float*)a.partial,
                       (size_t)grid.x * grid.y * grid.z, n, lambda, out);
  } else {
    hipLaunchKernelGGL(loss_grad_kernel, ggrid, dim3(256), 0, s, a);
    hipLaunchKernelGGL(loss_final_kernel, dim3(1), dim3(256), 0, s, (const float*)a.partial,
                       (size_t)ggrid.x * ggrid.y * ggrid.z, n, lambda, out);
  }
  return hipGetLastError();
}

}  // namespace omnigs
