// Per-iteration training-state updates around the rasterizer.
//
// densification statistics: one pass doing what the reference's training
// step does with three torch index ops (gaussian_mapper.cpp:419-428 and
// GaussianModel::addDensificationStats, gaussian_model.cpp:839-853): for
// every Gaussian with radius > 0 (the visibility filter of renderLonlat),
//   max_radii2D = max(max_radii2D, radius)
//   xyz_gradient_accum += |dL/dmean2D.xy|   (frobenius norm of the first two
//                                             components of the screen-space
//                                             points' gradient)
//   denom += 1
// HBM-bound: 32 B read and 12 B written per visible Gaussian.
#include <cmath>

#include "kernels.hpp"

namespace omnigs {

__global__ __launch_bounds__(256) void densify_stats_kernel(int P, const int* __restrict__ radii,
                                                            const float* __restrict__ dmean2D,
                                                            float* __restrict__ max_radii2D,
                                                            float* __restrict__ accum, float* __restrict__ denom) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const int r = radii[i];
  if (r <= 0) return;
  max_radii2D[i] = fmaxf(max_radii2D[i], (float)r);
  const float gx = dmean2D[3 * i], gy = dmean2D[3 * i + 1];
  accum[i] += sqrtf(gx * gx + gy * gy);
  denom[i] += 1.f;
}

hipError_t launch_densify_stats(int P, const int* radii, const float* dmean2D, float* max_radii2D, float* accum,
                                float* denom, hipStream_t s) {
  if (P == 0) return hipSuccess;
  hipLaunchKernelGGL(densify_stats_kernel, dim3((P + 255) / 256), dim3(256), 0, s, P, radii, dmean2D, max_radii2D,
                     accum, denom);
  return hipGetLastError();
}

// exponLrFunc (gaussian_model.cpp:1140-1153), in float as the reference
// evaluates it: log-linear from lr_init (step 0) to lr_final (max_steps),
// scaled during the first lr_delay_steps by a sine ramp from lr_delay_mult.
float expon_lr(int step, float lr_init, float lr_final, int lr_delay_steps, float lr_delay_mult, int max_steps) {
  if (step < 0 || (lr_init == 0.0f && lr_final == 0.0f)) return 0.0f;
  float delay_rate;
  if (lr_delay_steps > 0) {
    const float x = std::fmin(std::fmax(static_cast<float>(step) / lr_delay_steps, 0.0f), 1.0f);
    delay_rate = lr_delay_mult + (1.0f - lr_delay_mult) * std::sin(1.57079632679489661923f * x);
  } else {
    delay_rate = 1.0f;
  }
  const float t = std::fmin(std::fmax(static_cast<float>(step) / max_steps, 0.0f), 1.0f);
  const float log_lerp = std::exp(std::log(lr_init) * (1 - t) + std::log(lr_final) * t);
  return delay_rate * log_lerp;
}

}  // namespace omnigs
