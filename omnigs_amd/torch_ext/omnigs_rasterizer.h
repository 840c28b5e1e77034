// Extensions of the LibTorch boundary that the reference does not have, kept
// out of its declarations (rasterize_points.h, gaussian_rasterizer.h) so that
// those stay exactly the reference's:
//
// * seam wrap: tile columns wrap around the longitude seam instead of being
//   clamped at the image edge (the reference clamps with getRect,
//   cuda_rasterizer/forward.cu:681; its cyclic getRectCyclic,
//   auxiliary.h:68-83, is unused).  The image width must then be a multiple of
//   the 16-pixel tile width.
// * checkpointing: the autograd context keeps only the image buffer (per-pixel
//   final T and contributor counts, per-tile ranges) and the radii; the
//   backward re-runs preprocess and the depth sort, then bins and runs the
//   render backward in bands of tile rows, so that the per-instance buffers
//   exist for one band at a time (the reference keeps all three buffers from
//   the forward to the backward, src/gaussian_rasterizer.cpp:85-98).
#pragma once

#include <torch/torch.h>

#include <tuple>

#include "gaussian_rasterizer.h"
#include "rasterize_points.h"

namespace omnigs {

// RasterizeGaussiansCUDA / RasterizeGaussiansBackwardCUDA with the seam-wrap
// switch; seam_wrap = false is the reference's behaviour (and what the global
// functions run).
std::tuple<int, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor> RasterizeGaussiansCUDA(
    const torch::Tensor& background, const torch::Tensor& means3D, const torch::Tensor& colors,
    const torch::Tensor& opacity, const torch::Tensor& scales, const torch::Tensor& rotations,
    const float scale_modifier, const torch::Tensor& cov3D_precomp, const torch::Tensor& viewmatrix,
    const torch::Tensor& projmatrix, const float tan_fovx, const float tan_fovy, const int image_height,
    const int image_width, const torch::Tensor& sh, const int degree, const torch::Tensor& campos,
    const bool prefiltered, const int camera_type, const bool render_depth, const bool seam_wrap);

std::tuple<torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor,
           torch::Tensor>
RasterizeGaussiansBackwardCUDA(const torch::Tensor& background, const torch::Tensor& means3D,
                               const torch::Tensor& radii, const torch::Tensor& colors, const torch::Tensor& scales,
                               const torch::Tensor& rotations, const float scale_modifier,
                               const torch::Tensor& cov3D_precomp, const torch::Tensor& viewmatrix,
                               const torch::Tensor& projmatrix, const float tan_fovx, const float tan_fovy,
                               const torch::Tensor& dL_dout_color, const torch::Tensor& sh, const int degree,
                               const torch::Tensor& campos, const torch::Tensor& geomBuffer, const int R,
                               const torch::Tensor& binningBuffer, const torch::Tensor& imageBuffer,
                               const int camera_type, const bool seam_wrap);

// The checkpointed form's forward: RasterizeGaussiansCUDA's outputs, but the
// binning buffer holds no gradient records (17 of 53 bytes per instance);
// its buffers serve RasterizeGaussiansBackwardCheckpointedCUDA only, which
// reads the image buffer alone.
std::tuple<int, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor>
RasterizeGaussiansCheckpointedCUDA(const torch::Tensor& background, const torch::Tensor& means3D,
                                   const torch::Tensor& colors, const torch::Tensor& opacity,
                                   const torch::Tensor& scales, const torch::Tensor& rotations,
                                   const float scale_modifier, const torch::Tensor& cov3D_precomp,
                                   const torch::Tensor& viewmatrix, const int image_height, const int image_width,
                                   const torch::Tensor& sh, const int degree, const torch::Tensor& campos,
                                   const bool prefiltered, const int camera_type, const bool seam_wrap);

// The checkpointed backward: the forward's inputs (as given to
// RasterizeGaussiansCUDA), its R and image buffer, and the band budget in
// instances (0: the library's default); returns the eight gradients of
// RasterizeGaussiansBackwardCUDA.
std::tuple<torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor,
           torch::Tensor>
RasterizeGaussiansBackwardCheckpointedCUDA(const torch::Tensor& background, const torch::Tensor& means3D,
                                           const torch::Tensor& radii, const torch::Tensor& colors,
                                           const torch::Tensor& opacity, const torch::Tensor& scales,
                                           const torch::Tensor& rotations, const float scale_modifier,
                                           const torch::Tensor& cov3D_precomp, const torch::Tensor& viewmatrix,
                                           const torch::Tensor& dL_dout_color, const torch::Tensor& sh,
                                           const int degree, const torch::Tensor& campos, const bool prefiltered,
                                           const int R, const torch::Tensor& imageBuffer, const int camera_type,
                                           const bool seam_wrap, const int64_t band_instances);

struct RasterizerOptions {
  bool checkpoint = false;     // keep only the image buffer; recompute in the backward, in bands of tile rows
  bool seam_wrap = false;      // wrap tile columns around the longitude seam
  int64_t band_instances = 0;  // checkpointed backward: instances per band (0: the library's default)
};

// Stored form with the seam-wrap switch (GaussianRasterizerFunction is this
// with seam_wrap = false).  backward returns the reference's nine gradients
// plus an undefined one for the switch.
class StoredRasterizerFunction : public torch::autograd::Function<StoredRasterizerFunction> {
 public:
  static torch::autograd::tensor_list forward(torch::autograd::AutogradContext* ctx, torch::Tensor means3D,
                                              torch::Tensor means2D, torch::Tensor sh, torch::Tensor colors_precomp,
                                              torch::Tensor opacities, torch::Tensor scales, torch::Tensor rotations,
                                              torch::Tensor cov3Ds_precomp,
                                              GaussianRasterizationSettings raster_settings, bool seam_wrap);
  static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* ctx,
                                               torch::autograd::tensor_list grad_out_color);
};

// Checkpointed form: same inputs, outputs and gradients as the stored form
// (backward: plus undefined ones for the two switches).
class CheckpointRasterizerFunction : public torch::autograd::Function<CheckpointRasterizerFunction> {
 public:
  static torch::autograd::tensor_list forward(torch::autograd::AutogradContext* ctx, torch::Tensor means3D,
                                              torch::Tensor means2D, torch::Tensor sh, torch::Tensor colors_precomp,
                                              torch::Tensor opacities, torch::Tensor scales, torch::Tensor rotations,
                                              torch::Tensor cov3Ds_precomp,
                                              GaussianRasterizationSettings raster_settings, bool seam_wrap,
                                              int64_t band_instances);
  static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* ctx,
                                               torch::autograd::tensor_list grad_out_color);
};

inline torch::autograd::tensor_list rasterizeGaussians(torch::Tensor& means3D, torch::Tensor& means2D,
                                                       torch::Tensor& sh, torch::Tensor& colors_precomp,
                                                       torch::Tensor& opacities, torch::Tensor& scales,
                                                       torch::Tensor& rotations, torch::Tensor& cov3Ds_precomp,
                                                       GaussianRasterizationSettings& raster_settings,
                                                       const RasterizerOptions& options) {
  if (options.checkpoint)
    return CheckpointRasterizerFunction::apply(means3D, means2D, sh, colors_precomp, opacities, scales, rotations,
                                               cov3Ds_precomp, raster_settings, options.seam_wrap,
                                               options.band_instances);
  return StoredRasterizerFunction::apply(means3D, means2D, sh, colors_precomp, opacities, scales, rotations,
                                         cov3Ds_precomp, raster_settings, options.seam_wrap);
}

// GaussianRasterizer with options (same forward and markVisibleGaussians).
class LonlatRasterizer : public torch::nn::Module {
 public:
  LonlatRasterizer(GaussianRasterizationSettings& raster_settings, RasterizerOptions options = {})
      : raster_settings_(raster_settings), options_(options) {}

  torch::Tensor markVisibleGaussians(torch::Tensor& positions);

  std::tuple<torch::Tensor, torch::Tensor> forward(torch::Tensor means3D, torch::Tensor means2D,
                                                   torch::Tensor opacities, bool has_shs, bool has_colors_precomp,
                                                   bool has_scales, bool has_rotations, bool has_cov3D_precomp,
                                                   torch::Tensor shs, torch::Tensor colors_precomp,
                                                   torch::Tensor scales, torch::Tensor rotations,
                                                   torch::Tensor cov3D_precomp);

  GaussianRasterizationSettings raster_settings_;
  RasterizerOptions options_;
};

}  // namespace omnigs
