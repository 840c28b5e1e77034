// torch::autograd::Function around the equirectangular rasterizer.
//
// GaussianRasterizationSettings / GaussianRasterizerFunction /
// GaussianRasterizer follow the reference's include/gaussian_rasterizer.h:32-141
// (argument order, the "exactly one of" checks, the gradient order
// {dmeans3D, dmeans2D, dsh, dcolors, dopacity, dscales, drot, dcov3D, -}).
// GaussianRasterizerCheckpointFunction is this build's memory-lean form: its
// context keeps only the image buffer (per-pixel final T, n_contrib and tile
// ranges) and radii, and the backward recomputes preprocess and binning.
#pragma once

#include <torch/torch.h>

#include <tuple>

#include "rasterize_points.h"

namespace omnigs_torch {

enum CameraModelType { INVALID = 0, PINHOLE = 1, FISHEYE = 2, LONLAT = 3 };

struct GaussianRasterizationSettings {
  GaussianRasterizationSettings(int image_height, int image_width, float tanfovx, float tanfovy, torch::Tensor bg,
                                float scale_modifier, torch::Tensor viewmatrix, torch::Tensor projmatrix,
                                torch::Tensor RwcT, int sh_degree, torch::Tensor campos, bool prefiltered,
                                CameraModelType camera_type = LONLAT, bool render_depth = false,
                                bool seam_wrap = false)
      : image_height_(image_height), image_width_(image_width), tanfovx_(tanfovx), tanfovy_(tanfovy), bg_(bg),
        scale_modifier_(scale_modifier), viewmatrix_(viewmatrix), projmatrix_(projmatrix), RwcT_(RwcT),
        sh_degree_(sh_degree), campos_(campos), prefiltered_(prefiltered), camera_type_(camera_type),
        render_depth_(render_depth), seam_wrap_(seam_wrap) {}

  int image_height_;
  int image_width_;
  float tanfovx_;
  float tanfovy_;
  torch::Tensor bg_;
  float scale_modifier_;
  torch::Tensor viewmatrix_;
  torch::Tensor projmatrix_;
  torch::Tensor RwcT_;
  int sh_degree_;
  torch::Tensor campos_;
  bool prefiltered_;
  CameraModelType camera_type_;
  bool render_depth_;
  bool seam_wrap_;  // extension: wrap tile columns around the longitude seam
};

class GaussianRasterizerFunction : public torch::autograd::Function<GaussianRasterizerFunction> {
 public:
  static torch::autograd::tensor_list forward(torch::autograd::AutogradContext* ctx, torch::Tensor means3D,
                                              torch::Tensor means2D, torch::Tensor sh, torch::Tensor colors_precomp,
                                              torch::Tensor opacities, torch::Tensor scales, torch::Tensor rotations,
                                              torch::Tensor cov3Ds_precomp,
                                              GaussianRasterizationSettings raster_settings);
  static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* ctx,
                                               torch::autograd::tensor_list grad_out_color);
};

class GaussianRasterizerCheckpointFunction
    : public torch::autograd::Function<GaussianRasterizerCheckpointFunction> {
 public:
  static torch::autograd::tensor_list forward(torch::autograd::AutogradContext* ctx, torch::Tensor means3D,
                                              torch::Tensor means2D, torch::Tensor sh, torch::Tensor colors_precomp,
                                              torch::Tensor opacities, torch::Tensor scales, torch::Tensor rotations,
                                              torch::Tensor cov3Ds_precomp,
                                              GaussianRasterizationSettings raster_settings);
  static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* ctx,
                                               torch::autograd::tensor_list grad_out_color);
};

inline torch::autograd::tensor_list rasterizeGaussians(torch::Tensor& means3D, torch::Tensor& means2D,
                                                       torch::Tensor& sh, torch::Tensor& colors_precomp,
                                                       torch::Tensor& opacities, torch::Tensor& scales,
                                                       torch::Tensor& rotations, torch::Tensor& cov3Ds_precomp,
                                                       GaussianRasterizationSettings& raster_settings,
                                                       bool checkpoint = false) {
  if (checkpoint)
    return GaussianRasterizerCheckpointFunction::apply(means3D, means2D, sh, colors_precomp, opacities, scales,
                                                       rotations, cov3Ds_precomp, raster_settings);
  return GaussianRasterizerFunction::apply(means3D, means2D, sh, colors_precomp, opacities, scales, rotations,
                                           cov3Ds_precomp, raster_settings);
}

class GaussianRasterizer : public torch::nn::Module {
 public:
  explicit GaussianRasterizer(GaussianRasterizationSettings& raster_settings, bool checkpoint = false)
      : raster_settings_(raster_settings), checkpoint_(checkpoint) {}

  torch::Tensor markVisibleGaussians(torch::Tensor& positions);

  std::tuple<torch::Tensor, torch::Tensor> forward(torch::Tensor means3D, torch::Tensor means2D,
                                                   torch::Tensor opacities, bool has_shs, bool has_colors_precomp,
                                                   bool has_scales, bool has_rotations, bool has_cov3D_precomp,
                                                   torch::Tensor shs, torch::Tensor colors_precomp,
                                                   torch::Tensor scales, torch::Tensor rotations,
                                                   torch::Tensor cov3D_precomp);

  GaussianRasterizationSettings raster_settings_;
  bool checkpoint_;
};

}  // namespace omnigs_torch
