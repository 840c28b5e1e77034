// Autograd Functions over the LibTorch boundary.
//
// GaussianRasterizerFunction / GaussianRasterizer are the reference's classes
// (src/gaussian_rasterizer.cpp:24-224): the same saved tensors in the same
// order, the same saved scalars, the same gradient order and the same
// "exactly one of" checks (their messages kept verbatim, typo included).
// omnigs::StoredRasterizerFunction is the same stored form with the seam-wrap
// switch; omnigs::CheckpointRasterizerFunction keeps only the image buffer
// and re-runs preprocess and binning in the backward, in bands of tile rows.
#include "gaussian_rasterizer.h"

#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>

#include <stdexcept>
#include <string>

#include "omnigs_lonlat.h"
#include "omnigs_rasterizer.h"

namespace {

using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

constexpr int kLonlat = 3;  // Camera::CameraModelType::LONLAT

tensor_list grads_in_reference_order(
    const std::tuple<torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor,
                     torch::Tensor, torch::Tensor>& r) {
  return {std::get<3>(r) /*dL_dmeans3D*/, std::get<0>(r) /*dL_dmeans2D*/,  std::get<5>(r) /*dL_dsh*/,
          std::get<1>(r) /*dL_dcolors*/,  std::get<2>(r) /*dL_dopacity*/,  std::get<6>(r) /*dL_dscales*/,
          std::get<7>(r) /*dL_drotations*/, std::get<4>(r) /*dL_dcov3D*/, torch::Tensor()};
}

void save_settings(AutogradContext* ctx, const GaussianRasterizationSettings& rs, int R, bool seam_wrap) {
  ctx->saved_data["num_rendered"] = R;
  ctx->saved_data["scale_modifier"] = rs.scale_modifier_;
  ctx->saved_data["tanfovx"] = rs.tanfovx_;
  ctx->saved_data["tanfovy"] = rs.tanfovy_;
  ctx->saved_data["sh_degree"] = rs.sh_degree_;
  ctx->saved_data["camera_type"] = (int)rs.camera_type_;
  ctx->saved_data["image_height"] = rs.image_height_;
  ctx->saved_data["image_width"] = rs.image_width_;
  ctx->saved_data["prefiltered"] = rs.prefiltered_;
  ctx->saved_data["seam_wrap"] = seam_wrap;
}

char* resize_tensor(void* ctx, size_t nbytes) {
  try {
    auto* t = static_cast<torch::Tensor*>(ctx);
    t->resize_({(int64_t)nbytes});
    return reinterpret_cast<char*>(t->data_ptr());
  } catch (...) {
    return nullptr;
  }
}

const float* fptr(const torch::Tensor& t) { return t.numel() ? t.data_ptr<float>() : nullptr; }

std::tuple<int, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor> run_forward(
    const GaussianRasterizationSettings& rs, const torch::Tensor& means3D, const torch::Tensor& sh,
    const torch::Tensor& colors_precomp, const torch::Tensor& opacities, const torch::Tensor& scales,
    const torch::Tensor& rotations, const torch::Tensor& cov3Ds_precomp, bool seam_wrap) {
  return omnigs::RasterizeGaussiansCUDA(rs.bg_, means3D, colors_precomp, opacities, scales, rotations,
                                        rs.scale_modifier_, cov3Ds_precomp, rs.viewmatrix_, rs.projmatrix_,
                                        rs.tanfovx_, rs.tanfovy_, rs.image_height_, rs.image_width_, sh,
                                        rs.sh_degree_, rs.campos_, rs.prefiltered_, (int)rs.camera_type_,
                                        rs.render_depth_, seam_wrap);
}

// The stored form (reference src/gaussian_rasterizer.cpp:34-101, :103-170).
tensor_list stored_forward(AutogradContext* ctx, const torch::Tensor& means3D, const torch::Tensor& sh,
                           const torch::Tensor& colors_precomp, const torch::Tensor& opacities,
                           const torch::Tensor& scales, const torch::Tensor& rotations,
                           const torch::Tensor& cov3Ds_precomp, const GaussianRasterizationSettings& rs,
                           bool seam_wrap) {
  auto res = run_forward(rs, means3D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp, seam_wrap);
  const int R = std::get<0>(res);
  auto color = std::get<1>(res);
  auto radii = std::get<2>(res);
  save_settings(ctx, rs, R, seam_wrap);
  ctx->save_for_backward({rs.bg_, rs.viewmatrix_, rs.projmatrix_, rs.campos_, colors_precomp, means3D, scales,
                          rotations, cov3Ds_precomp, radii, sh, std::get<3>(res), std::get<4>(res),
                          std::get<5>(res)});
  return {color, radii};
}

tensor_list stored_backward(AutogradContext* ctx, const tensor_list& grad_outputs) {
  const int R = ctx->saved_data["num_rendered"].toInt();
  const float scale_modifier = (float)ctx->saved_data["scale_modifier"].toDouble();
  const float tanfovx = (float)ctx->saved_data["tanfovx"].toDouble();
  const float tanfovy = (float)ctx->saved_data["tanfovy"].toDouble();
  const int sh_degree = ctx->saved_data["sh_degree"].toInt();
  const int camera_type = ctx->saved_data["camera_type"].toInt();
  const bool seam_wrap = ctx->saved_data["seam_wrap"].toBool();
  auto s = ctx->get_saved_variables();
  auto r = omnigs::RasterizeGaussiansBackwardCUDA(s[0], s[5], s[9], s[4], s[6], s[7], scale_modifier, s[8], s[1],
                                                  s[2], tanfovx, tanfovy, grad_outputs[0].contiguous(), s[10],
                                                  sh_degree, s[3], s[11], R, s[12], s[13], camera_type, seam_wrap);
  return grads_in_reference_order(r);
}

// Empty tensors for the absent inputs, and the reference's argument checks
// (src/gaussian_rasterizer.cpp:172-224).
tensor_list module_forward(torch::Tensor& means3D, torch::Tensor& means2D, torch::Tensor& opacities, bool has_shs,
                           bool has_colors_precomp, bool has_scales, bool has_rotations, bool has_cov3D_precomp,
                           torch::Tensor& shs, torch::Tensor& colors_precomp, torch::Tensor& scales,
                           torch::Tensor& rotations, torch::Tensor& cov3D_precomp,
                           GaussianRasterizationSettings& rs, const omnigs::RasterizerOptions* options) {
  if ((!has_shs && !has_colors_precomp) || (has_shs && has_colors_precomp))
    throw std::runtime_error("Please provide excatly one of either SHs or precomputed colors!");
  if (((!has_scales || !has_rotations) && !has_cov3D_precomp) ||
      ((has_scales || has_rotations) && has_cov3D_precomp))
    throw std::runtime_error(
        "Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!");
  auto empty = torch::empty({0}, means3D.options());
  if (!has_shs) shs = empty;
  if (!has_colors_precomp) colors_precomp = empty;
  if (!has_scales) scales = empty;
  if (!has_rotations) rotations = empty;
  if (!has_cov3D_precomp) cov3D_precomp = empty;
  if (!options)
    return rasterizeGaussians(means3D, means2D, shs, colors_precomp, opacities, scales, rotations, cov3D_precomp, rs);
  return omnigs::rasterizeGaussians(means3D, means2D, shs, colors_precomp, opacities, scales, rotations,
                                    cov3D_precomp, rs, *options);
}

// markVisibleGaussians: the reference calls markVisible with its default
// camera_type (1, the pinhole frustum test) whatever the settings hold
// (src/gaussian_rasterizer.cpp:24-32); this build serves camera_type 3 only,
// so the settings' own type is passed.
torch::Tensor mark_visible(GaussianRasterizationSettings& rs, torch::Tensor& positions) {
  torch::NoGradGuard no_grad;
  return markVisible(positions, rs.viewmatrix_, rs.projmatrix_, (int)rs.camera_type_);
}

}  // namespace

// ---- the reference's classes ----

tensor_list GaussianRasterizerFunction::forward(AutogradContext* ctx, torch::Tensor means3D,
                                                torch::Tensor /*means2D*/, torch::Tensor sh,
                                                torch::Tensor colors_precomp, torch::Tensor opacities,
                                                torch::Tensor scales, torch::Tensor rotations,
                                                torch::Tensor cov3Ds_precomp,
                                                GaussianRasterizationSettings raster_settings) {
  return stored_forward(ctx, means3D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp,
                        raster_settings, /*seam_wrap=*/false);
}

tensor_list GaussianRasterizerFunction::backward(AutogradContext* ctx, tensor_list grad_outputs) {
  return stored_backward(ctx, grad_outputs);
}

torch::Tensor GaussianRasterizer::markVisibleGaussians(torch::Tensor& positions) {
  return mark_visible(raster_settings_, positions);
}

std::tuple<torch::Tensor, torch::Tensor> GaussianRasterizer::forward(
    torch::Tensor means3D, torch::Tensor means2D, torch::Tensor opacities, bool has_shs, bool has_colors_precomp,
    bool has_scales, bool has_rotations, bool has_cov3D_precomp, torch::Tensor shs, torch::Tensor colors_precomp,
    torch::Tensor scales, torch::Tensor rotations, torch::Tensor cov3D_precomp) {
  auto rs = raster_settings_;
  auto result = module_forward(means3D, means2D, opacities, has_shs, has_colors_precomp, has_scales, has_rotations,
                               has_cov3D_precomp, shs, colors_precomp, scales, rotations, cov3D_precomp, rs, nullptr);
  return std::make_tuple(result[0], result[1]);
}

// ---- extensions ----

namespace omnigs {

tensor_list StoredRasterizerFunction::forward(AutogradContext* ctx, torch::Tensor means3D, torch::Tensor /*means2D*/,
                                              torch::Tensor sh, torch::Tensor colors_precomp,
                                              torch::Tensor opacities, torch::Tensor scales,
                                              torch::Tensor rotations, torch::Tensor cov3Ds_precomp,
                                              GaussianRasterizationSettings raster_settings, bool seam_wrap) {
  auto out = stored_forward(ctx, means3D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp,
                            raster_settings, seam_wrap);
  ctx->mark_non_differentiable({out[1]});
  return out;
}

tensor_list StoredRasterizerFunction::backward(AutogradContext* ctx, tensor_list grad_outputs) {
  auto g = stored_backward(ctx, grad_outputs);
  g.emplace_back();  // the seam_wrap switch
  return g;
}

tensor_list CheckpointRasterizerFunction::forward(AutogradContext* ctx, torch::Tensor means3D,
                                                  torch::Tensor /*means2D*/, torch::Tensor sh,
                                                  torch::Tensor colors_precomp, torch::Tensor opacities,
                                                  torch::Tensor scales, torch::Tensor rotations,
                                                  torch::Tensor cov3Ds_precomp,
                                                  GaussianRasterizationSettings rs, bool seam_wrap,
                                                  int64_t band_instances) {
  // the binning buffer is carved without gradient records: the backward bins again
  auto res = omnigs::RasterizeGaussiansCheckpointedCUDA(
      rs.bg_, means3D, colors_precomp, opacities, scales, rotations, rs.scale_modifier_, cov3Ds_precomp,
      rs.viewmatrix_, rs.image_height_, rs.image_width_, sh, rs.sh_degree_, rs.campos_, rs.prefiltered_,
      (int)rs.camera_type_, seam_wrap);
  const int R = std::get<0>(res);
  auto color = std::get<1>(res);
  auto radii = std::get<2>(res);
  save_settings(ctx, rs, R, seam_wrap);
  ctx->saved_data["band_instances"] = band_instances;
  // geometry and binning buffers are dropped here; only the image buffer is kept
  ctx->save_for_backward({rs.bg_, rs.viewmatrix_, rs.projmatrix_, rs.campos_, colors_precomp, means3D, scales,
                          rotations, cov3Ds_precomp, radii, sh, opacities, std::get<5>(res)});
  ctx->mark_non_differentiable({radii});
  return {color, radii};
}

tensor_list CheckpointRasterizerFunction::backward(AutogradContext* ctx, tensor_list grad_outputs) {
  const int R = ctx->saved_data["num_rendered"].toInt();
  const float scale_modifier = (float)ctx->saved_data["scale_modifier"].toDouble();
  const int sh_degree = ctx->saved_data["sh_degree"].toInt();
  const int camera_type = ctx->saved_data["camera_type"].toInt();
  const bool prefiltered = ctx->saved_data["prefiltered"].toBool();
  const bool seam_wrap = ctx->saved_data["seam_wrap"].toBool();
  const int64_t band_instances = ctx->saved_data["band_instances"].toInt();
  auto s = ctx->get_saved_variables();
  // preprocess, depth sort, then binning and render backward per band of tile rows
  auto r = omnigs::RasterizeGaussiansBackwardCheckpointedCUDA(
      s[0], s[5], s[9], s[4], s[11], s[6], s[7], scale_modifier, s[8], s[1], grad_outputs[0].contiguous(), s[10],
      sh_degree, s[3], prefiltered, R, s[12], camera_type, seam_wrap, band_instances);
  auto g = grads_in_reference_order(r);
  g.emplace_back();  // the seam_wrap switch
  g.emplace_back();  // the band budget
  return g;
}

torch::Tensor LonlatRasterizer::markVisibleGaussians(torch::Tensor& positions) {
  return mark_visible(raster_settings_, positions);
}

std::tuple<torch::Tensor, torch::Tensor> LonlatRasterizer::forward(
    torch::Tensor means3D, torch::Tensor means2D, torch::Tensor opacities, bool has_shs, bool has_colors_precomp,
    bool has_scales, bool has_rotations, bool has_cov3D_precomp, torch::Tensor shs, torch::Tensor colors_precomp,
    torch::Tensor scales, torch::Tensor rotations, torch::Tensor cov3D_precomp) {
  auto rs = raster_settings_;
  auto result = module_forward(means3D, means2D, opacities, has_shs, has_colors_precomp, has_scales, has_rotations,
                               has_cov3D_precomp, shs, colors_precomp, scales, rotations, cov3D_precomp, rs,
                               &options_);
  return std::make_tuple(result[0], result[1]);
}

}  // namespace omnigs
