// Autograd Functions over the LibTorch boundary; see gaussian_rasterizer.h.
// The stored form follows the reference's src/gaussian_rasterizer.cpp:34-170.
#include "gaussian_rasterizer.h"

#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>

#include <stdexcept>
#include <string>

#include "omnigs_lonlat.h"

namespace omnigs_torch {

namespace {

torch::autograd::tensor_list grads_in_reference_order(
    const std::tuple<torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor,
                     torch::Tensor, torch::Tensor>& r) {
  return {std::get<3>(r) /*dL_dmeans3D*/, std::get<0>(r) /*dL_dmeans2D*/,  std::get<5>(r) /*dL_dsh*/,
          std::get<1>(r) /*dL_dcolors*/,  std::get<2>(r) /*dL_dopacity*/,  std::get<6>(r) /*dL_dscales*/,
          std::get<7>(r) /*dL_drotations*/, std::get<4>(r) /*dL_dcov3D*/, torch::Tensor()};
}

void save_settings(torch::autograd::AutogradContext* ctx, const GaussianRasterizationSettings& rs, int R) {
  ctx->saved_data["num_rendered"] = R;
  ctx->saved_data["scale_modifier"] = rs.scale_modifier_;
  ctx->saved_data["tanfovx"] = rs.tanfovx_;
  ctx->saved_data["tanfovy"] = rs.tanfovy_;
  ctx->saved_data["sh_degree"] = rs.sh_degree_;
  ctx->saved_data["camera_type"] = (int)rs.camera_type_;
  ctx->saved_data["image_height"] = rs.image_height_;
  ctx->saved_data["image_width"] = rs.image_width_;
  ctx->saved_data["prefiltered"] = rs.prefiltered_;
  ctx->saved_data["seam_wrap"] = rs.seam_wrap_;
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

}  // namespace

torch::autograd::tensor_list GaussianRasterizerFunction::forward(
    torch::autograd::AutogradContext* ctx, torch::Tensor means3D, torch::Tensor /*means2D*/, torch::Tensor sh,
    torch::Tensor colors_precomp, torch::Tensor opacities, torch::Tensor scales, torch::Tensor rotations,
    torch::Tensor cov3Ds_precomp, GaussianRasterizationSettings rs) {
  auto res = RasterizeGaussiansCUDA(rs.bg_, means3D, colors_precomp, opacities, scales, rotations,
                                    rs.scale_modifier_, cov3Ds_precomp, rs.viewmatrix_, rs.projmatrix_, rs.tanfovx_,
                                    rs.tanfovy_, rs.image_height_, rs.image_width_, sh, rs.sh_degree_, rs.campos_,
                                    rs.prefiltered_, (int)rs.camera_type_, rs.render_depth_, rs.seam_wrap_);
  const int R = std::get<0>(res);
  auto color = std::get<1>(res);
  auto radii = std::get<2>(res);
  save_settings(ctx, rs, R);
  ctx->save_for_backward({rs.bg_, rs.viewmatrix_, rs.projmatrix_, rs.campos_, colors_precomp, means3D, scales,
                          rotations, cov3Ds_precomp, radii, sh, std::get<3>(res), std::get<4>(res),
                          std::get<5>(res)});
  ctx->mark_non_differentiable({radii});
  return {color, radii};
}

torch::autograd::tensor_list GaussianRasterizerFunction::backward(torch::autograd::AutogradContext* ctx,
                                                                  torch::autograd::tensor_list grad_outputs) {
  const int R = ctx->saved_data["num_rendered"].toInt();
  const float scale_modifier = (float)ctx->saved_data["scale_modifier"].toDouble();
  const float tanfovx = (float)ctx->saved_data["tanfovx"].toDouble();
  const float tanfovy = (float)ctx->saved_data["tanfovy"].toDouble();
  const int sh_degree = ctx->saved_data["sh_degree"].toInt();
  const int camera_type = ctx->saved_data["camera_type"].toInt();
  auto s = ctx->get_saved_variables();
  auto r = RasterizeGaussiansBackwardCUDA(s[0], s[5], s[9], s[4], s[6], s[7], scale_modifier, s[8], s[1], s[2],
                                          tanfovx, tanfovy, grad_outputs[0].contiguous(), s[10], sh_degree, s[3],
                                          s[11], R, s[12], s[13], camera_type,
                                          ctx->saved_data["seam_wrap"].toBool());
  return grads_in_reference_order(r);
}

// ---- checkpointed form ----

torch::autograd::tensor_list GaussianRasterizerCheckpointFunction::forward(
    torch::autograd::AutogradContext* ctx, torch::Tensor means3D, torch::Tensor /*means2D*/, torch::Tensor sh,
    torch::Tensor colors_precomp, torch::Tensor opacities, torch::Tensor scales, torch::Tensor rotations,
    torch::Tensor cov3Ds_precomp, GaussianRasterizationSettings rs) {
  auto res = RasterizeGaussiansCUDA(rs.bg_, means3D, colors_precomp, opacities, scales, rotations,
                                    rs.scale_modifier_, cov3Ds_precomp, rs.viewmatrix_, rs.projmatrix_, rs.tanfovx_,
                                    rs.tanfovy_, rs.image_height_, rs.image_width_, sh, rs.sh_degree_, rs.campos_,
                                    rs.prefiltered_, (int)rs.camera_type_, rs.render_depth_, rs.seam_wrap_);
  const int R = std::get<0>(res);
  auto color = std::get<1>(res);
  auto radii = std::get<2>(res);
  save_settings(ctx, rs, R);
  // geometry and binning buffers are dropped here; only the image buffer is kept
  ctx->save_for_backward({rs.bg_, rs.viewmatrix_, rs.projmatrix_, rs.campos_, colors_precomp, means3D, scales,
                          rotations, cov3Ds_precomp, radii, sh, opacities, std::get<5>(res)});
  ctx->mark_non_differentiable({radii});
  return {color, radii};
}

torch::autograd::tensor_list GaussianRasterizerCheckpointFunction::backward(
    torch::autograd::AutogradContext* ctx, torch::autograd::tensor_list grad_outputs) {
  const int R = ctx->saved_data["num_rendered"].toInt();
  const float scale_modifier = (float)ctx->saved_data["scale_modifier"].toDouble();
  const float tanfovx = (float)ctx->saved_data["tanfovx"].toDouble();
  const float tanfovy = (float)ctx->saved_data["tanfovy"].toDouble();
  const int sh_degree = ctx->saved_data["sh_degree"].toInt();
  const int camera_type = ctx->saved_data["camera_type"].toInt();
  const int H = ctx->saved_data["image_height"].toInt();
  const int W = ctx->saved_data["image_width"].toInt();
  const bool prefiltered = ctx->saved_data["prefiltered"].toBool();
  auto s = ctx->get_saved_variables();
  torch::Tensor bg = s[0].contiguous(), vm = s[1].contiguous(), cp = s[3].contiguous(), cols = s[4].contiguous(),
                means3D = s[5].contiguous(), scales = s[6].contiguous(), rots = s[7].contiguous(),
                cov = s[8].contiguous(), radii = s[9], sh = s[10].contiguous(), opac = s[11].contiguous(),
                img = s[12];
  const int P = (int)means3D.size(0);
  torch::TensorOptions byte_opts = torch::TensorOptions(torch::kByte).device(means3D.device());
  torch::Tensor geom = torch::empty({0}, byte_opts), binning = torch::empty({0}, byte_opts);
  if (P != 0 && camera_type == LONLAT) {
    const c10::DeviceGuard guard(means3D.device());
    omnigs_lonlat_forward_args a{};
    a.P = P;
    a.D = sh_degree;
    a.M = sh.size(0) != 0 ? (int)sh.size(1) : 0;
    a.width = W;
    a.height = H;
    a.background = fptr(bg);
    a.means3D = fptr(means3D);
    a.shs = fptr(sh);
    a.colors_precomp = fptr(cols);
    a.opacities = fptr(opac);
    a.scales = fptr(scales);
    a.scale_modifier = scale_modifier;
    a.rotations = fptr(rots);
    a.cov3D_precomp = fptr(cov);
    a.viewmatrix = fptr(vm);
    a.campos = fptr(cp);
    a.prefiltered = prefiltered ? 1 : 0;
    a.seam_wrap = ctx->saved_data["seam_wrap"].toBool() ? 1 : 0;
    omnigs_buffer gb{&resize_tensor, &geom}, bb{&resize_tensor, &binning};
    int R2 = 0;
    void* stream = reinterpret_cast<void*>(c10::hip::getCurrentHIPStream(means3D.device().index()).stream());
    const int st = omnigs_lonlat_recompute(&a, gb, bb, reinterpret_cast<char*>(img.data_ptr()), stream, &R2);
    if (st != OMNIGS_OK)
      throw std::runtime_error(std::string("checkpointed backward: ") + omnigs_status_string(st) + " (" +
                               omnigs_last_error() + ")");
    if (R2 != R) throw std::runtime_error("checkpointed backward: recomputed instance count differs from forward");
  }
  auto r = RasterizeGaussiansBackwardCUDA(bg, means3D, radii, cols, scales, rots, scale_modifier, cov, vm, s[2],
                                          tanfovx, tanfovy, grad_outputs[0].contiguous(), sh, sh_degree, cp, geom, R,
                                          binning, img, camera_type, ctx->saved_data["seam_wrap"].toBool());
  return grads_in_reference_order(r);
}

// ---- module ----

torch::Tensor GaussianRasterizer::markVisibleGaussians(torch::Tensor& positions) {
  torch::NoGradGuard no_grad;
  return markVisible(positions, raster_settings_.viewmatrix_, raster_settings_.projmatrix_,
                     (int)raster_settings_.camera_type_);
}

std::tuple<torch::Tensor, torch::Tensor> GaussianRasterizer::forward(
    torch::Tensor means3D, torch::Tensor means2D, torch::Tensor opacities, bool has_shs, bool has_colors_precomp,
    bool has_scales, bool has_rotations, bool has_cov3D_precomp, torch::Tensor shs, torch::Tensor colors_precomp,
    torch::Tensor scales, torch::Tensor rotations, torch::Tensor cov3D_precomp) {
  auto rs = raster_settings_;
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
  auto result = rasterizeGaussians(means3D, means2D, shs, colors_precomp, opacities, scales, rotations, cov3D_precomp,
                                   rs, checkpoint_);
  return std::make_tuple(result[0], result[1]);
}

}  // namespace omnigs_torch
