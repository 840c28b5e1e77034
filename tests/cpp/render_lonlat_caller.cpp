// A caller of the rasterizer written against the reference's interface alone:
// it includes gaussian_rasterizer.h (GaussianRasterizationSettings,
// GaussianRasterizer, Camera::CameraModelType) and nothing of this library's
// own API, and follows GaussianRenderer::renderLonlat
// (src/gaussian_renderer.cpp:166-291) step for step: a zero screen-space
// points tensor that keeps its gradient, the settings built with
// Camera::CameraModelType::LONLAT, the model's activations
// (src/gaussian_model.cpp:54-77), precomputed covariances or scale/rotation,
// SH or override colours, and the (render, viewspace_points,
// visibility_filter, radii) tuple.  It is compiled and linked against
// libomnigs_torch.so by the build, so a reference-built caller of the same
// shape is known to compile and link; the GPU tests drive it through this
// module and compare it with the CPU restatement of the reference.
//
// TEST INFRASTRUCTURE: not part of the product.
#include <torch/extension.h>

#include <tuple>

#include "gaussian_rasterizer.h"

namespace {

std::tuple<torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor> render_lonlat(
    torch::Tensor xyz, torch::Tensor features_dc, torch::Tensor features_rest, torch::Tensor scaling,
    torch::Tensor rotation, torch::Tensor opacity, torch::Tensor world_view_transform, torch::Tensor camera_center,
    int image_height, int image_width, torch::Tensor bg_color, int active_sh_degree, float scaling_modifier,
    bool compute_cov3D, torch::Tensor override_color, bool use_override_color) {
  // Create zero tensor. Its gradient receives dL/dmeans2D.
  auto screenspace_points = torch::zeros_like(
      xyz, torch::TensorOptions().dtype(xyz.dtype()).requires_grad(true).device(torch::kCUDA));
  try {
    screenspace_points.retain_grad();
  } catch (const std::exception&) {
  }

  torch::Tensor full_proj_transform = world_view_transform;  // unused by the equirectangular camera
  torch::Tensor RwcT = torch::Tensor();
  GaussianRasterizationSettings raster_settings(image_height, image_width, 0.0f, 0.0f, bg_color, scaling_modifier,
                                                world_view_transform, full_proj_transform, RwcT, active_sh_degree,
                                                camera_center, false, Camera::CameraModelType::LONLAT);

  GaussianRasterizer rasterizer(raster_settings);

  auto means3D = xyz;
  auto means2D = screenspace_points;
  auto opacity_act = torch::sigmoid(opacity);

  bool has_scales = false, has_rotations = false, has_cov3D_precomp = false;
  torch::Tensor scales, rotations, cov3D_precomp;
  if (compute_cov3D) {
    // getCovarianceActivation: the upper triangle of (S R)^T (S R)
    auto s = torch::exp(scaling) * scaling_modifier;
    auto q = torch::nn::functional::normalize(rotation);
    auto w = q.select(1, 0), x = q.select(1, 1), y = q.select(1, 2), z = q.select(1, 3);
    auto Rm = torch::stack({torch::stack({1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)}, -1),
                            torch::stack({2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)}, -1),
                            torch::stack({2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}, -1)},
                           1);
    auto L = Rm * s.unsqueeze(1);
    auto Sig = torch::matmul(L, L.transpose(1, 2).contiguous());
    cov3D_precomp = torch::stack({Sig.select(1, 0).select(1, 0), Sig.select(1, 0).select(1, 1),
                                  Sig.select(1, 0).select(1, 2), Sig.select(1, 1).select(1, 1),
                                  Sig.select(1, 1).select(1, 2), Sig.select(1, 2).select(1, 2)},
                                 -1)
                        .contiguous();
    has_cov3D_precomp = true;
  } else {
    scales = torch::exp(scaling);
    rotations = torch::nn::functional::normalize(rotation);
    has_scales = true;
    has_rotations = true;
  }

  bool has_shs = false, has_color_precomp = false;
  torch::Tensor shs, colors_precomp;
  if (use_override_color) {
    colors_precomp = override_color;
    has_color_precomp = true;
  } else {
    shs = torch::cat({features_dc, features_rest}, 1);
    has_shs = true;
  }

  auto rasterizer_result = rasterizer.forward(means3D, means2D, opacity_act, has_shs, has_color_precomp, has_scales,
                                              has_rotations, has_cov3D_precomp, shs, colors_precomp, scales,
                                              rotations, cov3D_precomp);
  auto rendered_image = std::get<0>(rasterizer_result);
  auto radii = std::get<1>(rasterizer_result);
  return std::make_tuple(rendered_image, screenspace_points, radii > 0, radii);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "renderLonlat-shaped caller of the reference rasterizer interface (test infrastructure)";
  m.def("render_lonlat", &render_lonlat);
}
