// Python bindings of the LibTorch boundary (module omnigs_amd.lib._omnigs_torch).
#include <torch/extension.h>

#include "gaussian_rasterizer.h"
#include "omnigs_rasterizer.h"
#include "rasterize_points.h"

namespace {

// Applies the stored or checkpointed autograd Function; returns (color, radii).
std::tuple<torch::Tensor, torch::Tensor> rasterize_gaussians(
    torch::Tensor means3D, torch::Tensor means2D, torch::Tensor opacities, torch::Tensor shs,
    torch::Tensor colors_precomp, torch::Tensor scales, torch::Tensor rotations, torch::Tensor cov3D_precomp,
    int image_height, int image_width, torch::Tensor bg, float scale_modifier, torch::Tensor viewmatrix,
    torch::Tensor projmatrix, int sh_degree, torch::Tensor campos, bool prefiltered, int camera_type,
    bool checkpoint, bool seam_wrap, int64_t band_instances) {
  torch::Tensor RwcT;
  GaussianRasterizationSettings rs(image_height, image_width, 0.f, 0.f, bg, scale_modifier, viewmatrix, projmatrix,
                                   RwcT, sh_degree, campos, prefiltered, (Camera::CameraModelType)camera_type, false);
  omnigs::RasterizerOptions opt;
  opt.checkpoint = checkpoint;
  opt.seam_wrap = seam_wrap;
  opt.band_instances = band_instances;
  omnigs::LonlatRasterizer rasterizer(rs, opt);
  const bool has_shs = shs.defined() && shs.numel() > 0;
  const bool has_cols = colors_precomp.defined() && colors_precomp.numel() > 0;
  const bool has_sc = scales.defined() && scales.numel() > 0;
  const bool has_rot = rotations.defined() && rotations.numel() > 0;
  const bool has_cov = cov3D_precomp.defined() && cov3D_precomp.numel() > 0;
  auto empty = torch::empty({0}, means3D.options());
  return rasterizer.forward(means3D, means2D, opacities, has_shs, has_cols, has_sc, has_rot, has_cov,
                            has_shs ? shs : empty, has_cols ? colors_precomp : empty, has_sc ? scales : empty,
                            has_rot ? rotations : empty, has_cov ? cov3D_precomp : empty);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "LibTorch-ROCm boundary of the gfx950 equirectangular Gaussian rasterizer";
  // the reference's functions; seam_wrap = true selects the omnigs:: extension
  m.def(
      "RasterizeGaussiansCUDA",
      [](const torch::Tensor& background, const torch::Tensor& means3D, const torch::Tensor& colors,
         const torch::Tensor& opacity, const torch::Tensor& scales, const torch::Tensor& rotations,
         float scale_modifier, const torch::Tensor& cov3D_precomp, const torch::Tensor& viewmatrix,
         const torch::Tensor& projmatrix, float tan_fovx, float tan_fovy, int image_height, int image_width,
         const torch::Tensor& sh, int degree, const torch::Tensor& campos, bool prefiltered, int camera_type,
         bool render_depth, bool seam_wrap) {
        if (!seam_wrap)
          return RasterizeGaussiansCUDA(background, means3D, colors, opacity, scales, rotations, scale_modifier,
                                        cov3D_precomp, viewmatrix, projmatrix, tan_fovx, tan_fovy, image_height,
                                        image_width, sh, degree, campos, prefiltered, camera_type, render_depth);
        return omnigs::RasterizeGaussiansCUDA(background, means3D, colors, opacity, scales, rotations, scale_modifier,
                                              cov3D_precomp, viewmatrix, projmatrix, tan_fovx, tan_fovy,
                                              image_height, image_width, sh, degree, campos, prefiltered,
                                              camera_type, render_depth, true);
      },
      py::arg("background"), py::arg("means3D"), py::arg("colors"), py::arg("opacity"), py::arg("scales"),
      py::arg("rotations"), py::arg("scale_modifier"), py::arg("cov3D_precomp"), py::arg("viewmatrix"),
      py::arg("projmatrix"), py::arg("tan_fovx"), py::arg("tan_fovy"), py::arg("image_height"),
      py::arg("image_width"), py::arg("sh"), py::arg("degree"), py::arg("campos"), py::arg("prefiltered"),
      py::arg("camera_type") = 1, py::arg("render_depth") = false, py::arg("seam_wrap") = false);
  m.def(
      "RasterizeGaussiansBackwardCUDA",
      [](const torch::Tensor& background, const torch::Tensor& means3D, const torch::Tensor& radii,
         const torch::Tensor& colors, const torch::Tensor& scales, const torch::Tensor& rotations,
         float scale_modifier, const torch::Tensor& cov3D_precomp, const torch::Tensor& viewmatrix,
         const torch::Tensor& projmatrix, float tan_fovx, float tan_fovy, const torch::Tensor& dL_dout_color,
         const torch::Tensor& sh, int degree, const torch::Tensor& campos, const torch::Tensor& geomBuffer, int R,
         const torch::Tensor& binningBuffer, const torch::Tensor& imageBuffer, int camera_type, bool seam_wrap) {
        if (!seam_wrap)
          return RasterizeGaussiansBackwardCUDA(background, means3D, radii, colors, scales, rotations,
                                                scale_modifier, cov3D_precomp, viewmatrix, projmatrix, tan_fovx,
                                                tan_fovy, dL_dout_color, sh, degree, campos, geomBuffer, R,
                                                binningBuffer, imageBuffer, camera_type);
        return omnigs::RasterizeGaussiansBackwardCUDA(background, means3D, radii, colors, scales, rotations,
                                                      scale_modifier, cov3D_precomp, viewmatrix, projmatrix,
                                                      tan_fovx, tan_fovy, dL_dout_color, sh, degree, campos,
                                                      geomBuffer, R, binningBuffer, imageBuffer, camera_type, true);
      },
      py::arg("background"), py::arg("means3D"), py::arg("radii"), py::arg("colors"), py::arg("scales"),
      py::arg("rotations"), py::arg("scale_modifier"), py::arg("cov3D_precomp"), py::arg("viewmatrix"),
      py::arg("projmatrix"), py::arg("tan_fovx"), py::arg("tan_fovy"), py::arg("dL_dout_color"), py::arg("sh"),
      py::arg("degree"), py::arg("campos"), py::arg("geomBuffer"), py::arg("R"), py::arg("binningBuffer"),
      py::arg("imageBuffer"), py::arg("camera_type") = 1, py::arg("seam_wrap") = false);
  m.def("markVisible", &markVisible, py::arg("means3D"), py::arg("viewmatrix"), py::arg("projmatrix"),
        py::arg("camera_type") = 1);
  m.def("rasterize_gaussians", &rasterize_gaussians, py::arg("means3D"), py::arg("means2D"), py::arg("opacities"),
        py::arg("shs"), py::arg("colors_precomp"), py::arg("scales"), py::arg("rotations"), py::arg("cov3D_precomp"),
        py::arg("image_height"), py::arg("image_width"), py::arg("bg"), py::arg("scale_modifier"),
        py::arg("viewmatrix"), py::arg("projmatrix"), py::arg("sh_degree"), py::arg("campos"),
        py::arg("prefiltered") = false, py::arg("camera_type") = 3, py::arg("checkpoint") = false,
        py::arg("seam_wrap") = false, py::arg("band_instances") = 0);
}
