// LibTorch-ROCm boundary: tensors -> the C ABI of include/omnigs_lonlat.h.
//
// Mirrors src/rasterize_points.cu of the reference (:49-164 forward,
// :166-285 backward, :287-319 markVisible): the same argument meaning, the
// same output tuple layout, zero-filled outputs, kByte scratch buffers grown
// through a resize callback, P == 0 skipping the rasterizer, and the same
// exceptions for a bad means3D shape or an invalid camera_type.  Work runs on
// the current HIP stream of means3D's device.  The global functions carry the
// reference's signatures; the omnigs:: overloads add the seam-wrap switch.
#include "rasterize_points.h"

#include "omnigs_rasterizer.h"

#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>

#include <stdexcept>
#include <string>

#include "omnigs_lonlat.h"

namespace {

constexpr int kLonlat = 3;   // Camera::CameraModelType::LONLAT (reference include/camera.h:54-58)
constexpr int kPinhole = 1;  // served by the reference's pinhole rasterizer, not by this build

void check(int status, const char* where) {
  if (status == OMNIGS_OK) return;
  throw std::runtime_error(std::string(where) + ": " + omnigs_status_string(status) + " (" + omnigs_last_error() +
                           ")");
}

[[noreturn]] void bad_camera(int camera_type) {
  if (camera_type == kPinhole)
    throw std::runtime_error("[CudaRasterizer]camera_type 1 (pinhole) is not provided by this equirectangular build");
  throw std::runtime_error("[CudaRasterizer]Invalid camera_type");
}

// resize_ functor in C form (reference resizeFunctional, rasterize_points.cu:41-47)
char* resize_tensor(void* ctx, size_t nbytes) {
  try {
    auto* t = static_cast<torch::Tensor*>(ctx);
    t->resize_({(int64_t)nbytes});
    return reinterpret_cast<char*>(t->data_ptr());
  } catch (...) {
    return nullptr;  // surfaces as OMNIGS_ERR_ALLOC
  }
}

// Contiguous FP32 view kept alive by the holder; an empty tensor means "absent".
struct F32 {
  torch::Tensor t;
  explicit F32(const torch::Tensor& x) : t(x.numel() ? x.contiguous() : x) {}
  const float* ptr() const { return t.numel() ? t.data_ptr<float>() : nullptr; }
  float* mut() { return t.numel() ? t.data_ptr<float>() : nullptr; }
};

void* current_stream(const torch::Tensor& ref) {
  return reinterpret_cast<void*>(c10::hip::getCurrentHIPStream(ref.device().index()).stream());
}

}  // namespace

namespace omnigs {

namespace {

// The forward of RasterizeGaussiansCUDA; checkpointed: through
// omnigs_lonlat_forward_checkpointed (a binning buffer without the
// backward's gradient records).
std::tuple<int, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor> forward_common(
    const torch::Tensor& background, const torch::Tensor& means3D, const torch::Tensor& colors,
    const torch::Tensor& opacity, const torch::Tensor& scales, const torch::Tensor& rotations,
    const float scale_modifier, const torch::Tensor& cov3D_precomp, const torch::Tensor& viewmatrix,
    const int image_height, const int image_width, const torch::Tensor& sh, const int degree,
    const torch::Tensor& campos, const bool prefiltered, const int camera_type, const bool seam_wrap,
    const bool checkpointed) {
  if (means3D.ndimension() != 2 || means3D.size(1) != 3) {
    AT_ERROR("means3D must have dimensions (num_points, 3)");
  }
  const int P = (int)means3D.size(0);
  const int H = image_height;
  const int W = image_width;
  const c10::DeviceGuard guard(means3D.device());
  auto float_opts = means3D.options().dtype(torch::kFloat32);
  torch::Tensor out_color = torch::full({3, H, W}, 0.0, float_opts);
  torch::Tensor radii = torch::full({P}, 0, means3D.options().dtype(torch::kInt32));
  torch::TensorOptions byte_opts = torch::TensorOptions(torch::kByte).device(means3D.device());
  torch::Tensor geomBuffer = torch::empty({0}, byte_opts);
  torch::Tensor binningBuffer = torch::empty({0}, byte_opts);
  torch::Tensor imgBuffer = torch::empty({0}, byte_opts);

  int rendered = 0;
  if (P != 0) {
    const int M = sh.size(0) != 0 ? (int)sh.size(1) : 0;
    if (camera_type != kLonlat) bad_camera(camera_type);
    F32 bg(background), m(means3D), s(sh), c(colors), o(opacity), sc(scales), r(rotations), cv(cov3D_precomp),
        vm(viewmatrix), cp(campos);
    omnigs_lonlat_forward_args a{};
    a.P = P;
    a.D = degree;
    a.M = M;
    a.width = W;
    a.height = H;
    a.background = bg.ptr();
    a.means3D = m.ptr();
    a.shs = s.ptr();
    a.colors_precomp = c.ptr();
    a.opacities = o.ptr();
    a.scales = sc.ptr();
    a.scale_modifier = scale_modifier;
    a.rotations = r.ptr();
    a.cov3D_precomp = cv.ptr();
    a.viewmatrix = vm.ptr();
    a.campos = cp.ptr();
    a.prefiltered = prefiltered ? 1 : 0;
    a.out_color = out_color.data_ptr<float>();
    a.radii = radii.data_ptr<int>();
    a.seam_wrap = seam_wrap ? 1 : 0;
    omnigs_buffer gb{&resize_tensor, &geomBuffer}, bb{&resize_tensor, &binningBuffer}, ib{&resize_tensor, &imgBuffer};
    check(checkpointed ? omnigs_lonlat_forward_checkpointed(&a, gb, bb, ib, current_stream(means3D), &rendered)
                       : omnigs_lonlat_forward(&a, gb, bb, ib, current_stream(means3D), &rendered),
          "RasterizeGaussiansCUDA");
  }
  return std::make_tuple(rendered, out_color, radii, geomBuffer, binningBuffer, imgBuffer);
}

}  // namespace

std::tuple<int, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor> RasterizeGaussiansCUDA(
    const torch::Tensor& background, const torch::Tensor& means3D, const torch::Tensor& colors,
    const torch::Tensor& opacity, const torch::Tensor& scales, const torch::Tensor& rotations,
    const float scale_modifier, const torch::Tensor& cov3D_precomp, const torch::Tensor& viewmatrix,
    const torch::Tensor& /*projmatrix*/, const float /*tan_fovx*/, const float /*tan_fovy*/, const int image_height,
    const int image_width, const torch::Tensor& sh, const int degree, const torch::Tensor& campos,
    const bool prefiltered, const int camera_type, const bool /*render_depth*/, const bool seam_wrap) {
  return forward_common(background, means3D, colors, opacity, scales, rotations, scale_modifier, cov3D_precomp,
                        viewmatrix, image_height, image_width, sh, degree, campos, prefiltered, camera_type,
                        seam_wrap, false);
}

std::tuple<int, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor>
RasterizeGaussiansCheckpointedCUDA(const torch::Tensor& background, const torch::Tensor& means3D,
                                   const torch::Tensor& colors, const torch::Tensor& opacity,
                                   const torch::Tensor& scales, const torch::Tensor& rotations,
                                   const float scale_modifier, const torch::Tensor& cov3D_precomp,
                                   const torch::Tensor& viewmatrix, const int image_height, const int image_width,
                                   const torch::Tensor& sh, const int degree, const torch::Tensor& campos,
                                   const bool prefiltered, const int camera_type, const bool seam_wrap) {
  return forward_common(background, means3D, colors, opacity, scales, rotations, scale_modifier, cov3D_precomp,
                        viewmatrix, image_height, image_width, sh, degree, campos, prefiltered, camera_type,
                        seam_wrap, true);
}

std::tuple<torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor,
           torch::Tensor>
RasterizeGaussiansBackwardCUDA(const torch::Tensor& background, const torch::Tensor& means3D,
                               const torch::Tensor& radii, const torch::Tensor& colors, const torch::Tensor& scales,
                               const torch::Tensor& rotations, const float scale_modifier,
                               const torch::Tensor& cov3D_precomp, const torch::Tensor& viewmatrix,
                               const torch::Tensor& /*projmatrix*/, const float /*tan_fovx*/,
                               const float /*tan_fovy*/, const torch::Tensor& dL_dout_color, const torch::Tensor& sh,
                               const int degree, const torch::Tensor& campos, const torch::Tensor& geomBuffer,
                               const int R, const torch::Tensor& binningBuffer, const torch::Tensor& imageBuffer,
                               const int camera_type, const bool seam_wrap) {
  const int P = (int)means3D.size(0);
  const int H = (int)dL_dout_color.size(1);
  const int W = (int)dL_dout_color.size(2);
  const int M = sh.size(0) != 0 ? (int)sh.size(1) : 0;
  const c10::DeviceGuard guard(means3D.device());
  auto opts = means3D.options();
  torch::Tensor dL_dmeans3D = torch::zeros({P, 3}, opts);
  torch::Tensor dL_dmeans2D = torch::zeros({P, 3}, opts);
  torch::Tensor dL_dcolors = torch::zeros({P, 3}, opts);
  torch::Tensor dL_dconic = torch::zeros({P, 2, 2}, opts);
  torch::Tensor dL_dopacity = torch::zeros({P, 1}, opts);
  torch::Tensor dL_dcov3D = torch::zeros({P, 6}, opts);
  torch::Tensor dL_dsh = torch::zeros({P, M, 3}, opts);
  torch::Tensor dL_dscales = torch::zeros({P, 3}, opts);
  torch::Tensor dL_drotations = torch::zeros({P, 4}, opts);

  if (P != 0) {
    if (camera_type != kLonlat) bad_camera(camera_type);
    F32 bg(background), m(means3D), s(sh), c(colors), sc(scales), r(rotations), cv(cov3D_precomp), vm(viewmatrix),
        cp(campos), dpix(dL_dout_color);
    torch::Tensor rad = radii.contiguous();
    torch::Tensor gbuf = geomBuffer.contiguous(), bbuf = binningBuffer.contiguous(), ibuf = imageBuffer.contiguous();
    omnigs_lonlat_backward_args a{};
    a.P = P;
    a.D = degree;
    a.M = M;
    a.R = R;
    a.width = W;
    a.height = H;
    a.background = bg.ptr();
    a.means3D = m.ptr();
    a.shs = s.ptr();
    a.colors_precomp = c.ptr();
    a.scales = sc.ptr();
    a.scale_modifier = scale_modifier;
    a.rotations = r.ptr();
    a.cov3D_precomp = cv.ptr();
    a.viewmatrix = vm.ptr();
    a.campos = cp.ptr();
    a.radii = rad.numel() ? rad.data_ptr<int>() : nullptr;
    a.geom_buffer = reinterpret_cast<char*>(gbuf.data_ptr());
    a.binning_buffer = bbuf.numel() ? reinterpret_cast<char*>(bbuf.data_ptr()) : nullptr;
    a.binning_bytes = (size_t)bbuf.numel();  // the forward's layout, identified by its size
    a.image_buffer = reinterpret_cast<char*>(ibuf.data_ptr());
    a.dL_dpix = dpix.ptr();
    a.dL_dmean2D = dL_dmeans2D.data_ptr<float>();
    a.dL_dconic = dL_dconic.data_ptr<float>();
    a.dL_dopacity = dL_dopacity.data_ptr<float>();
    a.dL_dcolor = dL_dcolors.data_ptr<float>();
    a.dL_dmean3D = dL_dmeans3D.data_ptr<float>();
    a.dL_dcov3D = dL_dcov3D.data_ptr<float>();
    a.dL_dsh = M ? dL_dsh.data_ptr<float>() : nullptr;
    a.dL_dscale = dL_dscales.data_ptr<float>();
    a.dL_drot = dL_drotations.data_ptr<float>();
    a.seam_wrap = seam_wrap ? 1 : 0;
    check(omnigs_lonlat_backward(&a, current_stream(means3D)), "RasterizeGaussiansBackwardCUDA");
  }
  return std::make_tuple(dL_dmeans2D, dL_dcolors, dL_dopacity, dL_dmeans3D, dL_dcov3D, dL_dsh, dL_dscales,
                         dL_drotations);
}

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
                                           const bool seam_wrap, const int64_t band_instances) {
  const int P = (int)means3D.size(0);
  const int H = (int)dL_dout_color.size(1);
  const int W = (int)dL_dout_color.size(2);
  const int M = sh.size(0) != 0 ? (int)sh.size(1) : 0;
  const c10::DeviceGuard guard(means3D.device());
  auto opts = means3D.options();
  torch::Tensor dL_dmeans3D = torch::zeros({P, 3}, opts);
  torch::Tensor dL_dmeans2D = torch::zeros({P, 3}, opts);
  torch::Tensor dL_dcolors = torch::zeros({P, 3}, opts);
  torch::Tensor dL_dconic = torch::zeros({P, 2, 2}, opts);
  torch::Tensor dL_dopacity = torch::zeros({P, 1}, opts);
  torch::Tensor dL_dcov3D = torch::zeros({P, 6}, opts);
  torch::Tensor dL_dsh = torch::zeros({P, M, 3}, opts);
  torch::Tensor dL_dscales = torch::zeros({P, 3}, opts);
  torch::Tensor dL_drotations = torch::zeros({P, 4}, opts);

  if (P != 0) {
    if (camera_type != kLonlat) bad_camera(camera_type);
    if (band_instances < 0 || band_instances > 0xFFFFFFFFll)
      throw std::runtime_error("RasterizeGaussiansBackwardCheckpointedCUDA: band_instances out of range");
    F32 bg(background), m(means3D), s(sh), c(colors), o(opacity), sc(scales), r(rotations), cv(cov3D_precomp),
        vm(viewmatrix), cp(campos), dpix(dL_dout_color);
    torch::Tensor rad = radii.contiguous(), ibuf = imageBuffer.contiguous();
    omnigs_lonlat_forward_args f{};
    f.P = P;
    f.D = degree;
    f.M = M;
    f.width = W;
    f.height = H;
    f.background = bg.ptr();
    f.means3D = m.ptr();
    f.shs = s.ptr();
    f.colors_precomp = c.ptr();
    f.opacities = o.ptr();
    f.scales = sc.ptr();
    f.scale_modifier = scale_modifier;
    f.rotations = r.ptr();
    f.cov3D_precomp = cv.ptr();
    f.viewmatrix = vm.ptr();
    f.campos = cp.ptr();
    f.prefiltered = prefiltered ? 1 : 0;
    f.radii = rad.numel() ? rad.data_ptr<int>() : nullptr;
    f.seam_wrap = seam_wrap ? 1 : 0;
    omnigs_lonlat_backward_args a{};
    a.P = P;
    a.D = degree;
    a.M = M;
    a.R = R;
    a.width = W;
    a.height = H;
    a.background = f.background;
    a.means3D = f.means3D;
    a.shs = f.shs;
    a.colors_precomp = f.colors_precomp;
    a.scales = f.scales;
    a.scale_modifier = scale_modifier;
    a.rotations = f.rotations;
    a.cov3D_precomp = f.cov3D_precomp;
    a.viewmatrix = f.viewmatrix;
    a.campos = f.campos;
    a.radii = f.radii;
    a.image_buffer = reinterpret_cast<char*>(ibuf.data_ptr());
    // resize_ made the tensor the size the forward asked for: with the
    // checkpointed forward's blend bytes at its end
    a.image_bytes = (size_t)ibuf.numel();
    a.dL_dpix = dpix.ptr();
    a.dL_dmean2D = dL_dmeans2D.data_ptr<float>();
    a.dL_dconic = dL_dconic.data_ptr<float>();
    a.dL_dopacity = dL_dopacity.data_ptr<float>();
    a.dL_dcolor = dL_dcolors.data_ptr<float>();
    a.dL_dmean3D = dL_dmeans3D.data_ptr<float>();
    a.dL_dcov3D = dL_dcov3D.data_ptr<float>();
    a.dL_dsh = M ? dL_dsh.data_ptr<float>() : nullptr;
    a.dL_dscale = dL_dscales.data_ptr<float>();
    a.dL_drot = dL_drotations.data_ptr<float>();
    a.seam_wrap = f.seam_wrap;
    torch::TensorOptions byte_opts = torch::TensorOptions(torch::kByte).device(means3D.device());
    torch::Tensor geom = torch::empty({0}, byte_opts), binning = torch::empty({0}, byte_opts);
    omnigs_buffer gb{&resize_tensor, &geom}, bb{&resize_tensor, &binning};
    check(omnigs_lonlat_backward_checkpointed(&f, &a, gb, bb, (unsigned int)band_instances, current_stream(means3D),
                                              nullptr),
          "checkpointed backward");
  }
  return std::make_tuple(dL_dmeans2D, dL_dcolors, dL_dopacity, dL_dmeans3D, dL_dcov3D, dL_dsh, dL_dscales,
                         dL_drotations);
}

}  // namespace omnigs

std::tuple<int, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor> RasterizeGaussiansCUDA(
    const torch::Tensor& background, const torch::Tensor& means3D, const torch::Tensor& colors,
    const torch::Tensor& opacity, const torch::Tensor& scales, const torch::Tensor& rotations,
    const float scale_modifier, const torch::Tensor& cov3D_precomp, const torch::Tensor& viewmatrix,
    const torch::Tensor& projmatrix, const float tan_fovx, const float tan_fovy, const int image_height,
    const int image_width, const torch::Tensor& sh, const int degree, const torch::Tensor& campos,
    const bool prefiltered, const int camera_type, const bool render_depth) {
  return omnigs::RasterizeGaussiansCUDA(background, means3D, colors, opacity, scales, rotations, scale_modifier,
                                        cov3D_precomp, viewmatrix, projmatrix, tan_fovx, tan_fovy, image_height,
                                        image_width, sh, degree, campos, prefiltered, camera_type, render_depth,
                                        /*seam_wrap=*/false);
}

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
                               const int camera_type) {
  return omnigs::RasterizeGaussiansBackwardCUDA(background, means3D, radii, colors, scales, rotations,
                                                scale_modifier, cov3D_precomp, viewmatrix, projmatrix, tan_fovx,
                                                tan_fovy, dL_dout_color, sh, degree, campos, geomBuffer, R,
                                                binningBuffer, imageBuffer, camera_type, /*seam_wrap=*/false);
}

torch::Tensor markVisible(torch::Tensor& means3D, torch::Tensor& /*viewmatrix*/, torch::Tensor& /*projmatrix*/,
                          const int camera_type) {
  const int P = (int)means3D.size(0);
  const c10::DeviceGuard guard(means3D.device());
  torch::Tensor present = torch::full({P}, false, means3D.options().dtype(at::kBool));
  if (P != 0) {
    if (camera_type != kLonlat) bad_camera(camera_type);
    check(omnigs_lonlat_mark_visible(P, reinterpret_cast<uint8_t*>(present.data_ptr<bool>()), current_stream(means3D)),
          "markVisible");
  }
  return present;
}
