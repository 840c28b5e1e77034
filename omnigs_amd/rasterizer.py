"""Python face of the C++ LibTorch boundary (omnigs_amd/lib/_omnigs_torch*.so).

The C++ extension carries the reference's boundary functions
(RasterizeGaussiansCUDA, RasterizeGaussiansBackwardCUDA, markVisible; reference
include/rasterize_points.h:29-80) and its torch::autograd::Function
(reference src/gaussian_rasterizer.cpp:34-224), both over the C ABI.  This
module loads it (no fallback: a missing extension raises) and adds the
caller-side renderLonlat contract of the reference's GaussianRenderer
(src/gaussian_renderer.cpp:166-291): activations, a zero screen-space points
tensor whose .grad receives dL/dmeans2D, and the (image, viewspace_points,
visibility_filter, radii) tuple.
"""
from __future__ import annotations

import glob
import importlib.machinery
import importlib.util
import os
from dataclasses import dataclass

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
_ext = None

LONLAT = 3  # Camera::CameraModelType::LONLAT


def ext():
    """The loaded C++ extension module."""
    global _ext
    if _ext is not None:
        return _ext
    cands = glob.glob(os.path.join(_HERE, "lib", "_omnigs_torch*.so"))
    if not cands:
        raise RuntimeError("omnigs_amd/lib/_omnigs_torch*.so is missing: build it with `python -m omnigs_amd.build`")
    loader = importlib.machinery.ExtensionFileLoader("_omnigs_torch", cands[0])
    spec = importlib.util.spec_from_file_location("_omnigs_torch", cands[0], loader=loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    _ext = mod
    return mod


def RasterizeGaussiansCUDA(*args, **kw):
    return ext().RasterizeGaussiansCUDA(*args, **kw)


def RasterizeGaussiansBackwardCUDA(*args, **kw):
    return ext().RasterizeGaussiansBackwardCUDA(*args, **kw)


def markVisible(*args, **kw):
    return ext().markVisible(*args, **kw)


@dataclass
class GaussianRasterizationSettings:
    """Fields of the reference's GaussianRasterizationSettings
    (include/gaussian_rasterizer.h:32-72)."""
    image_height: int
    image_width: int
    bg: torch.Tensor
    viewmatrix: torch.Tensor
    campos: torch.Tensor
    sh_degree: int = 3
    scale_modifier: float = 1.0
    projmatrix: torch.Tensor | None = None
    tanfovx: float = 0.0
    tanfovy: float = 0.0
    prefiltered: bool = False
    camera_type: int = LONLAT
    seam_wrap: bool = False  # extension: wrap tile columns around the longitude seam


def _empty_like(ref: torch.Tensor) -> torch.Tensor:
    return torch.empty(0, dtype=torch.float32, device=ref.device)


def rasterize_gaussians(means3D, means2D, opacities, settings: GaussianRasterizationSettings, shs=None,
                        colors_precomp=None, scales=None, rotations=None, cov3D_precomp=None, checkpoint=False,
                        band_instances=0):
    """Runs the C++ autograd Function (stored buffers, or checkpointed when
    ``checkpoint``: the backward recomputes and bins in bands of tile rows of
    at most ``band_instances`` instances, 0 = the library's default); returns
    (color [3,H,W], radii [P])."""
    e = _empty_like(means3D)
    pm = settings.projmatrix if settings.projmatrix is not None else settings.viewmatrix
    return ext().rasterize_gaussians(
        means3D, means2D, opacities, shs if shs is not None else e, colors_precomp if colors_precomp is not None else e,
        scales if scales is not None else e, rotations if rotations is not None else e,
        cov3D_precomp if cov3D_precomp is not None else e, settings.image_height, settings.image_width, settings.bg,
        float(settings.scale_modifier), settings.viewmatrix, pm, int(settings.sh_degree), settings.campos,
        bool(settings.prefiltered), int(settings.camera_type), bool(checkpoint), bool(settings.seam_wrap),
        int(band_instances))


class GaussianRasterizer(torch.nn.Module):
    """Module form (reference include/gaussian_rasterizer.h:105-141)."""

    def __init__(self, settings: GaussianRasterizationSettings, checkpoint: bool = False, band_instances: int = 0):
        super().__init__()
        self.settings = settings
        self.checkpoint = checkpoint
        self.band_instances = band_instances

    def mark_visible(self, positions: torch.Tensor) -> torch.Tensor:
        with torch.no_grad():
            pm = self.settings.projmatrix if self.settings.projmatrix is not None else self.settings.viewmatrix
            return markVisible(positions, self.settings.viewmatrix, pm, self.settings.camera_type)

    def forward(self, means3D, means2D, opacities, shs=None, colors_precomp=None, scales=None, rotations=None,
                cov3D_precomp=None):
        if (shs is None) == (colors_precomp is None):
            raise RuntimeError("Please provide excatly one of either SHs or precomputed colors!")
        if ((scales is None or rotations is None) and cov3D_precomp is None) or (
                (scales is not None or rotations is not None) and cov3D_precomp is not None):
            raise RuntimeError("Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!")
        return rasterize_gaussians(means3D, means2D, opacities, self.settings, shs=shs, colors_precomp=colors_precomp,
                                   scales=scales, rotations=rotations, cov3D_precomp=cov3D_precomp,
                                   checkpoint=self.checkpoint, band_instances=self.band_instances)


def render_lonlat(params: dict, viewmatrix: torch.Tensor, campos: torch.Tensor, image_height: int,
                  image_width: int, bg_color: torch.Tensor, active_sh_degree: int = 3, scaling_modifier: float = 1.0,
                  checkpoint: bool = False, seam_wrap: bool = False, band_instances: int = 0):
    """The caller contract of GaussianRenderer::renderLonlat
    (src/gaussian_renderer.cpp:166-291) over raw parameters
    {xyz, features_dc [P,1,3], features_rest [P,15,3], scaling (log), rotation
    (raw quaternion), opacity (logit)} with the model's activations
    (src/gaussian_model.cpp:54-77).  Returns (image, viewspace_points,
    visibility_filter, radii); viewspace_points.grad receives dL/dmeans2D."""
    xyz = params["xyz"]
    screenspace_points = torch.zeros_like(xyz, requires_grad=True)
    try:
        screenspace_points.retain_grad()
    except RuntimeError:
        pass
    settings = GaussianRasterizationSettings(image_height=image_height, image_width=image_width, bg=bg_color,
                                             viewmatrix=viewmatrix, campos=campos, sh_degree=active_sh_degree,
                                             scale_modifier=scaling_modifier, camera_type=LONLAT,
                                             seam_wrap=seam_wrap)
    opacity = torch.sigmoid(params["opacity"])
    scales = torch.exp(params["scaling"])
    rotations = torch.nn.functional.normalize(params["rotation"])
    shs = torch.cat([params["features_dc"], params["features_rest"]], dim=1)
    raster = GaussianRasterizer(settings, checkpoint=checkpoint, band_instances=band_instances)
    image, radii = raster(xyz, screenspace_points, opacity, shs=shs, scales=scales, rotations=rotations)
    return image, screenspace_points, radii > 0, radii
