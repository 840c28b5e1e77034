"""The LibTorch boundary is a drop-in for the reference's.

CPU: libomnigs_torch.so exports RasterizeGaussiansCUDA,
RasterizeGaussiansBackwardCUDA, markVisible, GaussianRasterizerFunction and
GaussianRasterizer with the reference's signatures (global scope, same
parameter types; include/rasterize_points.h:29-80,
include/gaussian_rasterizer.h:32-141); where the reference tree is present, a
unit that includes the reference's own rasterize_points.h links against the
library with no undefined symbol, and the settings / module members are the
reference's, in its order.

GPU: tests/cpp/render_lonlat_caller.cpp -- a caller shaped like
GaussianRenderer::renderLonlat (src/gaussian_renderer.cpp:166-291) that sees
only gaussian_rasterizer.h -- renders and back-propagates through the
library, and matches the CPU oracle and the package's own path.
"""
from __future__ import annotations

import glob
import importlib.machinery
import importlib.util
import os
import re
import shutil
import subprocess
import sysconfig

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TORCH_LIB = os.path.join(ROOT, "omnigs_amd", "lib", "libomnigs_torch.so")
REF_INCLUDE = "/root/reference/include"

# Demangled signatures of the reference's declarations (the interface itself).
REFERENCE_SIGNATURES = [
    "RasterizeGaussiansCUDA(at::Tensor const&, at::Tensor const&, at::Tensor const&, at::Tensor const&, "
    "at::Tensor const&, at::Tensor const&, float, at::Tensor const&, at::Tensor const&, at::Tensor const&, float, "
    "float, int, int, at::Tensor const&, int, at::Tensor const&, bool, int, bool)",
    "RasterizeGaussiansBackwardCUDA(at::Tensor const&, at::Tensor const&, at::Tensor const&, at::Tensor const&, "
    "at::Tensor const&, at::Tensor const&, float, at::Tensor const&, at::Tensor const&, at::Tensor const&, float, "
    "float, at::Tensor const&, at::Tensor const&, int, at::Tensor const&, at::Tensor const&, int, "
    "at::Tensor const&, at::Tensor const&, int)",
    "markVisible(at::Tensor&, at::Tensor&, at::Tensor&, int)",
    "GaussianRasterizerFunction::forward(torch::autograd::AutogradContext*, at::Tensor, at::Tensor, at::Tensor, "
    "at::Tensor, at::Tensor, at::Tensor, at::Tensor, at::Tensor, GaussianRasterizationSettings)",
    "GaussianRasterizerFunction::backward(torch::autograd::AutogradContext*, std::vector<at::Tensor, "
    "std::allocator<at::Tensor> >)",
    "GaussianRasterizer::markVisibleGaussians(at::Tensor&)",
    "GaussianRasterizer::forward(at::Tensor, at::Tensor, at::Tensor, bool, bool, bool, bool, bool, at::Tensor, "
    "at::Tensor, at::Tensor, at::Tensor, at::Tensor)",
]


def _exported():
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("no nm")
    if not os.path.exists(TORCH_LIB):
        pytest.fail("libomnigs_torch.so not built (run __graft_entry__.build())")
    out = subprocess.run([nm, "-DC", "--defined-only", TORCH_LIB], check=True, capture_output=True, text=True).stdout
    return {line.split(" ", 2)[2] for line in out.splitlines() if " T " in line}


def test_reference_signatures_exported():
    syms = _exported()
    for sig in REFERENCE_SIGNATURES:
        assert sig in syms, sig


def _members(text: str, cls: str) -> list[str]:
    """Data members (type, name) of a struct/class body, in order."""
    text = re.sub(r"//[^\n]*|/\*.*?\*/", "", text, flags=re.S)
    m = re.search(r"(?:struct|class)\s+" + cls + r"\b[^{;]*\{", text)
    assert m, cls
    depth, i = 1, m.end()
    while depth:
        depth += {"{": 1, "}": -1}.get(text[i], 0)
        i += 1
    body = text[m.end():i - 1]
    # drop nested brace blocks (function bodies, initialiser lists) and member functions
    while "{" in body:
        body = re.sub(r"\{[^{}]*\}", ";", body)
    out = []
    for stmt in body.split(";"):
        stmt = re.sub(r"\b(public|private|protected)\s*:", "", stmt).strip()
        if not stmt or "(" in stmt:
            continue
        out.append(" ".join(stmt.split()))
    return out


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="reference tree absent")
def test_settings_and_module_members_are_the_references():
    ours = open(os.path.join(ROOT, "omnigs_amd", "torch_ext", "gaussian_rasterizer.h")).read()
    ref = open(os.path.join(REF_INCLUDE, "gaussian_rasterizer.h")).read()
    for cls in ("GaussianRasterizationSettings", "GaussianRasterizer"):
        assert _members(ours, cls) == _members(ref, cls), cls
    # the constructor's camera-type default, and the enumeration's values
    pin = "Camera::CameraModelType camera_type = Camera::CameraModelType::PINHOLE"
    assert pin in " ".join(ref.split()) and pin in " ".join(ours.split())
    cam_ref = " ".join(open(os.path.join(REF_INCLUDE, "camera.h")).read().split())
    cam_ours = " ".join(open(os.path.join(ROOT, "omnigs_amd", "torch_ext", "camera.h")).read().split())
    enum = "enum CameraModelType{ INVALID = 0, PINHOLE = 1, FISHEYE = 2, LONLAT = 3};"
    assert enum in cam_ref
    assert "enum CameraModelType { INVALID = 0, PINHOLE = 1, FISHEYE = 2, LONLAT = 3 };" in cam_ours


@pytest.mark.slow
@pytest.mark.skipif(not os.path.isfile(os.path.join(REF_INCLUDE, "rasterize_points.h")), reason="reference tree absent")
def test_unit_with_reference_header_links(tmp_path):
    """Compiled against the reference's own rasterize_points.h and linked
    with --no-undefined: every declaration resolves to this library."""
    import torch
    from torch.utils import cpp_extension as ce
    _exported()
    src = tmp_path / "uses_reference_header.cpp"
    src.write_text(
        f'#include "{REF_INCLUDE}/rasterize_points.h"\n'
        "void* omnigs_link_check[] = {\n"
        "  reinterpret_cast<void*>(&RasterizeGaussiansCUDA),\n"
        "  reinterpret_cast<void*>(&RasterizeGaussiansBackwardCUDA),\n"
        "  reinterpret_cast<void*>(&markVisible),\n"
        "};\n")
    torch_lib = os.path.join(os.path.dirname(torch.__file__), "lib")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    out = tmp_path / "uses_reference_header.so"
    cmd = ["g++", "-std=c++17", "-fPIC", "-shared", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           *[f"-I{p}" for p in ce.include_paths()], f"-I{sysconfig.get_paths()['include']}", str(src), "-o",
           str(out), f"-L{os.path.dirname(TORCH_LIB)}", "-lomnigs_torch", "-lomnigs_lonlat", f"-L{torch_lib}",
           "-lc10", "-ltorch", "-ltorch_cpu", "-lc10_hip", "-ltorch_hip", "-Wl,--no-undefined",
           f"-Wl,-rpath,{os.path.dirname(TORCH_LIB)}", f"-Wl,-rpath,{torch_lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_caller_module_built_without_library_api():
    """The renderLonlat-shaped caller includes the reference interface only."""
    src = open(os.path.join(ROOT, "tests", "cpp", "render_lonlat_caller.cpp")).read()
    includes = re.findall(r'#include\s+[<"]([^>"]+)[>"]', src)
    assert includes == ["torch/extension.h", "tuple", "gaussian_rasterizer.h"], includes
    assert "omnigs" not in re.sub(r"//[^\n]*", "", src)
    assert glob.glob(os.path.join(ROOT, "tests", "cpp", "lib", "_render_lonlat_caller*.so"))


# ----------------------------------------------------------------------------- GPU


def _caller():
    so = glob.glob(os.path.join(ROOT, "tests", "cpp", "lib", "_render_lonlat_caller*.so"))
    if not so:
        pytest.fail("tests/cpp/lib/_render_lonlat_caller*.so not built (run __graft_entry__.build())")
    loader = importlib.machinery.ExtensionFileLoader("_render_lonlat_caller", so[0])
    spec = importlib.util.spec_from_file_location("_render_lonlat_caller", so[0], loader=loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    return mod


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["sh", "cov3D_precomp", "override_color"])
def test_render_lonlat_caller_against_oracle(gpu, variant):
    import torch

    from oracle import Oracle
    from omnigs_amd.rasterizer import render_lonlat
    from tests.gpu_util import announce
    from tests.helpers import activate64, chain_activations, tiny_scene
    from tests.test_gpu_parity import FLIP_GRAD_FRACTION, FLIP_MAX_ABS, FLIP_PIXEL_FRACTION, GRAD_RTOL, \
        GRAD_SCALE_ATOL, IMG_ATOL, _frac_outside
    announce(f"renderLonlat-shaped C++ caller, 2000 Gaussians 256x128, {variant}")
    mod = _caller()
    sc = tiny_scene(P=2000, W=256, H=128, seed=12, log_scale=(-3.6, 0.3), radius=(1.5, 4.0))
    P = sc["means3D"].shape[0]
    f32 = lambda a: torch.tensor(np.asarray(a), dtype=torch.float32, device=gpu)  # noqa: E731
    leaf = lambda a: f32(a).requires_grad_(True)  # noqa: E731
    params = {"xyz": leaf(sc["means3D"]), "features_dc": leaf(sc["shs"][:, :1]),
              "features_rest": leaf(sc["shs"][:, 1:]), "scaling": leaf(sc["log_scales"]),
              "rotation": leaf(sc["raw_rotations"]), "opacity": leaf(sc["logits"])}
    colors = np.random.default_rng(1).uniform(0, 1, size=(P, 3))
    override = leaf(colors)
    vm, cp, bg = f32(sc["viewmatrix"]), f32(sc["campos"]), f32(sc["background"])
    dL = f32(sc["dL_dpix"])
    img, vs, vis, radii = mod.render_lonlat(params["xyz"], params["features_dc"], params["features_rest"],
                                            params["scaling"], params["rotation"], params["opacity"], vm, cp, 128,
                                            256, bg, 3, 1.0, variant == "cov3D_precomp", override,
                                            variant == "override_color")
    (img * dL).sum().backward()
    assert img.shape == (3, 128, 256) and radii.dtype == torch.int32 and vis.dtype == torch.bool
    assert torch.equal(vis, radii > 0) and vs.grad is not None and vs.grad.shape == (P, 3)

    # the CPU oracle on the same (FP32-rounded) activations
    s64 = activate64(sc)
    osc = {k: v for k, v in s64.items() if k not in ("log_scales", "raw_rotations", "logits")}
    if variant == "override_color":
        osc.pop("shs")
        osc["colors_precomp"] = colors
    if variant == "cov3D_precomp":
        q, s = osc.pop("rotations"), osc.pop("scales")
        w, x, y, z = q.T
        Rm = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                       np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                       np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
        Sg = np.einsum("pij,pj,pkj->pik", Rm, s ** 2, Rm)
        osc["cov3D_precomp"] = np.stack([Sg[:, 0, 0], Sg[:, 0, 1], Sg[:, 0, 2], Sg[:, 1, 1], Sg[:, 1, 2],
                                         Sg[:, 2, 2]], -1)
    o = Oracle(osc, 32).forward().backward(sc["dL_dpix"])
    d = np.abs(img.detach().cpu().numpy().reshape(-1) - o.get("out_color"))
    assert np.mean(d > IMG_ATOL) <= FLIP_PIXEL_FRACTION and d.max() <= FLIP_MAX_ABS, (d.max(), np.mean(d > IMG_ATOL))
    r_o = o.get("radii")
    assert np.mean(radii.cpu().numpy() != r_o) <= 1e-3
    f = _frac_outside(vs.grad.cpu().numpy().reshape(-1), o.get("dL_dmean2D"), GRAD_RTOL, GRAD_SCALE_ATOL)
    assert f <= FLIP_GRAD_FRACTION, ("dL_dmeans2D", f)
    f = _frac_outside(params["xyz"].grad.cpu().numpy().reshape(-1), o.get("dL_dmeans3D"), GRAD_RTOL,
                      GRAD_SCALE_ATOL)
    assert f <= FLIP_GRAD_FRACTION, ("dL_dmeans3D", f)
    if variant == "sh":
        g = {k: o.get(k) for k in ("dL_dmeans3D", "dL_dscales", "dL_drot", "dL_dopacity", "dL_dsh")}
        ch = chain_activations(sc, g)
        for mine, ref in (("scaling", "log_scales"), ("rotation", "raw_rotations"), ("opacity", "logits")):
            f = _frac_outside(params[mine].grad.cpu().numpy().reshape(-1), ch[ref].reshape(-1), GRAD_RTOL,
                              GRAD_SCALE_ATOL)
            assert f <= FLIP_GRAD_FRACTION, (mine, f)
        sh = torch.cat([params["features_dc"].grad, params["features_rest"].grad], 1).cpu().numpy().reshape(-1)
        assert _frac_outside(sh, ch["shs"].reshape(-1), GRAD_RTOL, GRAD_SCALE_ATOL) <= FLIP_GRAD_FRACTION
        # and the package's own Python path through the same library gives the same image
        p2 = {k: leaf(v.detach().cpu().numpy()) for k, v in params.items()}
        img2, vs2, _, radii2 = render_lonlat(p2, vm, cp, 128, 256, bg, 3)
        assert torch.equal(img2, img.detach()) and torch.equal(radii2, radii)
    elif variant == "override_color":
        f = _frac_outside(override.grad.cpu().numpy().reshape(-1), o.get("dL_dcolors"), GRAD_RTOL, GRAD_SCALE_ATOL)
        assert f <= FLIP_GRAD_FRACTION, ("dL_dcolors", f)
