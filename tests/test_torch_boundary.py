"""The C++ LibTorch boundary (RasterizeGaussiansCUDA / ...BackwardCUDA /
markVisible) and its autograd Functions.

CPU tests: the extension loads and raises the reference's errors (which happen
before any device work).  GPU tests: the boundary returns exactly what the C
ABI returns, in the reference's tuple layout; gradients through the autograd
Function (with the caller's activations) match central finite differences of
the FP64 oracle; the checkpointed Function matches the stored one.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from omnigs_amd import scenes


def _args(P=5, camera_type=3, dev="cpu"):
    f = dict(dtype=torch.float32, device=dev)
    return [torch.zeros(3, **f), torch.zeros(P, 3, **f), torch.empty(0, **f), torch.zeros(P, **f),
            torch.ones(P, 3, **f), torch.zeros(P, 4, **f), 1.0, torch.empty(0, **f), torch.eye(4, **f),
            torch.eye(4, **f), 0.0, 0.0, 32, 64, torch.zeros(P, 16, 3, **f), 3, torch.zeros(3, **f), False,
            camera_type, False]


def test_extension_loads_and_exports():
    from omnigs_amd import rasterizer
    e = rasterizer.ext()
    for name in ("RasterizeGaussiansCUDA", "RasterizeGaussiansBackwardCUDA", "markVisible", "rasterize_gaussians"):
        assert hasattr(e, name)


def test_reference_errors_before_device_work():
    from omnigs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer, RasterizeGaussiansCUDA
    a = _args()
    a[1] = torch.zeros(5, 2)
    with pytest.raises(RuntimeError, match=r"means3D must have dimensions \(num_points, 3\)"):
        RasterizeGaussiansCUDA(*a)
    with pytest.raises(RuntimeError, match="Invalid camera_type"):
        RasterizeGaussiansCUDA(*_args(camera_type=7))
    with pytest.raises(RuntimeError, match="pinhole"):
        RasterizeGaussiansCUDA(*_args(camera_type=1))
    # P == 0: the rasterizer is skipped, outputs are zeros, R = 0 (any camera_type)
    R, img, radii, g, b, i = RasterizeGaussiansCUDA(*_args(P=0, camera_type=7))
    assert R == 0 and img.shape == (3, 32, 64) and float(img.abs().sum()) == 0 and radii.numel() == 0
    s = GaussianRasterizationSettings(32, 64, torch.zeros(3), torch.eye(4), torch.zeros(3))
    r = GaussianRasterizer(s)
    x = torch.zeros(4, 3)
    with pytest.raises(RuntimeError, match="one of either SHs or precomputed colors"):
        r(x, x, torch.zeros(4), scales=x, rotations=torch.zeros(4, 4))
    with pytest.raises(RuntimeError, match="scale/rotation pair or precomputed 3D covariance"):
        r(x, x, torch.zeros(4), colors_precomp=x, scales=x)


# ----------------------------------------------------------------------------- GPU


def _dev_scene(sc, dev):
    return {k: (torch.from_numpy(np.ascontiguousarray(v.astype(np.float32) if v.dtype == np.float64 else v)).to(dev)
                if isinstance(v, np.ndarray) else v) for k, v in sc.items()}


@pytest.mark.gpu
def test_boundary_matches_c_abi(gpu):
    from omnigs_amd import capi
    from omnigs_amd.rasterizer import RasterizeGaussiansBackwardCUDA, RasterizeGaussiansCUDA, markVisible
    from tests.gpu_util import announce
    announce("boundary vs C ABI, 20k @ 512x256")
    sc = scenes.small_scene(P=20000, width=512, height=256, seed=0)
    sc["background"] = np.array([0.1, 0.2, 0.3], dtype=np.float32)
    ds = _dev_scene(sc, gpu)
    op = ds["opacities"]  # [P,1] as the caller passes it
    R, img, radii, gb, bb, ib = RasterizeGaussiansCUDA(ds["background"], ds["means3D"], torch.empty(0, device=gpu),
                                                       op, ds["scales"], ds["rotations"], 1.0,
                                                       torch.empty(0, device=gpu), ds["viewmatrix"].reshape(4, 4),
                                                       ds["viewmatrix"].reshape(4, 4), 0.0, 0.0, 256, 512, ds["shs"],
                                                       3, ds["campos"], False, 3, False)
    R2, img2, radii2, bufs = capi.forward_scene(ds)
    assert R == R2 and torch.equal(img, img2) and torch.equal(radii, radii2)
    dL = torch.from_numpy(sc["dL_dpix"]).to(gpu)
    out = RasterizeGaussiansBackwardCUDA(ds["background"], ds["means3D"], radii, torch.empty(0, device=gpu),
                                         ds["scales"], ds["rotations"], 1.0, torch.empty(0, device=gpu),
                                         ds["viewmatrix"].reshape(4, 4), ds["viewmatrix"].reshape(4, 4), 0.0, 0.0, dL,
                                         ds["shs"], 3, ds["campos"], gb, R, bb, ib, 3)
    names = ("dL_dmeans2D", "dL_dcolors", "dL_dopacity", "dL_dmeans3D", "dL_dcov3D", "dL_dsh", "dL_dscales",
             "dL_drotations")
    shapes = ((20000, 3), (20000, 3), (20000, 1), (20000, 3), (20000, 6), (20000, 16, 3), (20000, 3), (20000, 4))
    g = capi.backward_scene(ds, R2, radii2, bufs, dL)
    for t, name, shp in zip(out, names, shapes):
        assert tuple(t.shape) == shp, name
        ref = g[name].reshape(shp)
        # the four waves of a tile add their sums into LDS in a run-dependent
        # order, so two runs agree to rounding, amplified where an entry is
        # a cancelling sum (rotation gradients): 1e-5 of the tensor's scale
        assert torch.allclose(t, ref, rtol=1e-5, atol=1e-5 * float(ref.abs().max()) + 1e-30), name
    assert bool(markVisible(ds["means3D"], ds["viewmatrix"], ds["viewmatrix"], 3).all())


def _fd_loss64(sc, raw):
    from oracle import Oracle
    from tests.helpers import activate64
    s = dict(sc); s.update(raw)
    s = activate64(s)
    o = Oracle(s, 64).forward()
    return float(np.sum(o.get("out_color").reshape(sc["dL_dpix"].shape) * sc["dL_dpix"])), o.decision_sig


@pytest.mark.gpu
@pytest.mark.parametrize("checkpoint", [False, True])
def test_autograd_gradients_vs_fp64_finite_differences(gpu, checkpoint):
    """2k Gaussians at 256x128: gradients of <dL, image> w.r.t. means,
    log-scales, raw quaternions, opacity logits and SH through the autograd
    Function (torch activations on the GPU) against central finite
    differences (step 1e-4) of the FP64 oracle; steps that flip a discrete
    decision are masked.  Bar: |g - fd| <= 2e-3 |fd| + 2e-4 max|fd| on >= 97%
    of the checked entries (FP32 atomics and FP32 arithmetic vs FP64 FD)."""
    from omnigs_amd.rasterizer import render_lonlat
    from tests.gpu_util import announce
    from tests.helpers import tiny_scene
    announce(f"autograd vs FD, checkpoint={checkpoint}")
    sc = tiny_scene(P=2000, W=256, H=128, seed=12, log_scale=(-3.6, 0.3), radius=(1.5, 4.0))
    P = sc["means3D"].shape[0]
    t = lambda a: torch.tensor(np.asarray(a), dtype=torch.float32, device=gpu).requires_grad_(True)  # noqa: E731
    params = {"xyz": t(sc["means3D"]), "features_dc": t(sc["shs"][:, :1]), "features_rest": t(sc["shs"][:, 1:]),
              "scaling": t(sc["log_scales"]), "rotation": t(sc["raw_rotations"]), "opacity": t(sc["logits"])}
    vm = torch.tensor(sc["viewmatrix"], dtype=torch.float32, device=gpu)
    cp = torch.tensor(sc["campos"], dtype=torch.float32, device=gpu)
    bg = torch.tensor(sc["background"], dtype=torch.float32, device=gpu)
    img, vs, vis, radii = render_lonlat(params, vm, cp, 128, 256, bg, 3, checkpoint=checkpoint)
    loss = (img * torch.tensor(sc["dL_dpix"], dtype=torch.float32, device=gpu)).sum()
    loss.backward()
    assert vs.grad is not None and vs.grad.shape == (P, 3)
    grads = {"means3D": params["xyz"].grad, "log_scales": params["scaling"].grad,
             "raw_rotations": params["rotation"].grad, "logits": params["opacity"].grad,
             "shs": torch.cat([params["features_dc"].grad, params["features_rest"].grad], 1)}
    rng = np.random.default_rng(3)
    _, sig0 = _fd_loss64(sc, {})
    h = 1e-4
    for name, g in grads.items():
        g = g.detach().cpu().numpy().astype(np.float64)
        base = sc[name]
        cand = np.nonzero(np.abs(g).reshape(-1) > 1e-3 * np.abs(g).max())[0]
        idx = rng.choice(cand, min(40, cand.size), replace=False)
        ok = checked = 0
        fds, ans = [], []
        for k in idx:
            e = np.zeros(base.size); e[k] = h
            lp, sp = _fd_loss64(sc, {name: base + e.reshape(base.shape)})
            lm, sm = _fd_loss64(sc, {name: base - e.reshape(base.shape)})
            if sp != sig0 or sm != sig0:
                continue
            fds.append((lp - lm) / (2 * h)); ans.append(g.reshape(-1)[k])
        fds, ans = np.array(fds), np.array(ans)
        checked = len(fds)
        assert checked >= 0.6 * len(idx), (name, checked)
        tol = 2e-3 * np.abs(fds) + 2e-4 * np.abs(fds).max()
        ok = int(np.sum(np.abs(ans - fds) <= tol))
        print(f"{name}: {ok}/{checked} within tol; max rel err {np.max(np.abs(ans - fds) / np.maximum(np.abs(fds), 1e-12)):.3g}")
        assert ok >= 0.97 * checked, (name, ok, checked)


@pytest.mark.gpu
@pytest.mark.parametrize("band_instances", [0, 4000])
def test_checkpointed_equals_stored(gpu, band_instances):
    """The checkpointed Function (backward binned in bands of tile rows: one
    band at the default budget, several at 4000 instances) against the stored
    one."""
    from omnigs_amd.rasterizer import render_lonlat
    from tests.gpu_util import announce
    announce(f"checkpointed (band budget {band_instances}) vs stored, 20k @ 512x256")
    sc = scenes.small_scene(P=20000, width=512, height=256, seed=4)
    out = {}
    for ck in (False, True):
        params = {"xyz": torch.tensor(sc["means3D"], device=gpu, requires_grad=True),
                  "features_dc": torch.tensor(sc["shs"][:, :1], device=gpu, requires_grad=True),
                  "features_rest": torch.tensor(sc["shs"][:, 1:], device=gpu, requires_grad=True),
                  "scaling": torch.tensor(sc["log_scales"], device=gpu, requires_grad=True),
                  "rotation": torch.tensor(sc["raw_rotations"], device=gpu, requires_grad=True),
                  "opacity": torch.tensor(sc["logits"], device=gpu, requires_grad=True)}
        vm = torch.tensor(sc["viewmatrix"], device=gpu)
        img, vs, vis, radii = render_lonlat(params, vm, torch.tensor(sc["campos"], device=gpu), 256, 512,
                                            torch.zeros(3, device=gpu), 3, checkpoint=ck,
                                            band_instances=band_instances)
        (img * torch.tensor(sc["dL_dpix"], device=gpu)).sum().backward()
        out[ck] = (img.detach(), vs.grad.clone(), {k: v.grad.clone() for k, v in params.items()})
    assert torch.equal(out[False][0], out[True][0])
    for k in out[False][2]:
        a, b = out[False][2][k], out[True][2][k]
        assert torch.allclose(a, b, rtol=1e-5, atol=1e-6 * float(a.abs().max())), k
    assert torch.allclose(out[False][1], out[True][1], rtol=1e-5, atol=1e-6 * float(out[False][1].abs().max()))
