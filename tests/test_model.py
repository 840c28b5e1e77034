"""Training state around the rasterizer (the reference's GaussianModel):
learning-rate schedule, fused Adam over the six groups, densification
statistics, and densify / clone / split / prune with optimizer-state carry.

Pins: exponLrFunc's closed forms (end points, geometric midpoint, delay ramp);
the six-group step against torch.optim.Adam with the reference's learning
rates (trainingSetup, gaussian_model.cpp:485-518); the statistics kernel
against numpy; densify/prune on a hand-built case whose outcome follows from
gaussian_model.cpp:605-853 (counts, copied rows, split scale log(s / 1.6),
moments kept / zeroed, step preserved).
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from omnigs_amd import capi
from omnigs_amd.model import GROUPS, GaussianModel, OptimizationParams


def _lr(step, a, b, ds, dm, ms):
    return float(capi.load().omnigs_expon_lr(step, a, b, ds, dm, ms))


def _lr_closed(step, a, b, ds, dm, ms):
    if step < 0 or (a == 0 and b == 0):
        return 0.0
    d = dm + (1 - dm) * math.sin(0.5 * math.pi * min(max(step / ds, 0.0), 1.0)) if ds > 0 else 1.0
    t = min(max(step / ms, 0.0), 1.0)
    return d * math.exp(math.log(a) * (1 - t) + math.log(b) * t)


def test_expon_lr_schedule():
    a, b, ms = 1.6e-4, 1.6e-6, 30000
    assert _lr(0, a, b, 0, 0.01, ms) == pytest.approx(a, rel=1e-6)
    assert _lr(ms, a, b, 0, 0.01, ms) == pytest.approx(b, rel=1e-6)
    assert _lr(2 * ms, a, b, 0, 0.01, ms) == pytest.approx(b, rel=1e-6)
    assert _lr(ms // 2, a, b, 0, 0.01, ms) == pytest.approx(math.sqrt(a * b), rel=1e-5)
    assert _lr(-1, a, b, 0, 0.01, ms) == 0.0 and _lr(5, 0.0, 0.0, 0, 0.01, ms) == 0.0
    assert _lr(0, a, b, 100, 0.01, ms) == pytest.approx(0.01 * a, rel=1e-6)
    for step in (0, 17, 100, 250, 12345, 30000):
        for ds in (0, 100):
            assert _lr(step, a, b, ds, 0.01, ms) == pytest.approx(_lr_closed(step, a, b, ds, 0.01, ms), rel=2e-6)


def _toy_model():
    """Six Gaussians: 0,1 small + high grad (clone), 2,3 big + high grad
    (split), 4 low opacity (pruned), 5 ordinary."""
    P = 6
    g = torch.Generator().manual_seed(0)
    logit = lambda p: math.log(p / (1 - p))  # noqa: E731
    params = {
        "xyz": torch.randn(P, 3, generator=g),
        "features_dc": torch.randn(P, 1, 3, generator=g),
        "features_rest": torch.randn(P, 15, 3, generator=g),
        "scaling": torch.log(torch.tensor([[0.001] * 3, [0.002] * 3, [0.5, 0.1, 0.1], [0.2, 0.3, 0.4], [0.001] * 3,
                                           [0.005] * 3])),
        "rotation": torch.randn(P, 4, generator=g),
        "opacity": torch.tensor([[logit(0.5)], [logit(0.6)], [logit(0.7)], [logit(0.8)], [logit(0.001)],
                                 [logit(0.9)]]),
    }
    m = GaussianModel(params, 3, 1.0, device="cpu")
    m.training_setup(OptimizationParams())
    m.xyz_gradient_accum = torch.tensor([[1.0], [2.0], [3.0], [4.0], [0.0], [0.0]])
    m.denom = torch.tensor([[1.0], [1.0], [1.0], [1.0], [0.0], [1.0]])  # row 4: 0/0 -> nan -> 0
    for k in GROUPS:
        st = m.state[k]
        st["exp_avg"] = torch.arange(P, dtype=torch.float32).view(-1, *([1] * (st["exp_avg"].dim() - 1))) \
            .expand_as(st["exp_avg"]).contiguous() + 1
        st["exp_avg_sq"] = st["exp_avg"] * 10
        st["step"] = 7
    return m, params


def test_densify_clone_split_prune():
    m, p0 = _toy_model()
    gen = torch.Generator().manual_seed(5)
    m.densify_and_prune(max_grad=0.5, min_opacity=0.005, extent=1.0, max_screen_size=0, generator=gen)
    # rows: originals 0,1,5 kept; clones of 0,1; 2 children each of 2,3; 4 pruned (low opacity)
    assert m.P == 3 + 2 + 4
    xyz = m.p["xyz"]
    torch.testing.assert_close(xyz[0], p0["xyz"][0], rtol=0, atol=0)
    torch.testing.assert_close(xyz[2], p0["xyz"][5], rtol=0, atol=0)
    torch.testing.assert_close(xyz[3:5], p0["xyz"][0:2], rtol=0, atol=0)  # clones keep position
    for k in ("features_dc", "features_rest", "opacity", "rotation"):
        torch.testing.assert_close(m.p[k][5:9], torch.cat([p0[k][2:4], p0[k][2:4]]), rtol=0, atol=0)
    torch.testing.assert_close(m.p["scaling"][5:9], torch.cat([p0["scaling"][2:4]] * 2) - math.log(1.6),
                               rtol=1e-6, atol=1e-6)
    parent = torch.cat([p0["xyz"][2:4]] * 2)
    sigma = torch.exp(torch.cat([p0["scaling"][2:4]] * 2)).max(1).values
    assert torch.all((xyz[5:9] - parent).norm(dim=1) < 6 * sigma)
    # optimizer state: originals keep their moments (rows 0,1,5 had 1,2,6), new rows start at 0, step kept
    for k in GROUPS:
        st = m.state[k]
        ea = st["exp_avg"].reshape(m.P, -1)[:, 0]
        assert ea[:3].tolist() == [1.0, 2.0, 6.0] and torch.all(ea[3:] == 0)
        assert st["step"] == 7 and st["exp_avg"].shape == m.p[k].shape
    assert m.xyz_gradient_accum.shape == (m.P, 1) and torch.all(m.max_radii2D == 0)


def test_reset_opacity_and_ply_round_trip(tmp_path):
    m, _ = _toy_model()
    before = m.get_opacity().clone()
    m.reset_opacity()
    # the reference clamps against a tensor of ones: opacities keep their values
    assert torch.allclose(m.get_opacity(), before, rtol=1e-5, atol=1e-7)
    assert float(before.max()) > 0.0101  # so a clamp at 0.01 would have shown
    assert torch.all(m.state["opacity"]["exp_avg"] == 0) and m.state["opacity"]["step"] == 7
    m.save_ply(tmp_path / "m.ply")
    m2 = GaussianModel.load_ply(tmp_path / "m.ply", 3)
    for k in GROUPS:
        assert torch.equal(m2.p[k], m.p[k]), k


@pytest.mark.gpu
def test_gpu_densification_stats_kernel(gpu):
    from tests.gpu_util import announce
    announce("densification statistics kernel")
    rng = np.random.default_rng(0)
    P = 100003
    radii = rng.integers(-2, 40, size=P).astype(np.int32)
    d2 = rng.normal(0, 1e-3, size=(P, 3)).astype(np.float32)
    mr = rng.uniform(0, 30, size=P).astype(np.float32)
    acc = rng.uniform(0, 1e-2, size=(P, 1)).astype(np.float32)
    den = rng.integers(0, 5, size=(P, 1)).astype(np.float32)
    params = {k: torch.zeros((P,) + s, device=gpu) for k, s in
              (("xyz", (3,)), ("features_dc", (1, 3)), ("features_rest", (15, 3)), ("scaling", (3,)),
               ("rotation", (4,)), ("opacity", (1,)))}
    m = GaussianModel(params, device=gpu)
    m.max_radii2D = torch.from_numpy(mr.copy()).to(gpu)
    m.xyz_gradient_accum = torch.from_numpy(acc.copy()).to(gpu)
    m.denom = torch.from_numpy(den.copy()).to(gpu)
    m.add_densification_stats(torch.from_numpy(d2).to(gpu), torch.from_numpy(radii).to(gpu))
    torch.cuda.synchronize()
    vis = radii > 0
    np.testing.assert_array_equal(m.max_radii2D.cpu().numpy(), np.where(vis, np.maximum(mr, radii), mr))
    np.testing.assert_array_equal(m.denom.cpu().numpy()[:, 0], den[:, 0] + vis)
    ref = acc[:, 0] + np.where(vis, np.sqrt(d2[:, 0].astype(np.float64) ** 2 + d2[:, 1] ** 2), 0.0)
    np.testing.assert_allclose(m.xyz_gradient_accum.cpu().numpy()[:, 0], ref, rtol=1e-6, atol=1e-12)


@pytest.mark.gpu
def test_gpu_six_group_step_matches_torch_adam(gpu):
    from tests.gpu_util import announce
    announce("six-group fused Adam vs torch.optim.Adam")
    g = torch.Generator().manual_seed(1)
    P = 1000
    shapes = {"xyz": (3,), "features_dc": (1, 3), "features_rest": (15, 3), "opacity": (1,), "scaling": (3,),
              "rotation": (4,)}
    params = {k: torch.randn((P,) + s, generator=g) for k, s in shapes.items()}
    m = GaussianModel({k: v.to(gpu) for k, v in params.items()}, spatial_lr_scale=2.5, device=gpu)
    opt = OptimizationParams()
    m.training_setup(opt)
    ref = {k: torch.nn.Parameter(v.clone().double()) for k, v in params.items()}
    lrs = [opt.position_lr_init * 2.5, opt.feature_lr, opt.feature_lr / 20.0, opt.opacity_lr, opt.scaling_lr,
           opt.rotation_lr]
    topt = torch.optim.Adam([{"params": [ref[k]], "lr": lr} for k, lr in zip(GROUPS, lrs)], eps=1e-15)
    for _ in range(3):
        grads = {k: torch.randn((P,) + s, generator=g) for k, s in shapes.items()}
        for k in GROUPS:
            ref[k].grad = grads[k].double()
        topt.step()
        m.step({k: v.to(gpu) for k, v in grads.items()})
    torch.cuda.synchronize()
    for k in GROUPS:
        np.testing.assert_allclose(m.p[k].cpu().numpy(), ref[k].detach().numpy(), rtol=1e-6, atol=1e-7, err_msg=k)
