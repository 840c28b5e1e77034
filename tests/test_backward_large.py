"""Render backward with Gaussians that cover many tiles, in clamp and seam-wrap
mode, against the oracle given identical forward state.

The per-Gaussian gradients are summed from per-(tile, Gaussian) records: a
Gaussian with up to 16 instances is summed by one lane, a larger one by a
whole wave striding over its records, and in wrap mode a Gaussian's rect can
straddle the longitude seam (its tile columns read modulo the grid width).
The small scenes of the other parity tests rarely reach either case; this
scene puts about a hundred Gaussians of several hundred tiles into a 512x256
view, a third of them on the seam.
"""
from __future__ import annotations

import numpy as np
import pytest

from omnigs_amd import scenes


def _large_scene(wrap: bool, P_small=3000, P_big=100, W=512, H=256, seed=11):
    rng = np.random.default_rng(seed)
    means_small = scenes.uniform_directions(rng, P_small) * rng.uniform(1.5, 5.0, size=(P_small, 1))
    lon = np.where(rng.uniform(size=P_big) < 0.35, np.pi + rng.normal(0.0, 0.05, size=P_big),
                   rng.uniform(-np.pi, np.pi, size=P_big))
    lat = rng.uniform(-0.8, 0.8, size=P_big)
    rho = rng.uniform(2.0, 4.0, size=(P_big, 1))
    means_big = rho * np.stack([np.cos(lat) * np.sin(lon), np.sin(lat), np.cos(lat) * np.cos(lon)], 1)
    means = np.concatenate([means_small, means_big], 0).astype(np.float32)
    sc = scenes._gaussians(rng, means)
    ls = sc["log_scales"].copy()
    ls[P_small:] = rng.normal(-0.6, 0.3, size=(P_big, 3))  # ~0.5 m at 2-4 m: tens of tiles across
    sc["log_scales"] = ls.astype(np.float32)
    sc = scenes.activate(sc)
    vm, c = scenes.camera_at(np.zeros(3))
    sc.update(width=W, height=H, viewmatrix=vm, campos=c, background=np.array([0.1, 0.2, 0.3], dtype=np.float32),
              seam_wrap=wrap)
    sc["dL_dpix"] = rng.normal(0.0, 1e-2, size=(3, H, W)).astype(np.float32)
    return sc


@pytest.mark.gpu
@pytest.mark.parametrize("wrap", [False, True])
def test_render_backward_large_gaussians(gpu, wrap):
    import torch
    from omnigs_amd import capi
    from oracle import Oracle
    from tests.gpu_util import announce, gpu_state, np_, to_dev
    from tests.helpers import allclose_grad
    sc = _large_scene(wrap)
    P, W, H = sc["means3D"].shape[0], sc["width"], sc["height"]
    announce(f"render backward, large Gaussians, wrap={wrap}: {P} Gaussians {W}x{H}")
    ds = to_dev(sc)
    R, img, radii, bufs = capi.forward(background=ds["background"], means3D=ds["means3D"],
                                       opacities=ds["opacities"].reshape(-1).contiguous(), viewmatrix=ds["viewmatrix"],
                                       campos=ds["campos"], width=W, height=H, shs=ds["shs"], scales=ds["scales"],
                                       rotations=ds["rotations"], seam_wrap=wrap)
    st = gpu_state(bufs, P, R, W, H)
    tt = st["tiles_touched"].astype(np.int64)
    assert (tt > 64).sum() >= 20, "the scene must hold Gaussians summed by the whole-wave path"
    assert ((tt > 1) & (tt <= 16)).sum() >= 500, "and many summed by one lane"
    g = capi.backward(R=R, background=ds["background"], means3D=ds["means3D"], radii=radii,
                      viewmatrix=ds["viewmatrix"], campos=ds["campos"], dL_dpix=torch.from_numpy(sc["dL_dpix"]).to(gpu),
                      buffers=bufs, shs=ds["shs"], scales=ds["scales"], rotations=ds["rotations"], seam_wrap=wrap)
    o = Oracle(dict(sc), 32).run("preprocess")
    for k, v in (("xy", st["means2D"]), ("conic", st["conic_opacity"]), ("rgb", st["rgb"]),
                 ("point_list", st["point_list"]), ("ranges", st["ranges"]), ("final_T", st["final_T"]),
                 ("n_contrib", st["n_contrib"])):
        o.set(k, v)
    o.run("render_backward", sc["dL_dpix"])
    for mine, ref in (("dL_dmeans2D", "dL_dmean2D"), ("dL_dconic", "dL_dconic"), ("dL_dopacity", "dL_dopacity"),
                      ("dL_dcolors", "dL_dcolors")):
        ok, worst = allclose_grad(np_(g[mine]).reshape(-1), o.get(ref))
        assert ok, (mine, worst)
    # the big Gaussians carry gradient in every term
    big = np.nonzero(tt > 64)[0]
    assert np.abs(np_(g["dL_dopacity"]).reshape(-1)[big]).max() > 0
