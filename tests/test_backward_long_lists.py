"""Render backward on tiles whose lists hold thousands of blended instances,
against the oracle given identical forward state.

The backward stages each batch from a queue of the list positions that some
wave blended in the forward, kept in a ring of 1024 entries in LDS; the ring
wraps once a tile has staged 1024 entries.  The other parity scenes have
tiles of a few hundred blended instances.  Here 4000 Gaussians share one
direction with an opacity just above the 1/255 cut-off, so the pixels at the
centre blend about 2000 of them before their transmittance runs out.
"""
from __future__ import annotations

import numpy as np
import pytest

from omnigs_amd import scenes


def _stacked_scene(P=4000, W=128, H=64, seed=5):
    rng = np.random.default_rng(seed)
    dirs = np.array([0.0, 0.0, 1.0]) + rng.normal(0.0, 0.02, size=(P, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    means = (dirs * rng.uniform(2.5, 3.5, size=(P, 1))).astype(np.float32)
    sc = scenes._gaussians(rng, means)
    sc["log_scales"] = np.log(rng.uniform(0.6, 0.9, size=(P, 3))).astype(np.float32)
    opacity = 0.0047  # alpha = o G >= 1/255 where G >= 0.83
    sc["logits"] = np.full((P, 1), np.log(opacity / (1.0 - opacity)), dtype=np.float32)
    sc = scenes.activate(sc)
    vm, c = scenes.camera_at(np.zeros(3))
    sc.update(width=W, height=H, viewmatrix=vm, campos=c, background=np.array([0.1, 0.2, 0.3], dtype=np.float32),
              seam_wrap=False)
    sc["dL_dpix"] = rng.normal(0.0, 1e-2, size=(3, H, W)).astype(np.float32)
    return sc


@pytest.mark.gpu
def test_render_backward_long_tile_lists(gpu):
    import torch
    from omnigs_amd import capi
    from oracle import Oracle
    from tests.gpu_util import announce, gpu_state, np_, to_dev
    from tests.helpers import allclose_grad
    sc = _stacked_scene()
    P, W, H = sc["means3D"].shape[0], sc["width"], sc["height"]
    announce(f"render backward, long tile lists: {P} stacked Gaussians {W}x{H}")
    ds = to_dev(sc)
    R, img, radii, bufs = capi.forward(background=ds["background"], means3D=ds["means3D"],
                                       opacities=ds["opacities"].reshape(-1).contiguous(), viewmatrix=ds["viewmatrix"],
                                       campos=ds["campos"], width=W, height=H, shs=ds["shs"], scales=ds["scales"],
                                       rotations=ds["rotations"], seam_wrap=False)
    st = gpu_state(bufs, P, R, W, H)
    assert int(st["n_contrib"].max()) > 1500, "some pixel must blend well past the queue's 1024 entries"
    g = capi.backward(R=R, background=ds["background"], means3D=ds["means3D"], radii=radii,
                      viewmatrix=ds["viewmatrix"], campos=ds["campos"], dL_dpix=torch.from_numpy(sc["dL_dpix"]).to(gpu),
                      buffers=bufs, shs=ds["shs"], scales=ds["scales"], rotations=ds["rotations"], seam_wrap=False)
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
