"""Longitude-seam wrap mode (an extension: the reference clamps tile columns at
the image edge, cuda_rasterizer/auxiliary.h:56-66, forward.cu:680-681).

Property: in wrap mode a camera yaw of 2*pi*k/W rolls the equirectangular
image by exactly k columns.  With k a multiple of the 16-pixel tile width
the tile grid shifts by whole tiles, so every per-pixel decision is the same
and the roll holds to rounding.  Clamp mode (the reference) breaks it at the
seam.
"""
from __future__ import annotations

import numpy as np
import pytest

from omnigs_amd import scenes


def _seam_scene(P=600, W=256, H=128, seed=2):
    rng = np.random.default_rng(seed)
    lon = np.pi + rng.normal(0.0, 0.25, size=P)  # clustered on the seam
    lat = rng.uniform(-0.7, 0.7, size=P)
    rho = rng.uniform(1.5, 4.0, size=(P, 1))
    means = rho * np.stack([np.cos(lat) * np.sin(lon), np.sin(lat), np.cos(lat) * np.cos(lon)], 1)
    sc = scenes._gaussians(rng, means)
    sc["log_scales"] = rng.normal(-3.0, 0.3, size=(P, 3)).astype(np.float32)
    sc = scenes.activate(sc)
    sc.update(width=W, height=H, background=np.array([0.1, 0.2, 0.3], dtype=np.float32))
    return sc


def _yaw_view(theta):
    c, s = np.cos(theta), np.sin(theta)
    Rcw = np.array([[c, 0.0, -s], [0.0, 1.0, 0.0], [s, 0.0, c]])  # lon' = lon - theta: image rolls by -theta W/2pi
    return scenes.viewmatrix_from_pose(Rcw, np.zeros(3)).astype(np.float64), np.zeros(3)


def _render64(sc, theta, wrap):
    from oracle import Oracle
    s = dict(sc)
    s["viewmatrix"], s["campos"] = _yaw_view(theta)
    s["seam_wrap"] = wrap
    s = {k: (v.astype(np.float64) if isinstance(v, np.ndarray) and v.dtype == np.float32 else v) for k, v in s.items()}
    o = Oracle(s, 64).forward()
    return o.get("out_color").reshape(3, sc["height"], sc["width"])


@pytest.mark.parametrize("k", [16, 48, -32])
def test_yaw_rolls_the_image_in_wrap_mode(oracle_lib, k):
    sc = _seam_scene()
    W = sc["width"]
    a = _render64(sc, 0.0, True)
    b = _render64(sc, 2 * np.pi * k / W, True)
    d = np.abs(np.roll(a, -k, axis=2) - b)
    assert d.max() < 1e-6 or np.mean(d > 1e-6) < 2e-4, (d.max(), np.mean(d > 1e-6))


def test_clamp_mode_breaks_at_the_seam(oracle_lib):
    sc = _seam_scene()
    W = sc["width"]
    a = _render64(sc, 0.0, False)
    b = _render64(sc, 2 * np.pi * 64 / W, False)
    d = np.abs(np.roll(a, -64, axis=2) - b)
    assert np.mean(d > 1e-3) > 1e-3  # the reference's clamp cuts Gaussians at lon = +-pi


def test_wrap_mode_binning_brute_force(oracle_lib):
    from oracle import Oracle
    sc = _seam_scene(P=300, W=128, H=64)
    vm, c = _yaw_view(0.0)
    s = dict(sc, viewmatrix=vm.astype(np.float32), campos=c.astype(np.float32), seam_wrap=True)
    o = Oracle(s, 32).run("preprocess").run("binning")
    radii, xy, depth = o.get("radii"), o.get("xy"), o.get("depth")
    gx, gy = 128 // 16, 64 // 16
    items = []
    f = np.float32
    for i in range(len(radii)):
        r = int(radii[i])
        if r <= 0:
            continue
        px, py = f(xy[2 * i]), f(xy[2 * i + 1])
        x0 = int(np.floor((px - f(r)) / f(16))); x1 = int(np.floor((((px + f(r)) + f(16)) - f(1)) / f(16)))
        if x1 - x0 >= gx:
            x0, x1 = 0, gx
        y0 = min(gy, max(0, int((py - f(r)) / f(16)))); y1 = min(gy, max(0, int((((py + f(r)) + f(16)) - f(1)) / f(16))))
        for y in range(y0, y1):
            for x in range(x0, x1):
                items.append((y * gx + (x % gx), int(f(depth[i]).view(np.uint32)), i))
    items.sort()
    np.testing.assert_array_equal(o.get("point_list"), np.array([it[2] for it in items], dtype=np.uint32))
    assert o.num_rendered == len(items)


@pytest.mark.gpu
def test_wrap_mode_gpu_matches_oracle(gpu):
    import torch
    from tests.gpu_util import announce, gpu_state, np_, run_backward, to_dev
    from omnigs_amd import capi
    from oracle import Oracle
    announce("seam wrap: GPU vs oracle")
    sc = _seam_scene(P=4000, W=512, H=256, seed=5)
    vm, c = _yaw_view(0.3)
    sc["viewmatrix"], sc["campos"] = vm.astype(np.float32), c.astype(np.float32)
    sc["dL_dpix"] = np.random.default_rng(1).normal(0, 1e-2, size=(3, 256, 512)).astype(np.float32)
    ds = to_dev(sc)
    R, img, radii, bufs = capi.forward(background=ds["background"], means3D=ds["means3D"],
                                       opacities=ds["opacities"].reshape(-1).contiguous(), viewmatrix=ds["viewmatrix"],
                                       campos=ds["campos"], width=512, height=256, shs=ds["shs"], scales=ds["scales"],
                                       rotations=ds["rotations"], seam_wrap=True)
    st = gpu_state(bufs, 4000, R, 512, 256)
    o = Oracle(dict(sc, seam_wrap=True), 32).run("preprocess")
    o.set("radii", np_(radii)); o.set("xy", st["means2D"]); o.set("depth", st["depth"])
    o.set("tiles_touched", st["tiles_touched"])
    o.run("binning")
    assert o.num_rendered == R
    np.testing.assert_array_equal(st["point_list"], o.get("point_list"))
    np.testing.assert_array_equal(st["ranges"], o.get("ranges"))
    for k, v in (("conic", st["conic_opacity"]), ("rgb", st["rgb"])):
        o.set(k, v)
    o.run("render")
    d = np.abs(np_(img).reshape(3, -1) - o.get("out_color").reshape(3, -1)).max(0)
    bad = d > 1e-4
    assert np.all(o.get("margin")[bad] < 1e-4) and bad.sum() <= 4
    g = capi.backward(R=R, background=ds["background"], means3D=ds["means3D"], radii=radii,
                      viewmatrix=ds["viewmatrix"], campos=ds["campos"], dL_dpix=torch.from_numpy(sc["dL_dpix"]).to(gpu),
                      buffers=bufs, shs=ds["shs"], scales=ds["scales"], rotations=ds["rotations"], seam_wrap=True)
    for k, v in (("final_T", st["final_T"]), ("n_contrib", st["n_contrib"])):
        o.set(k, v)
    o.run("render_backward", sc["dL_dpix"])
    from tests.helpers import allclose_grad
    for mine, ref in (("dL_dmeans2D", "dL_dmean2D"), ("dL_dconic", "dL_dconic"), ("dL_dopacity", "dL_dopacity"),
                      ("dL_dcolors", "dL_dcolors")):
        ok, worst = allclose_grad(np_(g[mine]).reshape(-1), o.get(ref))
        assert ok, (mine, worst)
