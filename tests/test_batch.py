"""Batched multi-view forward / backward (extension): V views of one Gaussian
set per launch, tile shapes 16x16 / 32x8 / 8x32, graph capture.

GPU checks:
  * per-view images bit-identical to V single-view forwards (16x16 tiles);
  * parameter gradients equal the sum of the single-view backwards
    (1e-3 relative + 1e-4 of the largest entry: atomics add in another order);
  * 32x8 and 8x32 tiles against the oracle at the same tile shape (point
    lists bit-exact; images within 1e-4 except pixels whose decisions lie
    within float rounding of a threshold);
  * graph replays equal the eager step, with eager work between replays;
  * an instance capacity below the need reports overflow without faulting;
  * the colour decoder path against single-view decoder runs.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from omnigs_amd import scenes

SIZES = [(512, 256), (256, 128), (384, 192)]


def _poses(V, seed=3):
    rng = np.random.default_rng(seed)
    vms, cps = [], []
    for _ in range(V):
        ang = rng.uniform(-np.pi, np.pi)
        R = np.array([[np.cos(ang), 0, -np.sin(ang)], [0, 1, 0], [np.sin(ang), 0, np.cos(ang)]])
        c = rng.uniform(-0.3, 0.3, size=3)
        vms.append(scenes.viewmatrix_from_pose(R, -R @ c))
        cps.append(c.astype(np.float32))
    return np.stack(vms).astype(np.float32), np.stack(cps).astype(np.float32)


def test_batch_workspace_validation_without_gpu():
    from omnigs_amd import capi
    lib = capi.load()
    w = (ctypes.c_int * 2)(512, 256)
    h = (ctypes.c_int * 2)(256, 128)
    assert lib.omnigs_lonlat_batch_workspace_bytes(1000, 2, w, h, 16, 16, 10000) > 0
    assert lib.omnigs_lonlat_batch_workspace_bytes(1000, 0, w, h, 16, 16, 10000) == 0   # no views
    assert lib.omnigs_lonlat_batch_workspace_bytes(1000, 33, w, h, 16, 16, 10000) == 0  # > 32 views
    assert lib.omnigs_lonlat_batch_workspace_bytes(1000, 2, w, h, 16, 8, 10000) == 0    # unsupported tile
    assert lib.omnigs_lonlat_batch_workspace_bytes(1000, 2, w, h, 32, 8, 10000) > 0
    a = capi.BatchArgs()
    assert lib.omnigs_lonlat_batch_forward(ctypes.byref(a), None) == capi.OMNIGS_ERR_INVALID_ARGUMENT


def _setup(gpu, P=20000, seed=4):
    import torch
    from tests.gpu_util import to_dev
    sc = scenes.small_scene(P=P, seed=seed)
    ds = to_dev(sc)
    vms, cps = _poses(len(SIZES))
    vm_t, cp_t = torch.from_numpy(vms).to(gpu), torch.from_numpy(cps).to(gpu)
    inputs = dict(means3D=ds["means3D"], opacities=ds["opacities"].reshape(-1).contiguous(), background=ds["background"],
                  shs=ds["shs"], scales=ds["scales"], rotations=ds["rotations"])
    return sc, ds, vms, cps, vm_t, cp_t, inputs


def _single(ds, vm, cp, W, H, gpu, dL=None, mlp=None, tile=None):
    from omnigs_amd import capi
    R, img, radii, bufs = capi.forward(background=ds["background"], means3D=ds["means3D"],
                                       opacities=ds["opacities"].reshape(-1).contiguous(), viewmatrix=vm, campos=cp,
                                       width=W, height=H, shs=ds["shs"], scales=ds["scales"],
                                       rotations=ds["rotations"], mlp=mlp)
    g = None
    if dL is not None:
        g = capi.backward(R=R, background=ds["background"], means3D=ds["means3D"], radii=radii, viewmatrix=vm,
                          campos=cp, dL_dpix=dL, buffers=bufs, shs=ds["shs"], scales=ds["scales"],
                          rotations=ds["rotations"], mlp=mlp)
    return img, g


@pytest.mark.gpu
def test_batch_matches_single_views(gpu):
    import torch
    from omnigs_amd.batch import BatchRenderer
    from tests.gpu_util import announce, np_
    from tests.helpers import allclose_grad
    announce("batch of 3 views vs single-view forwards/backwards, 20k Gaussians")
    sc, ds, vms, cps, vm_t, cp_t, inputs = _setup(gpu)
    r = BatchRenderer(20000, SIZES, tile=(16, 16), device=gpu).bind(viewmatrices=vm_t, campos=cp_t, **inputs)
    r.fit_capacity()
    rng = np.random.default_rng(9)
    dL = torch.from_numpy(rng.normal(0, 1e-2, size=r.out.numel()).astype(np.float32)).to(gpu)
    g = r.backward(dL)
    torch.cuda.synchronize()
    sums = {}
    for v, (W, H) in enumerate(SIZES):
        o = 3 * r.pix_offsets[v]
        img, gv = _single(ds, vm_t[v].contiguous(), cp_t[v].contiguous(), W, H, gpu,
                          dL=dL[o:o + 3 * W * H].view(3, H, W).contiguous())
        torch.cuda.synchronize()
        assert torch.equal(r.image(v), img), f"view {v} differs"
        for k in ("dL_dmeans3D", "dL_dopacity", "dL_dsh", "dL_dscales", "dL_drotations", "dL_dcov3D"):
            sums[k] = sums.get(k, 0) + np_(gv[k]).reshape(-1).astype(np.float64)
    for k, ref in sums.items():
        ok, worst = allclose_grad(np_(g[k]).reshape(-1), ref)
        assert ok, (k, worst)


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [(32, 8), (8, 32)])
def test_batch_tile_shapes_match_oracle(gpu, oracle_lib, tile):
    import torch
    from omnigs_amd.batch import BatchRenderer
    from oracle import Oracle
    from tests.gpu_util import announce, np_
    announce(f"batch tile shape {tile[0]}x{tile[1]} vs oracle, 5000 Gaussians at 256x128")
    sc = scenes.small_scene(P=5000, width=256, height=128, seed=7)
    from tests.gpu_util import to_dev
    ds = to_dev(sc)
    vm_t = ds["viewmatrix"].reshape(1, 16).contiguous()
    cp_t = ds["campos"].reshape(1, 3).contiguous()
    r = BatchRenderer(5000, [(256, 128)], tile=tile, device=gpu).bind(
        viewmatrices=vm_t, campos=cp_t, means3D=ds["means3D"], opacities=ds["opacities"], background=ds["background"],
        shs=ds["shs"], scales=ds["scales"], rotations=ds["rotations"])
    r.fit_capacity()
    torch.cuda.synchronize()
    o = Oracle(sc, 32, block=tile).forward()
    need, over = r.needed()
    assert not over and need == o.num_rendered
    img = np_(r.image(0)).reshape(3, -1)
    d = np.abs(img - o.get("out_color").reshape(3, -1)).max(0)
    bad = d > 1e-4
    assert np.all(o.get("margin")[bad] < 1e-4) and bad.sum() <= 4, (bad.sum(), d.max())


@pytest.mark.gpu
def test_batch_graph_replay_matches_eager(gpu):
    import torch
    from omnigs_amd import capi
    from omnigs_amd.batch import BatchRenderer, capture, make_step
    from tests.gpu_util import announce
    announce("batch graph capture: replays vs eager")
    sc, ds, vms, cps, vm_t, cp_t, inputs = _setup(gpu, P=10000)
    r = BatchRenderer(10000, SIZES, tile=(16, 16), device=gpu).bind(viewmatrices=vm_t, campos=cp_t, **inputs)
    r.fit_capacity()
    target = torch.rand(r.out.shape, generator=torch.Generator(device=gpu).manual_seed(1), device=gpu)
    dl = torch.empty_like(r.out)
    step = make_step([r], [target], [capi.PhotometricLoss(0.0)], [dl])
    step()
    torch.cuda.synchronize()
    ref_img = r.out.clone()
    ref_g = {k: v.clone() for k, v in r.grads().items() if not k.startswith("_")}
    graph = capture(step)
    for rep in range(3):
        r.out.zero_()
        for v in r.grads().values():
            v.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(r.out, ref_img), f"replay {rep} image"
        for k, v in ref_g.items():
            assert torch.allclose(r.grads()[k], v, rtol=1e-5, atol=1e-7 * float(v.abs().max()) + 1e-12), (rep, k)
        _ = (torch.randn(1000, device=gpu) * 2).sum().item()  # eager work between replays


@pytest.mark.gpu
def test_batch_overflow_is_reported(gpu):
    import torch
    from omnigs_amd.batch import BatchRenderer
    from tests.gpu_util import announce
    announce("batch capacity overflow flag")
    sc, ds, vms, cps, vm_t, cp_t, inputs = _setup(gpu, P=5000)
    r = BatchRenderer(5000, SIZES, tile=(16, 16), device=gpu, capacity=1000).bind(viewmatrices=vm_t, campos=cp_t,
                                                                                   **inputs)
    r.forward()
    need, over = r.needed()
    assert over and need > 1000
    r.fit_capacity()
    need2, over2 = r.needed()
    assert not over2 and need2 == need and r.capacity >= need


@pytest.mark.gpu
def test_batch_decoder_matches_single_views(gpu):
    import torch
    from omnigs_amd.batch import BatchRenderer
    from omnigs_amd.decoder import ColourDecoder
    from tests.gpu_util import announce, np_
    announce("batch with colour decoder vs single views")
    sc, ds, vms, cps, vm_t, cp_t, inputs = _setup(gpu, P=8000)
    dec = ColourDecoder(8000, gpu, seed=3, latent_std=0.7)
    r = BatchRenderer(8000, SIZES, device=gpu).bind(viewmatrices=vm_t, campos=cp_t, mlp=dec.mlp_args(), **inputs)
    r.fit_capacity()
    dL = torch.from_numpy(np.random.default_rng(2).normal(0, 1e-2, size=r.out.numel()).astype(np.float32)).to(gpu)
    g = r.backward(dL)
    torch.cuda.synchronize()
    lat = 0
    mlpg = 0
    for v, (W, H) in enumerate(SIZES):
        o = 3 * r.pix_offsets[v]
        img, gv = _single(ds, vm_t[v].contiguous(), cp_t[v].contiguous(), W, H, gpu,
                          dL=dL[o:o + 3 * W * H].view(3, H, W).contiguous(), mlp=dec.mlp_args())
        torch.cuda.synchronize()
        assert torch.equal(r.image(v), img), f"view {v} differs"
        lat = lat + np_(gv["dL_dlatent"]).astype(np.float64)
        mlpg = mlpg + np_(gv["dL_dmlp"]).astype(np.float64)
    for name, got, ref in (("dL_dlatent", np_(g["dL_dlatent"]), lat), ("dL_dmlp", np_(g["dL_dmlp"]), mlpg)):
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert rel < 1e-4, (name, rel)
