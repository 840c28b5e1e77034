"""The workload configurations at their full sizes on the GPU, against the
oracle.

* 1.5M Gaussians at 1920x960 (the room scene, L1 upstream gradient): every
  stage given identical upstream outputs, with the OpenMP oracle -- render
  forward given identical binning (1e-4 abs, threshold-margin exceptions),
  render backward given identical forward and preprocess backward given
  identical inputs (both to the rounding-aware bar of tests/helpers.py).
* 3M Gaussians at 4096x2048: the checkpointed autograd Function equals the
  stored one (to a summation-order bar), holds less HBM from forward to
  backward and peaks below 0.8 of the stored form's peak (its backward bins
  in bands of tile rows);
  image(bg=1) - image(bg=0) = final T; binning bit-exact given the GPU's
  preprocess.
* Batches of 8, 16 and 32 uneven views per launch (tile 16x16, and 8 views at
  32x8 and 8x32), 20k Gaussians, views drawn from {240x120, 480x240,
  512x256} (the 960x480 / 1920x960 / 2048x1024 mix at a quarter scale) with
  poses in a 2 m cube: per view, binning bit-exact given the GPU's
  preprocess, image given identical binning, per-item 2D gradients given
  identical forward; the summed 3D gradients against the per-view oracle
  backwards.  And 1.5M Gaussians x 8 views of {960x480, 1920x960,
  2048x1024}: no overflow, per-view binning bit-exact, determinism, the
  background identity and linearity of the backward.
"""
from __future__ import annotations

import numpy as np
import pytest

from omnigs_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _name(request):
    from tests.gpu_util import announce
    announce(request.node.nodeid)
    yield


def _oracle(sc, prec=32, block=(16, 16)):
    from oracle import Oracle
    from tests.helpers import host_threads
    o = Oracle(sc, prec, block=block)
    o.set_threads(host_threads())
    return o


def _clamp3(clamped):
    return np.stack([clamped & 1, (clamped >> 1) & 1, (clamped >> 2) & 1], 1).astype(np.uint8)


def test_room_scene_full_size_stages(gpu):
    import torch

    from oracle import RUNNING_ERROR
    from tests.gpu_util import gpu_state, np_, record, run_backward, run_forward, to_dev
    from tests.helpers import host_threads, running_error_oracle
    from tests.test_gpu_parity import _check_image_with_margin, _rounding_grad_check
    sc = scenes.room_scene()
    P, W, H = sc["means3D"].shape[0], sc["width"], sc["height"]
    ds = to_dev(sc)
    R, img, radii, bufs = run_forward(ds)
    st = gpu_state(bufs, P, R, W, H)
    target = torch.from_numpy(sc["target"]).to(gpu)
    dL = (torch.sign(img - target) / (3.0 * H * W)).contiguous()
    g = run_backward(ds, R, radii, bufs, dL)
    torch.cuda.synchronize()
    dLn = np_(dL)
    osc = {k: v for k, v in sc.items() if k != "target"}
    o = _oracle(osc).run("preprocess")
    for k, v in (("xy", st["means2D"]), ("conic", st["conic_opacity"]), ("rgb", st["rgb"]),
                 ("point_list", st["point_list"]), ("ranges", st["ranges"])):
        o.set(k, v)
    o.run("render")
    _check_image_with_margin(np_(img), st["final_T"], st["n_contrib"], o, W * H)
    m = {"R": int(R)}
    # render backward given identical forward
    o.set("final_T", st["final_T"])
    o.set("n_contrib", st["n_contrib"])
    o.run("render_backward", dLn)
    from oracle import Oracle
    oe = Oracle(osc, RUNNING_ERROR)
    oe.set_threads(host_threads())
    for k in ("xy", "conic", "rgb", "point_list", "ranges", "final_T", "n_contrib"):
        oe.set(k, o.get(k))
    oe.run("render_backward", dLn)
    fragile = oe.get("fragile").astype(bool)
    m.update(fragile_gaussians=int(fragile.sum()), fragile_pixels=oe.fragile_pixels)
    for mine, ref, w in (("dL_dmeans2D", "dL_dmean2D", 3), ("dL_dconic", "dL_dconic", 4),
                         ("dL_dopacity", "dL_dopacity", 1), ("dL_dcolors", "dL_dcolors", 3)):
        _rounding_grad_check(ref, np_(g[mine]).reshape(-1), o.get(ref), oe, w, None, m)
    del oe
    # preprocess backward given identical inputs
    g2 = {"dL_dmean2D": np_(g["dL_dmeans2D"]), "dL_dconic": np_(g["dL_dconic"]),
          "dL_dcolors": np_(g["dL_dcolors"]), "dL_dopacity": np_(g["dL_dopacity"])}
    o.set("radii", np_(radii))
    o.set("clamped", _clamp3(st["clamped"]))
    for k, v in g2.items():
        o.set(k, v)
    o.run("cov2d_backward").run("preprocess_backward")
    oe2 = running_error_oracle(osc, {"radii": np_(radii), "clamped": _clamp3(st["clamped"])}, grads2d=g2,
                               threads=host_threads())
    for mine, ref in (("dL_dcov3D", "dL_dcov3D"), ("dL_dmeans3D", "dL_dmeans3D"), ("dL_dsh", "dL_dsh"),
                      ("dL_dscales", "dL_dscales"), ("dL_drotations", "dL_drot")):
        _rounding_grad_check(ref, np_(g[mine]).reshape(-1), o.get(ref), oe2, metrics=m)
    record("room scene 1.5M 1920x960, stages given identical inputs", **m)


def test_checkpoint_3M_4096x2048(gpu):
    import torch

    from omnigs_amd import rasterizer
    from tests.gpu_util import gpu_state, np_, record, run_forward, to_dev
    P, W, H = 3_000_000, 4096, 2048
    sc = scenes.room_scene(P=P, width=W, height=H)
    rng = np.random.default_rng(5)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(gpu)  # noqa: E731
    raw = {"xyz": sc["means3D"], "features_dc": sc["shs"][:, :1, :], "features_rest": sc["shs"][:, 1:, :],
           "scaling": np.log(sc["scales"]), "rotation": sc["rotations"] * rng.uniform(0.5, 2.0, size=(P, 1)),
           "opacity": np.log(sc["opacities"] / (1.0 - sc["opacities"])).reshape(P, 1)}
    vm, cp, bg, target = t(sc["viewmatrix"]), t(sc["campos"]), t(sc["background"]), t(sc["target"])
    out = {}
    for ck in (False, True):
        params = {k: t(v).requires_grad_() for k, v in raw.items()}
        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated(gpu)
        torch.cuda.reset_peak_memory_stats(gpu)
        img, vs, _, radii = rasterizer.render_lonlat(params, vm, cp, H, W, bg, checkpoint=ck)
        held = torch.cuda.memory_allocated(gpu) - base
        (img - target).abs().mean().backward()
        torch.cuda.synchronize()
        out[ck] = (img.detach(), radii, vs.grad.clone(), {k: v.grad.clone() for k, v in params.items()},
                   held, torch.cuda.max_memory_allocated(gpu) - base)
        del params, img, vs
    assert torch.equal(out[False][0], out[True][0]) and torch.equal(out[False][1], out[True][1])
    # The banded backward adds a Gaussian's per-tile terms band by band, so
    # Gaussians spanning several bands sum them in another order than the
    # stored form's one gather.  Summation order alone moves the stored form
    # by up to 8e-7 max|g| between two runs (the render backward's in-tile
    # LDS sums); the bar allows 1e-5 max|g| per entry and 1e-5 in L2.
    diffs = {}
    for k in out[False][3]:
        a, b = out[False][3][k], out[True][3][k]
        diffs[k] = float((a - b).abs().max() / a.abs().max())
        assert torch.allclose(a, b, rtol=1e-5, atol=1e-5 * float(a.abs().max())), k
        assert float((a - b).norm()) <= 1e-5 * float(a.norm()), k
    assert torch.allclose(out[False][2], out[True][2], rtol=1e-5, atol=1e-5 * float(out[False][2].abs().max()))
    # the checkpointed form holds less between forward and backward, and its
    # peak (now set by autograd's parameter gradients after the rasterizer's
    # backward, its binning running in bands of tile rows) is well below the
    # stored form's
    assert out[True][4] < out[False][4]
    assert out[True][5] < 0.8 * out[False][5], (out[True][5], out[False][5])
    record("checkpointed vs stored 3M 4096x2048", held_stored=out[False][4], held_checkpointed=out[True][4],
           peak_stored=out[False][5], peak_checkpointed=out[True][5],
           grad_maxdiff_over_max={k: round(v, 9) for k, v in diffs.items()})
    # background identity and binning bit-exact given the GPU's preprocess
    ds = to_dev(sc)
    R, img0, radii, bufs = run_forward(ds)
    st = gpu_state(bufs, P, R, W, H)
    ds1 = dict(ds)
    ds1["background"] = torch.ones(3, device=gpu)
    _, img1, _, _ = run_forward(ds1)
    T = torch.from_numpy(st["final_T"]).to(gpu).reshape(H, W)
    assert float((img1 - img0 - T).abs().max()) <= 2.5e-7
    osc = {k: v for k, v in sc.items() if k != "target"}
    o = _oracle(osc).run("preprocess")
    o.set("radii", np_(radii))
    o.set("xy", st["means2D"])
    o.set("depth", st["depth"])
    o.set("tiles_touched", st["tiles_touched"])
    o.run("binning")
    assert o.num_rendered == R
    np.testing.assert_array_equal(st["point_list"], o.get("point_list"))
    np.testing.assert_array_equal(st["ranges"], o.get("ranges"))


QUARTER_SIZES = [(240, 120), (480, 240), (512, 256)]
FULL_SIZES = [(960, 480), (1920, 960), (2048, 1024)]


def _views(V, choice, seed):
    rng = np.random.default_rng(seed)
    sizes, vms, cps = [], [], []
    for _ in range(V):
        sizes.append(tuple(choice[rng.integers(len(choice))]))
        vm, c = scenes.random_pose_camera(rng, 1.0)
        vms.append(vm)
        cps.append(c)
    return sizes, np.stack(vms).astype(np.float32), np.stack(cps).astype(np.float32)


def _bind(gpu, sc, sizes, vms, cps, tile, P):
    import torch

    from omnigs_amd.batch import BatchRenderer
    from tests.gpu_util import to_dev
    ds = to_dev(sc)
    r = BatchRenderer(P, sizes, tile=tile, device=gpu)
    r.bind(viewmatrices=torch.from_numpy(vms).to(gpu), campos=torch.from_numpy(cps).to(gpu),
           means3D=ds["means3D"], opacities=ds["opacities"].reshape(-1).contiguous(), background=ds["background"],
           shs=ds["shs"], scales=ds["scales"], rotations=ds["rotations"])
    r.fit_capacity()
    return r, ds


def _view_items(st, v, P):
    """The GPU's per-item preprocess outputs of view v (Gaussian order)."""
    sl = slice(v * P, (v + 1) * P)
    per = np.empty(st["depth_key"].shape, np.uint32)
    per[st["depth_order"].view(np.uint32)] = st["depth_key"].view(np.uint32)
    return {"radii": st["radii"][sl], "xy": st["means2D"].reshape(-1, 2)[sl], "conic": st["conic_opacity"].reshape(-1, 4)[sl],
            "rgb": st["rgb"].reshape(-1, 3)[sl], "depth": per[sl].view(np.float32),
            "tiles_touched": st["tiles_touched"].view(np.uint32)[sl], "clamped": st["clamped"][sl]}


def _view_binning(st, v, P, sizes, tile, key_shift):
    """View v's sorted list (Gaussian ids) and tile ranges (relative to that
    list) from the batch's arrays: its tiles' keys start at v << key_shift
    (padded keys) or after the earlier views' tiles (compact keys), and each
    key's range holds that (view, tile) group of the point list in depth
    order."""
    ntiles = [((w + tile[0] - 1) // tile[0]) * ((h + tile[1] - 1) // tile[1]) for w, h in sizes]
    kb = v << key_shift if key_shift else int(sum(ntiles[:v]))
    rg = st["ranges"].view(np.uint32).reshape(-1, 2)[kb:kb + ntiles[v]].astype(np.int64)
    cnt = np.maximum(rg[:, 1] - rg[:, 0], 0)
    plist = st["point_list"].view(np.uint32).astype(np.int64)
    pl = np.concatenate([plist[a:a + c] for (a, _), c in zip(rg, cnt) if c > 0] or [np.zeros(0, np.int64)]) - v * P
    start = np.cumsum(cnt) - cnt
    rel = np.where((cnt > 0)[:, None], np.stack([start, start + cnt], axis=1), 0)
    return pl, rel.astype(np.uint32), int(cnt.sum())


@pytest.mark.parametrize("V,tile,full", [(8, (16, 16), False), (16, (16, 16), False), (32, (16, 16), False),
                                         (8, (32, 8), False), (8, (8, 32), False), (16, (16, 16), True)])
def test_batched_views_against_oracle(gpu, V, tile, full):
    """full: 16 views of the full sizes, whose 13-bit tile ids per view sort
    in one pass fewer than the batch's compact 17-bit ids: padded keys."""
    import torch

    from tests.gpu_util import np_, record
    from tests.helpers import host_threads, rounding_check, running_error_oracle
    from tests.test_gpu_parity import _check_image_with_margin, _radii_agree, _rounding_grad_check
    P = 20000
    sc = scenes.small_scene(P=P, seed=11)
    sizes, vms, cps = _views(V, FULL_SIZES if full else QUARTER_SIZES, 200 + V + tile[0])
    r, ds = _bind(gpu, sc, sizes, vms, cps, tile, P)
    rng = np.random.default_rng(V)
    dL = torch.from_numpy(rng.normal(0, 1e-2, size=r.out.numel()).astype(np.float32)).to(gpu)
    g = r.backward(dL)
    torch.cuda.synchronize()
    st = {k: np_(v) for k, v in r.state().items()}
    assert (r.key_shift > 0) == full, r.key_shift
    need, over = r.needed()
    assert not over
    dLn = np_(dL)
    names = (("dL_dmeans3D", "dL_dmeans3D"), ("dL_dsh", "dL_dsh"), ("dL_dscales", "dL_dscales"),
             ("dL_drotations", "dL_drot"), ("dL_dopacity", "dL_dopacity"), ("dL_dcov3D", "dL_dcov3D"))
    sums = {ref: 0.0 for _, ref in names}
    errs = {ref: 0.0 for _, ref in names}
    m = {"views": V, "tile": f"{tile[0]}x{tile[1]}", "instances": need, "key_shift": r.key_shift}
    worst = {}
    for v, (W, H) in enumerate(sizes):
        scv = {k: x for k, x in sc.items() if k != "dL_dpix"}
        scv.update(viewmatrix=vms[v], campos=cps[v], width=W, height=H)
        it = _view_items(st, v, P)
        o = _oracle(scv, block=tile).run("preprocess")
        _radii_agree(it["radii"], o.get("radii"))
        for k in ("radii", "xy", "depth", "tiles_touched"):
            o.set(k, it[k])
        o.run("binning")
        pl, rg, Rv = _view_binning(st, v, P, sizes, tile, r.key_shift)
        assert o.num_rendered == Rv, (v, o.num_rendered, Rv)
        np.testing.assert_array_equal(pl, o.get("point_list"), err_msg=f"view {v}")
        np.testing.assert_array_equal(rg.reshape(-1), o.get("ranges"), err_msg=f"view {v}")
        # render given identical binning
        o.set("conic", it["conic"])
        o.set("rgb", it["rgb"])
        o.run("render")
        p0 = r.pix_offsets[v]
        fT, nc = st["final_T"][p0:p0 + W * H], st["n_contrib"][p0:p0 + W * H].view(np.uint32)
        _check_image_with_margin(np_(r.image(v)), fT, nc, o, W * H)
        # per-item 2D gradients given identical forward
        dLv = dLn[3 * p0:3 * p0 + 3 * W * H].reshape(3, H, W)
        o.set("final_T", fT)
        o.set("n_contrib", nc)
        o.run("render_backward", dLv)
        state = {k: o.get(k) for k in ("xy", "conic", "rgb", "point_list", "ranges", "final_T", "n_contrib")}
        oe = running_error_oracle(scv, dict(state, radii=it["radii"], clamped=_clamp3(it["clamped"])), dLv,
                                  threads=host_threads(), block=tile)
        fragile = oe.get("fragile").astype(bool)
        sl = slice(v * P, (v + 1) * P)
        g2 = {"dL_dmean2D": st["dL_dmean2D"].reshape(-1, 3)[sl], "dL_dconic": st["dL_dconic"].reshape(-1, 4)[sl],
              "dL_dopacity": st["dL_dopacity"][sl], "dL_dcolors": st["dL_dcolor"].reshape(-1, 3)[sl]}
        vm_ = {}
        for ref, w in (("dL_dmean2D", 3), ("dL_dconic", 4), ("dL_dopacity", 1), ("dL_dcolors", 3)):
            _rounding_grad_check(ref, g2[ref].reshape(-1), o.get(ref), oe, w, None, vm_)
        for k, x in vm_.items():
            if k.endswith("worst_ratio"):
                worst[k] = max(worst.get(k, 0.0), x)
        # the view's 3D terms from the GPU's 2D gradients (identical inputs)
        o.set("radii", it["radii"])
        o.set("clamped", _clamp3(it["clamped"]))
        for k, x in g2.items():
            o.set(k, x)
        o.run("cov2d_backward").run("preprocess_backward")
        oe3 = running_error_oracle(scv, {"radii": it["radii"], "clamped": _clamp3(it["clamped"])}, grads2d=g2,
                                   threads=host_threads(), block=tile)
        for _, ref in names:
            val = o.get(ref).astype(np.float64)
            errs[ref] = errs[ref] + oe3.get("err:" + ref) + V * np.abs(val) + np.abs(sums[ref] if np.ndim(sums[ref]) else 0)
            sums[ref] = sums[ref] + val
    for mine, ref in names:
        ok, w, n = rounding_check(np_(g[mine]).reshape(-1), sums[ref], errs[ref])
        assert ok, (mine, w)
        worst[f"{ref}_sum_worst_ratio"] = round(w, 4)
    m.update(worst)
    record(f"batch {V} views {tile[0]}x{tile[1]} 20k Gaussians", **m)


def test_batched_full_size_properties(gpu):
    import torch

    from tests.gpu_util import np_, record
    P = 1_500_000
    sc = scenes.small_scene(P=P, width=64, height=32, seed=0)
    sizes, vms, cps = _views(8, FULL_SIZES, 108)
    r, ds = _bind(gpu, sc, sizes, vms, cps, (16, 16), P)
    torch.cuda.synchronize()
    need, over = r.needed()
    assert not over and need <= r.capacity
    st = {k: np_(v) for k, v in r.state().items()}
    assert need == int(st["tiles_touched"].view(np.uint32).astype(np.int64).sum())
    # per-view binning bit-exact given the GPU's preprocess
    for v, (W, H) in enumerate(sizes):
        scv = {k: x for k, x in sc.items() if k != "dL_dpix"}
        scv.update(viewmatrix=vms[v], campos=cps[v], width=W, height=H)
        it = _view_items(st, v, P)
        o = _oracle(scv).run("preprocess")
        for k in ("radii", "xy", "depth", "tiles_touched"):
            o.set(k, it[k])
        o.run("binning")
        pl, rg, Rv = _view_binning(st, v, P, sizes, (16, 16), r.key_shift)
        assert o.num_rendered == Rv
        np.testing.assert_array_equal(pl, o.get("point_list"), err_msg=f"view {v}")
        np.testing.assert_array_equal(rg.reshape(-1), o.get("ranges"), err_msg=f"view {v}")
        del o
    # determinism
    img0 = r.out.clone()
    r.forward()
    torch.cuda.synchronize()
    assert torch.equal(r.out, img0)
    # linearity of the backward
    dL = torch.from_numpy(np.random.default_rng(3).normal(0, 1e-3, size=r.out.numel()).astype(np.float32)).to(gpu)
    g1 = {k: v.clone() for k, v in r.backward(dL).items() if not k.startswith("_")}
    g2 = r.backward(2.0 * dL)
    torch.cuda.synchronize()
    for k, a in g1.items():
        err = float((2 * a - g2[k]).abs().max()) / max(float(g2[k].abs().max()), 1e-30)
        assert err < 1e-5, (k, err)
    # background identity: image(bg = 1) - image(bg = 0) = final T in every view and channel
    T = torch.from_numpy(st["final_T"]).to(gpu)
    r.bind(viewmatrices=torch.from_numpy(vms).to(gpu), campos=torch.from_numpy(cps).to(gpu),
           means3D=ds["means3D"], opacities=ds["opacities"].reshape(-1).contiguous(),
           background=torch.ones(3, device=gpu), shs=ds["shs"], scales=ds["scales"], rotations=ds["rotations"])
    r.forward()
    torch.cuda.synchronize()
    worst = 0.0
    for v, (W, H) in enumerate(sizes):
        p0 = r.pix_offsets[v]
        o = 3 * p0
        d = r.out[o:o + 3 * W * H].view(3, H, W) - img0[o:o + 3 * W * H].view(3, H, W) - T[p0:p0 + W * H].view(H, W)
        worst = max(worst, float(d.abs().max()))
    assert worst <= 2.5e-7, worst
    record("batch 1.5M Gaussians x 8 full-size views", instances=need, capacity=r.capacity,
           sizes=[f"{w}x{h}" for w, h in sizes], bg_identity_max=worst)
