"""Parity of the gfx950 HIP path (called through the C ABI) with the CPU oracle.

Bars:
* integer work (radii, tile counts, instance lists, tile ranges, n_contrib):
  bit-exact given identical upstream outputs;
* image / final_T: |gpu - oracle| <= 1e-4 absolute;
* gradients given identical inputs: |gpu - o32| <= 1e-3 |o32| + 4 u e per
  element, e being the running-error oracle's bound of that entry's FP32
  rounding error (tests/helpers.py), which also carries the possible effect
  of a backward decision within rounding of its threshold (the pair's own
  terms, and the recovered T of the instances in front of it);
* the GPU's error against the FP64 oracle is recorded (gpu_util.record).
"""
from __future__ import annotations

import numpy as np
import pytest

from omnigs_amd import scenes
from tests.helpers import allclose_grad

pytestmark = pytest.mark.gpu

IMG_ATOL = 1e-4
GRAD_RTOL = 1e-3
GRAD_SCALE_ATOL = 1e-4


@pytest.fixture(autouse=True)
def _name(request):
    from tests.gpu_util import announce
    announce(request.node.nodeid)
    yield


def _oracle(sc, prec=32):
    from oracle import Oracle
    return Oracle(sc, prec)


def _setup(sc):
    from tests.gpu_util import gpu_state, run_forward, to_dev
    ds = to_dev(sc)
    R, img, radii, bufs = run_forward(ds)
    st = gpu_state(bufs, sc["means3D"].shape[0], R, sc["width"], sc["height"])
    return ds, R, img, radii, bufs, st


def _radii_agree(gr, orr, oracle_obj=None):
    mism = np.nonzero(gr != orr)[0]
    assert len(mism) <= max(2, int(1e-4 * len(gr))), f"{len(mism)} radii differ"
    assert np.all(np.abs(gr[mism].astype(np.int64) - orr[mism]) <= 1) or np.all((gr[mism] == 0) | (orr[mism] == 0))
    return mism


def test_c_abi_loads_and_reports(gpu):
    from omnigs_amd import capi
    lib = capi.load()
    assert b"gfx950" in lib.omnigs_version()
    present = capi.mark_visible(1000, gpu)
    assert bool(present.all())


@pytest.mark.parametrize("P,W,H,seed", [(20000, 512, 256, 0), (3000, 200, 90, 1)])
def test_preprocess_parity(gpu, P, W, H, seed):
    from tests.gpu_util import np_
    sc = scenes.small_scene(P=P, width=W, height=H, seed=seed)
    ds, R, img, radii, bufs, st = _setup(sc)
    o = _oracle(sc).run("preprocess")
    gr, orr = np_(radii), o.get("radii")
    mism = _radii_agree(gr, orr)
    vis = (gr > 0) & (orr > 0)
    vis[mism] = False
    np.testing.assert_array_equal(st["tiles_touched"][vis], o.get("tiles_touched")[vis])
    np.testing.assert_allclose(st["means2D"].reshape(-1, 2)[vis], o.get("xy").reshape(-1, 2)[vis], atol=2e-3)
    gc, oc = st["conic_opacity"].reshape(-1, 4)[vis], o.get("conic").reshape(-1, 4)[vis]
    np.testing.assert_allclose(gc, oc, rtol=2e-4, atol=2e-4 * np.abs(oc).max())
    np.testing.assert_allclose(st["rgb"].reshape(-1, 3)[vis], o.get("rgb").reshape(-1, 3)[vis], atol=2e-5)
    np.testing.assert_allclose(st["depth"][vis], o.get("depth")[vis], rtol=1e-6)
    cl = st["clamped"][vis]
    ocl = o.get("clamped").reshape(-1, 3)[vis]
    ocl_bits = ocl[:, 0] | (ocl[:, 1] << 1) | (ocl[:, 2] << 2)
    assert np.mean(cl != ocl_bits) < 1e-4


@pytest.mark.parametrize("P,W,H,seed", [(20000, 512, 256, 0), (3000, 200, 90, 1), (800, 64, 32, 2)])
def test_binning_bit_exact_given_identical_preprocess(gpu, P, W, H, seed):
    sc = scenes.small_scene(P=P, width=W, height=H, seed=seed)
    ds, R, img, radii, bufs, st = _setup(sc)
    from tests.gpu_util import np_
    o = _oracle(sc).run("preprocess")
    o.set("radii", np_(radii))
    o.set("xy", st["means2D"])
    o.set("depth", st["depth"])
    o.set("tiles_touched", st["tiles_touched"])
    o.run("binning")
    assert R == o.num_rendered
    np.testing.assert_array_equal(st["point_list"], o.get("point_list"))
    np.testing.assert_array_equal(st["ranges"], o.get("ranges"))
    np.testing.assert_array_equal(st["tile_keys"], (o.get("keys") >> np.uint64(32)).astype(np.uint32))


@pytest.mark.parametrize("P,W,H,seed", [(20000, 512, 256, 0), (3000, 200, 90, 1)])
def test_render_forward_given_identical_binning(gpu, P, W, H, seed):
    sc = scenes.small_scene(P=P, width=W, height=H, seed=seed)
    sc["background"] = np.array([0.1, 0.2, 0.3], dtype=np.float32)
    ds, R, img, radii, bufs, st = _setup(sc)
    from tests.gpu_util import np_
    o = _oracle(sc).run("preprocess")
    for k, v in (("xy", st["means2D"]), ("conic", st["conic_opacity"]), ("rgb", st["rgb"]),
                 ("point_list", st["point_list"]), ("ranges", st["ranges"])):
        o.set(k, v)
    o.run("render")
    _check_image_with_margin(np_(img), st["final_T"], st["n_contrib"], o, W * H)


# A pixel whose alpha >= 1/255 or T >= 1e-4 decision lies within this relative
# distance of its threshold in the FP32 oracle may differ legitimately between
# two correct FP32 evaluations (the reference's own expf is not correctly
# rounded either); such pixels are the only ones allowed past 1e-4.
DECISION_MARGIN = 1e-4


def _check_image_with_margin(img, final_T, n_contrib, o, N):
    d = np.abs(img.reshape(3, N) - o.get("out_color").reshape(3, N)).max(0)
    dT = np.abs(final_T - o.get("final_T"))
    bad = (d > IMG_ATOL) | (dT > IMG_ATOL) | (n_contrib != o.get("n_contrib"))
    margin = o.get("margin")
    print(f"pixels past 1e-4 or with another n_contrib: {int(bad.sum())}; all near a threshold: "
          f"{bool(np.all(margin[bad] < DECISION_MARGIN))}; max|d| elsewhere {d[~bad].max() if (~bad).any() else 0:.3g}")
    assert np.all(margin[bad] < DECISION_MARGIN), (np.nonzero(bad)[0][:10], margin[bad][:10])
    assert bad.sum() <= max(4, int(1e-4 * N))


def _grad_check(name, g, o_val):
    ok, worst = allclose_grad(g, o_val, GRAD_RTOL, GRAD_SCALE_ATOL)
    assert ok, f"{name}: worst tolerance ratio {worst:.3g}"


def _rounding_grad_check(name, g, o32, oerr, width=1, fragile=None, metrics=None):
    """Rounding-aware bar (tests/helpers.py) on every entry (entries of
    Gaussians flagged in ``fragile``, when given, are held to the looser
    1e-3 |o| + 1e-4 max|o| bar instead)."""
    from tests.gpu_util import fp64_errors
    from tests.helpers import per_gaussian_mask, rounding_check
    skip = per_gaussian_mask(fragile, width) if fragile is not None else None
    ok, worst, n = rounding_check(g, o32, oerr.get("err:" + name), skip=skip)
    assert ok, f"{name}: worst rounding-aware ratio {worst:.3g} over {n} entries"
    if skip is not None and skip.any():
        ok2, worst2 = allclose_grad(np.asarray(g).ravel()[skip], np.asarray(o32).ravel()[skip], GRAD_RTOL,
                                    GRAD_SCALE_ATOL)
        assert ok2, f"{name} (fragile Gaussians): worst ratio {worst2:.3g}"
    if metrics is not None:
        metrics[f"{name}_worst_ratio"] = round(worst, 4)
        metrics.update(fp64_errors(name, g, oerr.get(name)))


@pytest.mark.parametrize("P,W,H,seed", [(20000, 512, 256, 0), (3000, 200, 90, 1)])
def test_render_backward_given_identical_forward(gpu, P, W, H, seed):
    import torch
    sc = scenes.small_scene(P=P, width=W, height=H, seed=seed)
    sc["background"] = np.array([0.1, 0.2, 0.3], dtype=np.float32)
    ds, R, img, radii, bufs, st = _setup(sc)
    from tests.gpu_util import np_, run_backward
    dL = torch.from_numpy(sc["dL_dpix"]).to(gpu)
    g = run_backward(ds, R, radii, bufs, dL)
    o = _oracle(sc).run("preprocess")
    for k, v in (("xy", st["means2D"]), ("conic", st["conic_opacity"]), ("rgb", st["rgb"]),
                 ("point_list", st["point_list"]), ("ranges", st["ranges"]), ("final_T", st["final_T"]),
                 ("n_contrib", st["n_contrib"])):
        o.set(k, v)
    o.run("render_backward", sc["dL_dpix"])
    from oracle import RUNNING_ERROR, Oracle
    from tests.gpu_util import record
    oe = Oracle(sc, RUNNING_ERROR)
    for k in ("xy", "conic", "rgb", "point_list", "ranges", "final_T", "n_contrib"):
        oe.set(k, o.get(k))
    oe.run("render_backward", sc["dL_dpix"])
    fragile = oe.get("fragile").astype(bool)
    m = {"fragile_gaussians": int(fragile.sum()), "fragile_pixels": oe.fragile_pixels}
    for mine, ref, w in (("dL_dmeans2D", "dL_dmean2D", 3), ("dL_dconic", "dL_dconic", 4),
                         ("dL_dopacity", "dL_dopacity", 1), ("dL_dcolors", "dL_dcolors", 3)):
        _rounding_grad_check(ref, np_(g[mine]).reshape(-1), o.get(ref), oe, w, None, m)
    record(f"render_backward_given_identical_forward P={P} {W}x{H}", **m)


@pytest.mark.parametrize("P,W,H,seed,degree", [(20000, 512, 256, 0, 3), (3000, 200, 90, 1, 1)])
def test_preprocess_backward_given_identical_inputs(gpu, P, W, H, seed, degree):
    import torch
    sc = scenes.small_scene(P=P, width=W, height=H, seed=seed)
    sc["degree"] = degree
    ds, R, img, radii, bufs, st = _setup(sc)
    from tests.gpu_util import np_, run_backward
    dL = torch.from_numpy(sc["dL_dpix"]).to(gpu)
    g = run_backward(ds, R, radii, bufs, dL)
    o = _oracle(sc).run("preprocess")
    clamped = st["clamped"]
    o.set("radii", np_(radii))
    o.set("clamped", np.stack([clamped & 1, (clamped >> 1) & 1, (clamped >> 2) & 1], 1).astype(np.uint8))
    o.set("dL_dmean2D", np_(g["dL_dmeans2D"]))
    o.set("dL_dconic", np_(g["dL_dconic"]))
    o.set("dL_dcolors", np_(g["dL_dcolors"]))
    o.set("dL_dopacity", np_(g["dL_dopacity"]))
    o.run("cov2d_backward").run("preprocess_backward")
    from tests.gpu_util import record
    from tests.helpers import running_error_oracle
    clamp3 = np.stack([clamped & 1, (clamped >> 1) & 1, (clamped >> 2) & 1], 1).astype(np.uint8)
    oe = running_error_oracle(sc, {"radii": np_(radii), "clamped": clamp3},
                              grads2d={"dL_dmean2D": np_(g["dL_dmeans2D"]), "dL_dconic": np_(g["dL_dconic"]),
                                       "dL_dcolors": np_(g["dL_dcolors"]), "dL_dopacity": np_(g["dL_dopacity"])})
    m = {}
    for mine, ref in (("dpx_dt", "dpx_dt"), ("dpy_dt", "dpy_dt"), ("dL_dcov3D", "dL_dcov3D"),
                      ("dL_dmeans3D", "dL_dmeans3D"), ("dL_dsh", "dL_dsh"), ("dL_dscales", "dL_dscales"),
                      ("dL_drotations", "dL_drot")):
        _rounding_grad_check(ref, np_(g[mine]).reshape(-1), o.get(ref), oe, metrics=m)
    record(f"preprocess_backward_given_identical_inputs P={P} {W}x{H} degree {degree}", **m)


# End to end (independent preprocess on both sides), ulp-level differences in
# the projection can move a pair across the reference's discrete thresholds
# (alpha >= 1/255, T >= 1e-4, the 3-sigma radius and tile rect); such a flip
# changes a pixel by up to one contribution (~alpha c T, alpha ~ 1/255) and the
# recovered T of every other Gaussian blended there.  The FP32 and FP64
# oracles disagree the same way.  end_to_end_check therefore masks exactly
# the decisions that differ, found with the oracle's blend signatures: the
# oracle run on the GPU's preprocess outputs blends the pairs the GPU blends
# (the stage tests show its render equals the GPU's), so a Gaussian or pixel
# whose signature differs from the oracle's own took another decision.
# Pixels whose decision margin is below DECISION_MARGIN on either side are
# masked too (exp implementations differ there).  Every pixel past
# 1e-4 + 8 |o32 - o64| must be masked (the second term covers pixels near the
# poles, where the projection is ill-conditioned and the FP32 oracle itself
# strays).  Gaussians blended in a masked pixel, and those whose signature
# differs, are masked (at most 2%); every other gradient entry is held to
#     |gpu - o32| <= 1e-3 |o32| + 4 u e,
# e being the running-error bound of the backward with the forward state's
# measured deviations (GPU and FP32 oracle from FP64) as input uncertainty.
# The old fraction-based constants remain for callers outside this file.
FLIP_PIXEL_FRACTION = 2e-4
FLIP_MAX_ABS = 0.05
FLIP_GRAD_FRACTION = 2e-3
E2E_FWD_K = 8.0




def _gpu_fwd(st):
    return {"xy": st["means2D"], "conic": st["conic_opacity"], "rgb": st["rgb"], "final_T": st["final_T"]}


def oracle_on_gpu_preprocess(sc, st, radii, tile=(16, 16)):
    """The FP32 oracle's binning and render run on the GPU's preprocess
    outputs: it takes the GPU's discrete decisions (up to exp rounding at
    DECISION_MARGIN), so comparing its blend signatures with the oracle's own
    finds the pairs whose decisions differ."""
    o = _oracle(sc, 32) if tile == (16, 16) else __import__("oracle").Oracle(sc, 32, block=tile)
    o.run("preprocess")
    for k, v in (("radii", radii), ("xy", st["means2D"]), ("depth", st["depth"]),
                 ("tiles_touched", st["tiles_touched"]), ("conic", st["conic_opacity"]), ("rgb", st["rgb"])):
        o.set(k, v)
    return o.run("binning").run("render")


def end_to_end_check(case, sc, img, radii, g, o32, o64, names, o_gpu, gpu_state, tile=(16, 16), max_masked=0.02):
    """Decision-masked end-to-end parity (see above).  g: GPU gradients by the
    oracle's names (numpy); img [3,H,W]; radii [P]; o_gpu: the oracle on the
    GPU's preprocess outputs; gpu_state: the GPU's xy, conic, rgb, final_T.
    Records the measured errors against FP32 and FP64 and the masked counts."""
    from tests.gpu_util import fp64_errors, record
    from tests.helpers import per_gaussian_mask, rounding_check, U32, ROUNDING_K
    from oracle import RUNNING_ERROR, Oracle
    W, H = int(sc["width"]), int(sc["height"])
    P = int(np.asarray(sc["means3D"]).shape[0])
    bx, by = tile
    gx = (W + bx - 1) // bx
    N = W * H
    # radii agree except at the ceil boundary
    rmis = radii != o32.get("radii")
    assert rmis.mean() <= 1e-3, (case, int(rmis.sum()))
    # decisions that differ: blend signatures, and pairs within rounding of a threshold
    g_mask = o_gpu.get("gsig") != o32.get("gsig")
    pix_mask = (o_gpu.get("psig") != o32.get("psig")) | (o32.get("margin") < DECISION_MARGIN) | \
        (o_gpu.get("margin") < DECISION_MARGIN)
    d = np.abs(img.reshape(3, N) - o32.get("out_color").reshape(3, N)).max(0)
    # where the projection is ill-conditioned (near the poles) the FP32 oracle
    # itself strays from FP64; a pixel is allowed that much (times 8) on top of 1e-4
    d_ref = np.abs(o32.get("out_color").reshape(3, N).astype(np.float64) - o64.get("out_color").reshape(3, N)).max(0)
    pix_bad = d > IMG_ATOL + E2E_FWD_K * d_ref
    unexplained = pix_bad & ~pix_mask
    assert not unexplained.any(), (case, np.nonzero(unexplained)[0][:10], d[unexplained][:10])
    assert d.max() <= FLIP_MAX_ABS, (case, float(d.max()))
    # Gaussians blended in a masked pixel, on either side
    taint = g_mask.copy()
    for oo in (o32, o_gpu):
        pl, rg, nc = oo.get("point_list"), oo.get("ranges").reshape(-1, 2), oo.get("n_contrib")
        xy_o, co_o = oo.get("xy").reshape(-1, 2).astype(np.float64), oo.get("conic").reshape(-1, 4).astype(np.float64)
        for pid in np.nonzero(pix_mask)[0]:
            py, px = divmod(int(pid), W)
            t = (py // by) * gx + px // bx
            b, e = rg[t]
            ids = pl[b:min(e, b + int(nc[pid]) + 1)]
            dx, dy = xy_o[ids, 0] - px, xy_o[ids, 1] - py
            if sc.get("seam_wrap"):
                dx = dx - W * np.round(dx / W)
            c = co_o[ids]
            power = -0.5 * (c[:, 0] * dx * dx + c[:, 2] * dy * dy) - c[:, 1] * dx * dy
            alpha = np.minimum(0.99, c[:, 3] * np.exp(np.minimum(power, 0.0)))
            taint[ids[alpha >= (1.0 / 255.0) * (1.0 - 1e-3)]] = True
    assert taint.mean() <= max_masked, (case, taint.mean())
    # running-error bound of the backward on the FP64 forward state, each
    # continuous forward quantity given as input uncertainty the larger of the
    # GPU's and the FP32 oracle's measured deviation from FP64 (so the bound
    # carries sum |d grad / d x| |dx| of the actual differences, to first order)
    oe = Oracle(sc, RUNNING_ERROR)
    if sc.get("scales") is not None and sc.get("cov3D_precomp") is None:
        oe.run("cov3d")
    for k in ("radii", "point_list", "ranges", "n_contrib", "clamped"):
        oe.set(k, o32.get(k))
    for k in ("xy", "conic", "rgb", "final_T"):
        v64 = o64.get(k).astype(np.float64)
        oe.set(k, v64)
        dev = np.maximum(np.abs(np.asarray(gpu_state[k], np.float64).ravel() - v64),
                         np.abs(o32.get(k).astype(np.float64) - v64))
        oe.set("err:" + k, dev / U32)
    oe.run("render_backward", sc["dL_dpix"]).run("cov2d_backward").run("preprocess_backward")
    m = {"masked_gaussians": int(taint.sum()), "signature_gaussians": int(g_mask.sum()),
         "masked_pixels": int(pix_mask.sum()),
         "pixels_past_1e-4": int(pix_bad.sum()), "image_abs_vs_fp32": float(d.max()),
         "image_abs_vs_fp64": float(np.abs(img.reshape(-1) - o64.get("out_color")).max()),
         "image_abs_vs_fp64_unmasked": float(np.abs(img.reshape(3, N) - o64.get("out_color").reshape(3, N))
                                             .max(0)[~pix_mask].max())}
    for name in names:
        a = np.asarray(g[name], np.float64).ravel()
        o = o32.get(name).astype(np.float64)
        w = a.size // P
        ok, worst, n = rounding_check(a, o, oe.get("err:" + name), skip=per_gaussian_mask(taint, w))
        assert ok, f"{case} {name}: worst ratio {worst:.3g} over {n} unmasked entries"
        m[f"{name}_worst_ratio"] = round(worst, 4)
        keep = ~per_gaussian_mask(taint, w)
        m.update(fp64_errors(name, a[keep], o64.get(name)[keep]))
    record(case, **m)
    return m


def _frac_outside(a, b, rtol, scale_atol):
    a = np.asarray(a, dtype=np.float64).ravel(); b = np.asarray(b, dtype=np.float64).ravel()
    tol = rtol * np.abs(b) + scale_atol * max(np.abs(b).max(), 1e-30)
    return float(np.mean(np.abs(a - b) > tol))


_GRAD_NAMES = (("dL_dmeans3D", "dL_dmeans3D"), ("dL_dsh", "dL_dsh"), ("dL_dscales", "dL_dscales"),
               ("dL_drotations", "dL_drot"), ("dL_dopacity", "dL_dopacity"), ("dL_dmeans2D", "dL_dmean2D"),
               ("dL_dconic", "dL_dconic"), ("dL_dcolors", "dL_dcolors"))


@pytest.mark.parametrize("P,W,H,seed", [(20000, 512, 256, 0), (20000, 512, 256, 1)])
def test_end_to_end_against_fp32_and_fp64(gpu, P, W, H, seed):
    """The 20k-Gaussian 512x256 case, fwd + bwd, all stages on the GPU."""
    import torch
    sc = scenes.small_scene(P=P, width=W, height=H, seed=seed)
    ds, R, img, radii, bufs, st = _setup(sc)
    from tests.gpu_util import np_, run_backward
    g = run_backward(ds, R, radii, bufs, torch.from_numpy(sc["dL_dpix"]).to(gpu))
    o32 = _oracle(sc, 32).forward().backward(sc["dL_dpix"])
    o64 = _oracle(sc, 64).forward().backward(sc["dL_dpix"])
    assert abs(R - o32.num_rendered) <= max(2, 1e-4 * R)
    gg = {ref: np_(g[mine]) for mine, ref in _GRAD_NAMES}
    end_to_end_check(f"end_to_end P={P} {W}x{H} seed {seed}", sc, np_(img), np_(radii), gg, o32, o64,
                     [ref for _, ref in _GRAD_NAMES], oracle_on_gpu_preprocess(sc, st, np_(radii)), _gpu_fwd(st))


def test_backward_repeatable_with_records_rewritten(gpu):
    """Each backward writes every staged (tile, Gaussian) record before the
    per-Gaussian gather reads them, so a backward with another dL/dpix in
    between leaves nothing stale: repeated calls on one forward agree with
    each other and with a backward on a fresh forward.  (The four waves of a
    tile combine their sums with LDS atomics, so agreement is to float
    rounding, not bitwise; a stale record would be off by a whole term.)"""
    import torch
    from tests.gpu_util import announce, run_backward
    sc = scenes.small_scene(P=20000, width=512, height=256, seed=4)
    ds, R, img, radii, bufs, st = _setup(sc)
    dL1 = torch.from_numpy(sc["dL_dpix"]).to(gpu)
    dL2 = torch.flip(dL1, dims=[2]).contiguous()
    announce("repeatable backward: 20000 Gaussians 512x256 seed 4")
    g1 = {k: v.clone() for k, v in run_backward(ds, R, radii, bufs, dL1).items()}
    g2 = {k: v.clone() for k, v in run_backward(ds, R, radii, bufs, dL2).items()}
    g1b = run_backward(ds, R, radii, bufs, dL1)
    ds2, R2, _, radii2, bufs2, _ = _setup(sc)
    g2b = run_backward(ds2, R2, radii2, bufs2, dL2)
    assert R2 == R
    for k in g1:
        for x, y in ((g1[k], g1b[k]), (g2[k], g2b[k])):
            scale = max(float(y.abs().max()), 1e-30)
            assert float((x - y).abs().max()) <= 1e-5 * scale, k
    d12 = float((g1["dL_dmeans2D"] - g2["dL_dmeans2D"]).abs().max())
    assert d12 > 1e-3 * float(g1["dL_dmeans2D"].abs().max())


def test_edge_cases_empty_and_all_culled(gpu):
    import torch
    from tests.gpu_util import np_, run_backward, run_forward, to_dev
    sc = scenes.small_scene(P=10, width=64, height=32, seed=3)
    # P == 0: no work, image untouched (zeros), R = 0
    empty = {k: (v[:0] if isinstance(v, np.ndarray) and v.ndim >= 1 and v.shape[0] == 10 else v)
             for k, v in sc.items()}
    ds = to_dev(empty)
    R, img, radii, bufs = run_forward(ds)
    assert R == 0 and float(img.abs().sum()) == 0.0 and radii.numel() == 0
    # every Gaussian inside the 0.2 m cull radius: R = 0, image = background
    sc2 = dict(sc)
    sc2["means3D"] = (sc["means3D"] / np.linalg.norm(sc["means3D"], axis=1, keepdims=True) * 0.1).astype(np.float32)
    sc2["background"] = np.array([0.25, 0.5, 0.75], dtype=np.float32)
    ds2 = to_dev(sc2)
    R, img, radii, bufs = run_forward(ds2)
    assert R == 0 and int(radii.abs().sum()) == 0
    for c in range(3):
        assert torch.all(img[c] == float(sc2["background"][c]))
    g = run_backward(ds2, R, radii, bufs, torch.ones((3, 32, 64), device=gpu))
    for k, v in g.items():
        assert float(v.abs().sum()) == 0.0, k


@pytest.mark.parametrize("variant", ["colors_precomp", "cov3D_precomp", "degree0", "M9_degree2", "whale",
                                     "seam"])
def test_variants_end_to_end(gpu, variant):
    import torch
    from tests.gpu_util import np_, run_backward
    sc = scenes.small_scene(P=2000, width=256, height=128, seed=21)
    if variant == "colors_precomp":
        sc.pop("shs")
        sc["colors_precomp"] = np.random.default_rng(0).uniform(0, 1, size=(2000, 3)).astype(np.float32)
    elif variant == "cov3D_precomp":
        q, s = sc.pop("rotations").astype(np.float64), sc.pop("scales").astype(np.float64)
        w, x, y, z = q.T
        R_ = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                       np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                       np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)
        Sg = np.einsum("pij,pj,pkj->pik", R_, s ** 2, R_)
        sc["cov3D_precomp"] = np.stack([Sg[:, 0, 0], Sg[:, 0, 1], Sg[:, 0, 2], Sg[:, 1, 1], Sg[:, 1, 2],
                                        Sg[:, 2, 2]], -1).astype(np.float32)
    elif variant == "degree0":
        sc["degree"] = 0
    elif variant == "M9_degree2":
        sc["shs"] = np.ascontiguousarray(sc["shs"][:, :9, :])
        sc["degree"] = 2
    elif variant == "whale":
        sc["means3D"][:5] = np.array([[0.0, 0.0, 0.5], [0.3, 0.1, -0.4], [0.0, 0.45, 0.1], [-0.4, -0.2, 0.0],
                                      [0.25, 0.0, 0.25]], dtype=np.float32)
        sc["scales"][:5] = 0.3
    elif variant == "seam":
        rng = np.random.default_rng(5)
        lon = np.pi + rng.normal(0, 0.03, size=2000)
        lat = rng.uniform(-0.5, 0.5, size=2000)
        rho = rng.uniform(1, 4, size=(2000, 1))
        sc["means3D"] = (rho * np.stack([np.cos(lat) * np.sin(lon), np.sin(lat), np.cos(lat) * np.cos(lon)],
                                        1)).astype(np.float32)
    sc["background"] = np.array([0.1, 0.2, 0.3], dtype=np.float32)
    ds, R, img, radii, bufs, st = _setup(sc)
    g = run_backward(ds, R, radii, bufs, torch.from_numpy(sc["dL_dpix"]).to(gpu))
    o = _oracle(sc, 32).forward().backward(sc["dL_dpix"])
    o64 = _oracle(sc, 64).forward().backward(sc["dL_dpix"])
    assert abs(R - o.num_rendered) <= max(2, 1e-4 * R)
    names = [("dL_dmeans3D", "dL_dmeans3D"), ("dL_dopacity", "dL_dopacity"), ("dL_dcov3D", "dL_dcov3D"),
             ("dL_dmeans2D", "dL_dmean2D"), ("dL_dconic", "dL_dconic"), ("dL_dcolors", "dL_dcolors")]
    if variant != "colors_precomp":
        names.append(("dL_dsh", "dL_dsh"))
    if variant != "cov3D_precomp":
        names += [("dL_dscales", "dL_dscales"), ("dL_drotations", "dL_drot")]
    gg = {ref: np_(g[mine]) for mine, ref in names}
    end_to_end_check(f"variant {variant}", sc, np_(img), np_(radii), gg, o, o64, [ref for _, ref in names],
                     oracle_on_gpu_preprocess(sc, st, np_(radii)), _gpu_fwd(st),
                     max_masked=0.05 if variant == "whale" else 0.02)


def test_full_size_properties(gpu):
    """1.5M Gaussians at 1920x960 (the room scene): the integer path is checked
    bit-exactly against the oracle's preprocess-fed binning, and
    size-independent properties are checked on the GPU outputs."""
    import torch
    from tests.gpu_util import gpu_state, np_, run_backward, run_forward, to_dev
    sc = scenes.room_scene()
    P, W, H = sc["means3D"].shape[0], sc["width"], sc["height"]
    ds = to_dev(sc)
    R, img, radii, bufs = run_forward(ds)
    st = gpu_state(bufs, P, R, W, H)
    tt = st["tiles_touched"].astype(np.int64)
    assert R == int(tt.sum())
    # binning bit-exact given the GPU's own preprocess outputs
    o = _oracle({k: v for k, v in sc.items() if k != "target"}, 32)
    o.set("radii", np_(radii)); o.set("xy", st["means2D"]); o.set("depth", st["depth"])
    o.set("tiles_touched", st["tiles_touched"])
    o.run("binning")
    assert o.num_rendered == R
    np.testing.assert_array_equal(st["point_list"], o.get("point_list"))
    np.testing.assert_array_equal(st["ranges"], o.get("ranges"))
    # determinism: a second forward is bit-identical
    R2, img2, radii2, bufs2 = run_forward(ds)
    assert R2 == R and torch.equal(img, img2) and torch.equal(radii, radii2)
    # background identity: image(bg=1) - image(bg=0) = final_T (within 1 ulp of 1)
    ds1 = dict(ds); ds1["background"] = torch.ones(3, device=gpu)
    _, img1, _, _ = run_forward(ds1)
    T = torch.from_numpy(st["final_T"]).to(gpu).reshape(H, W)
    assert float((img1 - img - T).abs().max()) <= 2.5e-7
    assert float(T.min()) >= 0.9999e-4 and float(T.max()) <= 1.0
    # backward is linear in dL/dpix
    target = torch.from_numpy(sc["target"]).to(gpu)
    dL = torch.sign(img - target) / (3.0 * H * W)
    g1 = run_backward(ds, R, radii, bufs, dL, want_jacobian=False)
    g2 = run_backward(ds, R, radii, bufs, 2.0 * dL, want_jacobian=False)
    for k in ("dL_dmeans3D", "dL_dsh", "dL_dopacity", "dL_dscales", "dL_drotations"):
        a, b = g1[k], g2[k]
        err = float((2 * a - b).abs().max()) / max(float(b.abs().max()), 1e-30)
        assert err < 1e-5, (k, err)
