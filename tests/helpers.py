"""Shared test helpers: FD harness around the FP64 oracle, tolerances."""
from __future__ import annotations

import numpy as np

from omnigs_amd import scenes


def tiny_scene(P=24, W=128, H=64, seed=3, bg=(0.3, 0.5, 0.7), log_scale=(-2.6, 0.25), degree=3,
               logit_clip=2.0, radius=(1.6, 3.0)):
    """A few mid-sized Gaussians (footprints of tens of pixels) with a
    non-black background so every backward term is exercised."""
    rng = np.random.default_rng(seed)
    dirs = scenes.uniform_directions(rng, P)
    # keep away from the poles and the seam so projections stay smooth
    dirs[:, 1] *= 0.6
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    means = dirs * rng.uniform(radius[0], radius[1], size=(P, 1))
    raw = {
        "means3D": means.astype(np.float64),
        "log_scales": rng.normal(log_scale[0], log_scale[1], size=(P, 3)),
        "raw_rotations": rng.normal(size=(P, 4)),
        "logits": np.clip(rng.normal(0.0, 1.0, size=(P, 1)), -logit_clip, logit_clip),
        "shs": np.concatenate([rng.uniform(-0.5, 0.5, size=(P, 1, 3)) / 0.2821,
                               rng.normal(0.0, 0.2, size=(P, 15, 3))], 1),
    }
    sc = {"width": W, "height": H, "degree": degree, "scale_modifier": 1.0,
          "background": np.asarray(bg, dtype=np.float64)}
    vm, c = scenes.camera_at(np.zeros(3))
    sc["viewmatrix"] = vm.astype(np.float64)
    sc["campos"] = c.astype(np.float64)
    sc.update(raw)
    sc["dL_dpix"] = rng.normal(0.0, 1.0, size=(3, H, W))
    return sc


def activate64(sc: dict) -> dict:
    out = dict(sc)
    q = sc["raw_rotations"]
    out["rotations"] = q / np.linalg.norm(q, axis=1, keepdims=True)
    out["scales"] = np.exp(sc["log_scales"])
    out["opacities"] = 1.0 / (1.0 + np.exp(-sc["logits"]))
    return out


def chain_activations(sc: dict, g: dict) -> dict:
    """Chains the rasterizer gradients through the caller-side activations
    (exp, normalise, sigmoid; gaussian_model.cpp:54-77)."""
    q = sc["raw_rotations"]
    n = np.linalg.norm(q, axis=1, keepdims=True)
    qh = q / n
    dq = g["dL_drot"].reshape(-1, 4)
    dq_raw = (dq - qh * np.sum(qh * dq, axis=1, keepdims=True)) / n
    s = np.exp(sc["log_scales"])
    o = 1.0 / (1.0 + np.exp(-sc["logits"]))
    return {
        "means3D": g["dL_dmeans3D"].reshape(-1, 3),
        "log_scales": g["dL_dscales"].reshape(-1, 3) * s,
        "raw_rotations": dq_raw,
        "logits": g["dL_dopacity"].reshape(-1, 1) * o * (1 - o),
        "shs": g["dL_dsh"].reshape(sc["shs"].shape),
    }


def allclose_grad(a, b, rtol=1e-3, scale_atol=1e-4):
    """Per-element |a-b| <= rtol |b| + scale_atol max|b|.  Returns (ok, worst ratio)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    if b.size == 0:
        return True, 0.0
    tol = rtol * np.abs(b) + scale_atol * max(np.abs(b).max(), 1e-30)
    ratio = np.abs(a - b) / tol
    return bool(np.all(ratio <= 1.0)), float(ratio.max())


# ---------------------------------------------------------------------------
# Rounding-aware gradient bars.
#
# The running-error oracle (oracle.RUNNING_ERROR) gives every gradient entry a
# first-order bound e of the rounding error an FP32 evaluation of the
# reference's operations accumulates, in units of u = 2^-24 (Higham, Sec. 3.3).
# A GPU entry a passes against the FP32 oracle o when
#     |a - o| <= GRAD_RTOL |o| + ROUNDING_K u e,
# i.e. within the 1e-3 relative gradient bar, or within a few times the rounding
# error both evaluations are entitled to.  ROUNDING_K = 4: the oracle's own
# FP32 error stays below 0.25 u e on the test scenes, and the GPU evaluates
# the same functions in another order (tree sums, FMAs, rcp).
U32 = 2.0 ** -24
ROUNDING_K = 4.0
GRAD_RTOL = 1e-3
STATE_KEYS_FWD = ("radii", "xy", "conic", "rgb", "point_list", "ranges", "final_T", "n_contrib", "clamped")


def rounding_check(a, o, err, rtol=GRAD_RTOL, k=ROUNDING_K, skip=None):
    """(ok, worst ratio, count checked): |a - o| <= rtol |o| + k u err on
    every element not in ``skip`` (a boolean mask of the same size)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    o = np.asarray(o, dtype=np.float64).ravel()
    err = np.asarray(err, dtype=np.float64).ravel()
    tol = rtol * np.abs(o) + k * U32 * err
    keep = np.ones(a.size, bool) if skip is None else ~np.asarray(skip, bool).ravel()
    d = np.abs(a - o)[keep]
    t = tol[keep]
    if d.size == 0:
        return True, 0.0, 0
    bad = d > t
    ratio = np.where(t > 0, d / np.where(t > 0, t, 1.0), np.where(d > 0, np.inf, 0.0))
    return (not bool(bad.any())), float(ratio.max()), int(d.size)


def per_gaussian_mask(flags, width):
    """Expands a per-Gaussian flag to the entries of a [P, width] array."""
    return np.repeat(np.asarray(flags, bool), width)


def running_error_oracle(scene, state: dict, dL_dpix=None, grads2d: dict | None = None, threads: int = 1,
                         block=(16, 16)):
    """The running-error oracle on a scene whose forward state (and, for the
    preprocess backward, whose 2D gradients) are given -- normally the GPU's
    own arrays, so that both sides read identical FP32 inputs.  Runs the
    render backward when dL_dpix is given, then cov2d / preprocess backward
    (on the 2D gradients of ``grads2d`` when given, else its own)."""
    from oracle import RUNNING_ERROR, Oracle
    o = Oracle(scene, RUNNING_ERROR, block=block)
    o.set_threads(threads)
    if scene.get("scales") is not None and scene.get("cov3D_precomp") is None:
        o.run("cov3d")
    for k, v in state.items():
        o.set(k, v)
    if dL_dpix is not None:
        o.run("render_backward", dL_dpix)
    if grads2d:
        for k, v in grads2d.items():
            o.set(k, v)
    return o.run("cov2d_backward").run("preprocess_backward")


def host_threads() -> int:
    """Threads for the OpenMP oracle: the process's CPU share, at most 16."""
    import os
    return max(1, min(16, len(os.sched_getaffinity(0))))
