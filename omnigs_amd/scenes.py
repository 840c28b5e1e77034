"""Seeded synthetic equirectangular scenes (host numpy, FP32).

Each generator returns a dict with the raw (pre-activation) parameters and the
activated tensors that the rasterizer consumes, mirroring the caller-side
activations of OmniGS (gaussian_model.cpp:54-77): scales = exp(log_scales),
rotations = q / |q|, opacities = sigmoid(logits), shs = cat(dc, rest).

The camera follows the keyframe convention (gaussian_keyframe.cpp:132-173):
``viewmatrix`` holds Tcw^T as a row-major [4,4] array, i.e. Tcw column-major,
so its flat entries 12..14 are the translation of Tcw.
"""
from __future__ import annotations

import numpy as np

SH_C0 = 0.28209479177387814


def viewmatrix_from_pose(Rcw: np.ndarray, tcw: np.ndarray) -> np.ndarray:
    """Flat [16] column-major Tcw (= row-major Tcw^T)."""
    T = np.eye(4, dtype=np.float64)
    T[:3, :3] = Rcw
    T[:3, 3] = tcw
    return np.ascontiguousarray(T.T.reshape(-1)).astype(np.float32)


def camera_at(center: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Identity rotation camera at ``center``: Tcw = [I | -c]."""
    c = np.asarray(center, dtype=np.float64)
    return viewmatrix_from_pose(np.eye(3), -c), c.astype(np.float32)


def activate(raw: dict) -> dict:
    """Caller-side activations (gaussian_model.cpp:54-77)."""
    q = raw["raw_rotations"].astype(np.float64)
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    out = dict(raw)
    out["scales"] = np.exp(raw["log_scales"].astype(np.float64)).astype(np.float32)
    out["rotations"] = q.astype(np.float32)
    out["opacities"] = (1.0 / (1.0 + np.exp(-raw["logits"].astype(np.float64)))).astype(np.float32)
    return out


def _gaussians(rng: np.random.Generator, means: np.ndarray, degree: int = 3) -> dict:
    P = means.shape[0]
    M = 16
    log_s = rng.normal(-4.0, 0.5, size=(P, 3)).astype(np.float32)
    q = rng.normal(0.0, 1.0, size=(P, 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    logits = rng.normal(0.0, 1.0, size=(P, 1)).astype(np.float32)
    shs = np.empty((P, M, 3), dtype=np.float32)
    shs[:, 0, :] = rng.uniform(-0.5, 0.5, size=(P, 3)) / 0.2821
    shs[:, 1:, :] = rng.normal(0.0, 0.05, size=(P, M - 1, 3))
    return activate({"means3D": means.astype(np.float32), "log_scales": log_s, "raw_rotations": q,
                     "logits": logits, "shs": shs, "degree": degree, "scale_modifier": 1.0})


def uniform_directions(rng: np.random.Generator, n: int) -> np.ndarray:
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def small_scene(P: int = 20000, width: int = 512, height: int = 256, seed: int = 0,
                grad_std: float = 1e-2) -> dict:
    """The 20k-Gaussian 512x256 case: camera at the origin, identity pose,
    means uniform in direction at radius U(1,5), log-scales N(-4,0.5),
    normalised N(0,1) quaternions, logits N(0,1), degree-3 SH with
    DC ~ U(-0.5,0.5)/0.2821 and higher orders N(0,0.05) (read as std devs),
    black background, upstream gradient N(0, 1e-2) (std)."""
    rng = np.random.default_rng(seed)
    means = uniform_directions(rng, P) * rng.uniform(1.0, 5.0, size=(P, 1))
    sc = _gaussians(rng, means)
    vm, c = camera_at(np.zeros(3))
    sc.update(width=width, height=height, viewmatrix=vm, campos=c,
              background=np.zeros(3, dtype=np.float32))
    sc["dL_dpix"] = rng.normal(0.0, grad_std, size=(3, height, width)).astype(np.float32)
    return sc


def room_scene(P: int = 1_500_000, width: int = 1920, height: int = 960, seed: int = 1,
               seam_fraction: float = 0.10) -> dict:
    """The 1.5M-Gaussian 1920x960 case: means uniform in the box
    [-6,6]x[-2,3]x[-6,6] m with ``seam_fraction`` of them clustered near the
    camera's longitude seam (lon = pi + N(0, 0.05) rad, lat U(-0.5, 0.5),
    range U(1, 6) m), camera at U([-0.5,0.5]^3) with identity rotation, other
    distributions as in :func:`small_scene`.  ``target`` is an N(0.5, 0.1)
    image for the L1 upstream gradient sign(img - target) / (3 H W)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.5, 0.5, size=3)
    n_seam = int(round(P * seam_fraction))
    n_box = P - n_seam
    box = rng.uniform([-6.0, -2.0, -6.0], [6.0, 3.0, 6.0], size=(n_box, 3))
    lon = np.pi + rng.normal(0.0, 0.05, size=n_seam)
    lat = rng.uniform(-0.5, 0.5, size=n_seam)
    rho = rng.uniform(1.0, 6.0, size=n_seam)
    seam = c + rho[:, None] * np.stack([np.cos(lat) * np.sin(lon), np.sin(lat), np.cos(lat) * np.cos(lon)], 1)
    means = np.concatenate([box, seam], 0)
    means = means[rng.permutation(P)]
    sc = _gaussians(rng, means)
    vm, cc = camera_at(c)
    sc.update(width=width, height=height, viewmatrix=vm, campos=cc,
              background=np.zeros(3, dtype=np.float32))
    sc["target"] = rng.normal(0.5, 0.1, size=(3, height, width)).astype(np.float32)
    return sc


def random_pose_camera(rng: np.random.Generator, half_extent: float = 1.0) -> tuple[np.ndarray, np.ndarray]:
    """Random rotation (normalised N(0,1) quaternion) at a uniform position in a cube."""
    c = rng.uniform(-half_extent, half_extent, size=3)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    Rwc = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                    [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                    [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    Rcw = Rwc.T
    return viewmatrix_from_pose(Rcw, -Rcw @ c), c.astype(np.float32)
