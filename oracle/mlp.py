"""CPU restatement of the colour decoder (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's CPU leg, never by
the ``omnigs_amd`` product path.

The decoder has no counterpart in the reference (it replaces the degree-3 SH
colour of computeColorFromSH, cuda_rasterizer/forward.cu:30-83).  This file
states what the kernels compute:
  x   = [latent (16), Y2(dir) (9)]           Y2 = degree-2 real SH basis
  h1  = relu(W1 x + b1)       (64)
  h2  = relu(W2 h1 + b2)      (64)
  rgb = sigmoid(W3 h2 + b3)   (3)
with BF16 rounding (nearest even) wherever the kernels feed the MFMA unit:
x, the weights, h1 and h2 as the next layer's input, and in the backward the
adjoints dO, dh2, dh1 and the staged activations of the weight gradients;
sums are FP64 here (FP32 on the GPU).  ``bf16=False`` gives the exact FP64
function, whose gradients the tests check by finite differences.

Parameter vector theta: W1 [64][25], b1 [64], W2 [64][64], b2 [64], W3 [3][64],
b3 [3].
"""
from __future__ import annotations

import numpy as np

LATENT, HIDDEN, IN, OUT = 16, 64, 25, 3
SIZES = [("W1", (HIDDEN, IN)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
         ("W3", (OUT, HIDDEN)), ("b3", (OUT,))]
PARAM_COUNT = sum(int(np.prod(s)) for _, s in SIZES)

SH_C0 = 0.28209479177387814
SH_C1 = 0.4886025119029199
SH_C2 = (1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396)


def bf16(x) -> np.ndarray:
    """Round FP32 values to BF16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(theta) -> dict:
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + n]).reshape(shape)
        o += n
    return out


def pack(p: dict) -> np.ndarray:
    return np.concatenate([np.asarray(p[n], dtype=np.float64).reshape(-1) for n, _ in SIZES])


def sh2_basis(d: np.ndarray) -> np.ndarray:
    """[N,3] unit directions -> [N,9] basis (constants and signs of auxiliary.h:32-49)."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    return np.stack([np.full_like(x, SH_C0), -SH_C1 * y, SH_C1 * z, -SH_C1 * x,
                     SH_C2[0] * x * y, SH_C2[1] * y * z, SH_C2[2] * (2 * z * z - x * x - y * y),
                     SH_C2[3] * x * z, SH_C2[4] * (x * x - y * y)], axis=1)


def sh2_basis_vjp(d: np.ndarray, g: np.ndarray) -> np.ndarray:
    """Gradient of sum(g * sh2_basis(d)) with respect to d (treated as free)."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    c20, c22, c24 = SH_C2[0], SH_C2[2], SH_C2[4]
    gx = -SH_C1 * g[:, 3] + c20 * y * g[:, 4] - c20 * z * g[:, 7] - 2 * c22 * x * g[:, 6] + 2 * c24 * x * g[:, 8]
    gy = -SH_C1 * g[:, 1] + c20 * x * g[:, 4] - c20 * z * g[:, 5] - 2 * c22 * y * g[:, 6] - 2 * c24 * y * g[:, 8]
    gz = SH_C1 * g[:, 2] - c20 * y * g[:, 5] - c20 * x * g[:, 7] + 4 * c22 * z * g[:, 6]
    return np.stack([gx, gy, gz], axis=1)


def _r(x, on):
    return bf16(x).astype(np.float64) if on else np.asarray(x, dtype=np.float64)


def forward(theta, latents, dirs, bf16_ops: bool = True):
    """Returns (rgb [N,3], cache) for FP32/FP64 latents [N,16] and unit dirs [N,3]."""
    p = unpack(np.asarray(theta, dtype=np.float32 if bf16_ops else np.float64))
    W1, W2, W3 = _r(p["W1"], bf16_ops), _r(p["W2"], bf16_ops), _r(p["W3"], bf16_ops)
    b1, b2, b3 = (np.asarray(p[k], dtype=np.float64) for k in ("b1", "b2", "b3"))
    basis = sh2_basis(np.asarray(dirs, dtype=np.float32 if bf16_ops else np.float64))
    x = np.concatenate([_r(latents, bf16_ops), _r(basis, bf16_ops)], axis=1)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h1 = h1.astype(np.float32).astype(np.float64) if bf16_ops else h1
    h1in = _r(h1, bf16_ops)
    h2 = np.maximum(h1in @ W2.T + b2, 0.0)
    h2 = h2.astype(np.float32).astype(np.float64) if bf16_ops else h2
    h2in = _r(h2, bf16_ops)
    o = h2in @ W3.T + b3
    rgb = 1.0 / (1.0 + np.exp(-o))
    return rgb, dict(x=x, h1=h1, h1in=h1in, h2=h2, h2in=h2in, rgb=rgb, W1=W1, W2=W2, W3=W3, bf16=bf16_ops)


def backward(cache, dL_drgb):
    """Adjoints of forward(): returns (dL/dlatent [N,16], dL/dbasis [N,9], dL/dtheta [PARAM_COUNT])."""
    on = cache["bf16"]
    s = cache["rgb"]
    dO = np.asarray(dL_drgb, dtype=np.float64) * s * (1.0 - s)
    dO = dO.astype(np.float32).astype(np.float64) if on else dO
    dOb = _r(dO, on)
    dh2 = (dOb @ cache["W3"]) * (cache["h2"] > 0)
    dh2 = dh2.astype(np.float32).astype(np.float64) if on else dh2
    dh2b = _r(dh2, on)
    dh1 = (dh2b @ cache["W2"]) * (cache["h1"] > 0)
    dh1 = dh1.astype(np.float32).astype(np.float64) if on else dh1
    dh1b = _r(dh1, on)
    dx = dh1b @ cache["W1"]
    g = {"W1": dh1b.T @ cache["x"], "b1": dh1b.sum(0), "W2": dh2b.T @ cache["h1in"], "b2": dh2b.sum(0),
         "W3": dOb.T @ cache["h2in"], "b3": dOb.sum(0)}
    return dx[:, :LATENT], dx[:, LATENT:IN], pack(g)


def direction_to_mean_vjp(v: np.ndarray, gd: np.ndarray) -> np.ndarray:
    """Gradient through d = v / |v| (dnormvdv, auxiliary.h:134-144)."""
    n = np.linalg.norm(v, axis=1, keepdims=True)
    d = v / n
    return (gd - d * np.sum(d * gd, axis=1, keepdims=True)) / n


# ----------------------------------------------------------------------------
# Distillation sampling (counter hash; the same integers as the kernels).

def _mix32(x):
    x = np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7feb352d) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846ca68b) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def seeded_directions(seed: int, g: np.ndarray, k: np.ndarray) -> np.ndarray:
    g = np.asarray(g, dtype=np.uint64)
    k = np.asarray(k, dtype=np.uint64)
    inner = (g * 0x9E3779B9 + k * 0x85EBCA6B + 0x27d4eb2f) & 0xFFFFFFFF
    h1 = _mix32(np.uint64(seed & 0xFFFFFFFF) ^ _mix32(inner))
    h2 = _mix32(h1 ^ 0x165667b1)
    u = (h1 >> 8).astype(np.float32) * np.float32(1.0 / 16777216.0)
    v = (h2 >> 8).astype(np.float32) * np.float32(1.0 / 16777216.0)
    z = np.float32(1.0) - np.float32(2.0) * u
    s = np.sqrt(np.maximum(np.float32(0.0), np.float32(1.0) - z * z))
    ph = np.float32(6.283185307179586) * v
    return np.stack([s * np.cos(ph), s * np.sin(ph), z], axis=1).astype(np.float32)


def seeded_view_centres(seed: int, g: np.ndarray, k: np.ndarray, view_box) -> np.ndarray:
    """View centres uniform in view_box = (lo xyz, hi xyz): three more words
    of the direction hash's counter; lo + u (hi - lo) with one FP32 rounding
    (the kernels' fmaf; exact product in FP64, then rounded)."""
    g = np.asarray(g, dtype=np.uint64)
    k = np.asarray(k, dtype=np.uint64)
    inner = (g * 0x9E3779B9 + k * 0x85EBCA6B + 0x27d4eb2f) & 0xFFFFFFFF
    h0 = _mix32(np.uint64(seed & 0xFFFFFFFF) ^ _mix32(inner))
    h1 = _mix32(h0 ^ 0x5bd1e995)
    h2 = _mix32(h1 ^ 0x165667b1)
    h3 = _mix32(h2 ^ 0xc2b2ae35)
    box = np.asarray(view_box, dtype=np.float32)
    out = []
    for i, h in enumerate((h1, h2, h3)):
        u = (h >> 8).astype(np.float32) * np.float32(1.0 / 16777216.0)
        span = np.float32(box[3 + i] - box[i])
        out.append((u.astype(np.float64) * np.float64(span) + np.float64(box[i])).astype(np.float32))
    return np.stack(out, axis=1)


def view_box_directions(means: np.ndarray, centres: np.ndarray) -> np.ndarray:
    """unit(mean - centre) in FP32 ((0, 0, 1) for a zero vector)."""
    v = (means.astype(np.float32) - centres.astype(np.float32)).astype(np.float32)
    n = np.sqrt(np.sum(v.astype(np.float64) ** 2, axis=1, keepdims=True)).astype(np.float32)
    d = np.where(n > 0, v / np.where(n > 0, n, 1), np.array([0, 0, 1], dtype=np.float32))
    return d.astype(np.float32)


def distill_dir_index(seed: int, step: int, g: np.ndarray, num_dirs: int) -> np.ndarray:
    g = np.asarray(g, dtype=np.uint64)
    x = ((seed * 0x632BE5AB) + (step * 0x9E3779B9) + g) & 0xFFFFFFFF
    return _mix32(x) % num_dirs


def sh3_colour(shs: np.ndarray, dirs: np.ndarray) -> np.ndarray:
    """Degree-3 SH colour + 0.5 clamped at 0 (computeColorFromSH, forward.cu:30-83) in FP64."""
    x, y, z = (dirs[:, i].astype(np.float64) for i in range(3))
    xx, yy, zz, xy, yz, xz = x * x, y * y, z * z, x * y, y * z, x * z
    C3 = (-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
          1.445305721320277, -0.5900435899266435)
    b = [np.full_like(x, SH_C0), -SH_C1 * y, SH_C1 * z, -SH_C1 * x,
         SH_C2[0] * xy, SH_C2[1] * yz, SH_C2[2] * (2 * zz - xx - yy), SH_C2[3] * xz, SH_C2[4] * (xx - yy),
         C3[0] * y * (3 * xx - yy), C3[1] * xy * z, C3[2] * y * (4 * zz - xx - yy),
         C3[3] * z * (2 * zz - 3 * xx - 3 * yy), C3[4] * x * (4 * zz - xx - yy), C3[5] * z * (xx - yy),
         C3[6] * x * (xx - 3 * yy)]
    B = np.stack(b, axis=1)  # [N,16]
    rgb = np.einsum("nk,nkc->nc", B, shs[:, :16, :].astype(np.float64)) + 0.5
    return np.maximum(rgb, 0.0)


def distill_step(theta, latents, shs, seed: int, step: int, num_dirs: int, means=None, view_box=None):
    """One distillation pass as the kernel runs it: (loss, dL/dlatent, dL/dtheta).
    With means [P,3] and view_box, the directions run from seeded view
    centres in the box to the means; else they are uniform on the sphere."""
    P = latents.shape[0]
    g = np.arange(P)
    k = distill_dir_index(seed, step, g, num_dirs)
    if means is None:
        dirs = seeded_directions(seed, g, k)
    else:
        dirs = view_box_directions(means, seeded_view_centres(seed, g, k, view_box))
    target = sh3_colour(shs, dirs)
    rgb, cache = forward(theta, latents, dirs)
    err = rgb - target
    loss = float(np.mean(err ** 2))
    dlat, _, dth = backward(cache, 2.0 * err / (3.0 * P))
    return loss, dlat, dth


def init_theta(seed: int = 0) -> np.ndarray:
    """torch.nn.Linear's default initialisation (uniform +-1/sqrt(fan_in))."""
    rng = np.random.default_rng(seed)
    p = {}
    for (wn, (o, i)), bn in ((("W1", (HIDDEN, IN)), "b1"), (("W2", (HIDDEN, HIDDEN)), "b2"),
                             (("W3", (OUT, HIDDEN)), "b3")):
        bound = 1.0 / np.sqrt(i)
        p[wn] = rng.uniform(-bound, bound, size=(o, i))
        p[bn] = rng.uniform(-bound, bound, size=(o,))
    return pack(p).astype(np.float32)
