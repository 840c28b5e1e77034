"""CPU restatement of the mapper's training loss (FP64, torch on the CPU).

TEST INFRASTRUCTURE ONLY: imported by tests/ and never by the ``omnigs_amd``
product path.

Follows the reference's ops one for one:
  * gaussian(11, 1.5)        include/loss_utils.h:53-67
  * create_window            include/loss_utils.h:69-78 (outer product, per channel)
  * _ssim                    include/loss_utils.h:80-113 (grouped conv2d, padding 5,
                             C1 = 0.01^2, C2 = 0.03^2, mean of the map)
  * l1_loss                  include/loss_utils.h:31-34
  * loss and bottom crop     src/gaussian_mapper.cpp:384-407
The image gradient comes from torch autograd through these ops, which is what
the reference's loss.backward() (gaussian_mapper.cpp:408) computes.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def gaussian_window_1d(window_size: int = 11, sigma: float = 1.5) -> torch.Tensor:
    vals = [math.exp(-((x - window_size // 2) ** 2) / (2.0 * sigma * sigma)) for x in range(window_size)]
    g = torch.tensor(vals, dtype=torch.float64)
    return g / g.sum()


def create_window(window_size: int, channel: int) -> torch.Tensor:
    w1 = gaussian_window_1d(window_size, 1.5).unsqueeze(1)
    w2 = w1.mm(w1.t()).unsqueeze(0).unsqueeze(0)
    return w2.expand(channel, 1, window_size, window_size).contiguous()


def ssim(img1: torch.Tensor, img2: torch.Tensor, window_size: int = 11) -> torch.Tensor:
    """Mean SSIM of two [C,H,W] (or [N,C,H,W]) images."""
    channel = img1.shape[-3]
    window = create_window(window_size, channel).to(img1.dtype)
    pad = window_size // 2
    x1, x2 = (img1.unsqueeze(0), img2.unsqueeze(0)) if img1.dim() == 3 else (img1, img2)
    conv = lambda t: torch.nn.functional.conv2d(t, window, padding=pad, groups=channel)  # noqa: E731
    mu1, mu2 = conv(x1), conv(x2)
    mu1_sq, mu2_sq, mu1_mu2 = mu1.pow(2), mu2.pow(2), mu1 * mu2
    s1 = conv(x1 * x1) - mu1_sq
    s2 = conv(x2 * x2) - mu2_sq
    s12 = conv(x1 * x2) - mu1_mu2
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    m = ((2 * mu1_mu2 + C1) * (2 * s12 + C2)) / ((mu1_sq + mu2_sq + C1) * (s1 + s2 + C2))
    return m.mean()


def l1_loss(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    return torch.abs(a - b).mean()


def skip_bottom_rows(height: int, ratio: float) -> int:
    return int(round(height * ratio)) if ratio > 0.0 else 0


def photometric_loss(image, gt_image, lambda_dssim: float = 0.2, skip_bottom: int = 0, mask=None):
    """Returns (loss, L1, SSIM, dL/dimage) as numpy float64 for [C,H,W] inputs.

    skip_bottom rows are dropped from the bottom before both terms, as the
    reference's Slice(0, -pix_skip_bottom) does; mask [H,W] multiplies the
    rendered image first (masked_image = rendered_image * mask)."""
    x = torch.tensor(np.asarray(image, dtype=np.float64), requires_grad=True)
    y = torch.tensor(np.asarray(gt_image, dtype=np.float64))
    xm = x * torch.tensor(np.asarray(mask, dtype=np.float64)) if mask is not None else x
    H = x.shape[1]
    if skip_bottom > 0:
        xm, y = xm[:, : H - skip_bottom, :], y[:, : H - skip_bottom, :]
    l1 = l1_loss(xm, y)
    if lambda_dssim != 0.0:
        s = ssim(xm, y)
        loss = (1.0 - lambda_dssim) * l1 + lambda_dssim * (1.0 - s)
    else:
        s = torch.zeros((), dtype=torch.float64)
        loss = l1
    loss.backward()
    return float(loss.detach()), float(l1.detach()), float(s.detach()), x.grad.numpy()


def ssim_brute_force(img1: np.ndarray, img2: np.ndarray, window_size: int = 11) -> float:
    """Mean SSIM by explicit per-pixel windows with zero padding (pure Python
    loops, tiny images only): the definition the grouped conv2d computes."""
    C, H, W = img1.shape
    half = window_size // 2
    g = gaussian_window_1d(window_size, 1.5).numpy()
    w = np.outer(g, g)
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    total = 0.0
    for c in range(C):
        p1 = np.zeros((H + 2 * half, W + 2 * half))
        p2 = np.zeros_like(p1)
        p1[half:half + H, half:half + W] = img1[c]
        p2[half:half + H, half:half + W] = img2[c]
        for i in range(H):
            for j in range(W):
                a = p1[i:i + window_size, j:j + window_size]
                b = p2[i:i + window_size, j:j + window_size]
                mu1, mu2 = (w * a).sum(), (w * b).sum()
                s1 = (w * a * a).sum() - mu1 * mu1
                s2 = (w * b * b).sum() - mu2 * mu2
                s12 = (w * a * b).sum() - mu1 * mu2
                total += ((2 * mu1 * mu2 + C1) * (2 * s12 + C2)) / ((mu1 ** 2 + mu2 ** 2 + C1) * (s1 + s2 + C2))
    return total / (C * H * W)
