"""Learned colour decoder for the equirectangular rasterizer (an extension with
no reference counterpart: it stands in for the degree-3 SH colour of
computeColorFromSH, cuda_rasterizer/forward.cu:30-83).

Each Gaussian carries a 16-dim latent; the decoder maps [latent, Y2(view
direction)] through 25 -> 64 -> 64 -> 3 (ReLU, ReLU, sigmoid) on the MFMA unit
(BF16 operands, FP32 accumulation).  The latents are BF16 in HBM (32 B per
Gaussian instead of 192 B of SH coefficients); FP32 masters and Adam states are
kept for training.

Training is distillation from the SH model's own colours: every pass gives
each Gaussian one of its ``num_dirs`` seeded view directions and regresses the
decoder onto the SH colour there (mean squared error), with Adam on the
parameters and the latents.  Everything runs through the C ABI
(omnigs_mlp_distill_step, omnigs_adam_step); no torch op touches the math.
"""
from __future__ import annotations

import math

import torch

from . import capi


def init_theta(generator: torch.Generator, device) -> torch.Tensor:
    """torch.nn.Linear's default initialisation (uniform +-1/sqrt(fan_in)) of
    the FP32 parameter vector W1 [64][25], b1, W2 [64][64], b2, W3 [3][64], b3."""
    parts = []
    for fan_out, fan_in in ((64, 25), (64, 64), (3, 64)):
        bound = 1.0 / math.sqrt(fan_in)
        parts.append((torch.rand(fan_out * fan_in, generator=generator) * 2 - 1) * bound)
        parts.append((torch.rand(fan_out, generator=generator) * 2 - 1) * bound)
    return torch.cat(parts).to(device=device, dtype=torch.float32).contiguous()


class ColourDecoder:
    """Decoder parameters, per-Gaussian latents and their optimizer state."""

    LATENT = 16

    def __init__(self, P: int, device="cuda:0", seed: int = 0, latent_std: float = 0.1):
        self.device = torch.device(device)
        g = torch.Generator().manual_seed(seed)
        self.theta = init_theta(g, self.device)
        self.latents = (torch.randn((P, self.LATENT), generator=g) * latent_std).to(self.device).contiguous()
        self.latents_bf16 = capi.f32_to_bf16(self.latents)
        self.packed = capi.mlp_pack(self.theta)
        self._adam = None
        self.step_count = 0

    @property
    def P(self) -> int:
        return self.latents.shape[0]

    def mlp_args(self):
        """(packed image, BF16 latents) for capi.forward / capi.backward."""
        return self.packed, self.latents_bf16

    def colours(self, dirs: torch.Tensor) -> torch.Tensor:
        """Decoder colours of every Gaussian for unit directions [P,3]."""
        return capi.mlp_eval(self.latents_bf16, dirs, self.packed)

    def _adam_state(self):
        if self._adam is None:
            z = torch.zeros_like
            self._adam = dict(m_t=z(self.theta), v_t=z(self.theta), m_l=z(self.latents), v_l=z(self.latents),
                              g_t=z(self.theta), g_l=z(self.latents))
        return self._adam

    def apply_gradients(self, d_theta: torch.Tensor, d_latent: torch.Tensor, lr: float = 1e-3,
                        lr_latent: float | None = None, eps: float = 1e-15) -> None:
        """One Adam step on (theta, latents) with the given gradients, then
        refreshes the packed image and the BF16 latents."""
        st = self._adam_state()
        self.step_count += 1
        capi.adam_step([
            dict(param=self.theta, grad=d_theta, exp_avg=st["m_t"], exp_avg_sq=st["v_t"], lr=lr, eps=eps,
                 step=self.step_count),
            dict(param=self.latents, grad=d_latent, exp_avg=st["m_l"], exp_avg_sq=st["v_l"],
                 lr=lr if lr_latent is None else lr_latent, eps=eps, step=self.step_count,
                 param_bf16=self.latents_bf16),
        ])
        capi.mlp_pack(self.theta, self.packed)

    def distill(self, shs: torch.Tensor, steps: int = 2000, lr: float = 1e-3, num_dirs: int = 64, seed: int = 1,
                lr_latent: float | None = 1e-2, log_every: int = 0, means3D: torch.Tensor | None = None,
                view_box=None):
        """Distills the decoder from SH colours (shs [P,16,3]); returns the
        per-step losses as a tensor (kept on the device until the end).
        lr is the network's Adam rate; the latents default to 1e-2 because
        each one sees a single sample per pass.

        Directions: uniform on the sphere, or, with means3D [P,3] and
        view_box (lo xyz, hi xyz), the view directions from ``num_dirs``
        seeded camera centres per Gaussian in the box where the cameras that
        render the scene stand (the colours a render can show)."""
        if means3D is not None:
            assert view_box is not None and len(view_box) == 6
            means3D = means3D.contiguous()
        assert shs.shape[0] == self.P and shs.dtype == torch.float32
        shs = shs.contiguous()
        st = self._adam_state()
        ws = capi.mlp_workspace(self.P, self.device)
        losses = torch.empty((steps,), dtype=torch.float32, device=self.device)
        for i in range(steps):
            capi.mlp_distill_step(shs, self.latents, self.packed, seed, self.step_count, num_dirs, st["g_l"],
                                  st["g_t"], losses[i:i + 1], ws, means3D=means3D, view_box=view_box)
            self.apply_gradients(st["g_t"], st["g_l"], lr=lr, lr_latent=lr_latent)
            if log_every and (i + 1) % log_every == 0:
                print(f"distill step {i + 1}: loss {float(losses[i]):.6f}", flush=True)
        return losses
