"""Training state of the Gaussians around the rasterizer: the host-side
mirror of the reference's GaussianModel (src/gaussian_model.cpp) for what a
training step touches.

* Parameters are the raw tensors of the reference: xyz [P,3], features_dc
  [P,1,3], features_rest [P,K,3], scaling (log) [P,3], rotation (raw
  quaternion) [P,4], opacity (logit) [P,1], with its activations
  (gaussian_model.cpp:54-77).
* ``training_setup`` builds the six Adam groups of
  GaussianModel::trainingSetup (gaussian_model.cpp:485-518); ``step`` runs
  them as one fused launch (omnigs_adam_step) instead of torch::optim::Adam.
* ``update_learning_rate`` is exponLrFunc (omnigs_expon_lr).
* ``add_densification_stats`` is one kernel for the max_radii2D update and
  addDensificationStats (omnigs_densification_stats).
* densify / clone / split / prune / reset_opacity follow
  gaussian_model.cpp:564-853 with torch index ops, as the reference does:
  they run every few hundred steps and are data-dependent compactions.
  Optimizer state is carried as the reference carries it: kept rows keep
  their moments, new rows start at zero, the step count is preserved.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import capi, ply

GROUPS = ("xyz", "features_dc", "features_rest", "opacity", "scaling", "rotation")


@dataclass
class OptimizationParams:
    """The Optimization.* fields the model uses (cfg/lonlat/*.yaml)."""
    position_lr_init: float = 0.00016
    position_lr_final: float = 0.0000016
    position_lr_delay_mult: float = 0.01
    position_lr_max_steps: int = 30000
    feature_lr: float = 0.0025
    opacity_lr: float = 0.05
    scaling_lr: float = 0.005
    rotation_lr: float = 0.001
    percent_dense: float = 0.01


def build_rotation(q: torch.Tensor) -> torch.Tensor:
    """Rotation matrices of (w, x, y, z) quaternions, normalised first
    (general_utils.h:34-60)."""
    q = q / torch.sqrt((q * q).sum(1, keepdim=True))
    r, x, y, z = q.unbind(1)
    R = torch.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y),
        2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x),
        2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)], 1)
    return R.view(-1, 3, 3)


def inverse_sigmoid(x: torch.Tensor) -> torch.Tensor:
    return torch.log(x / (1 - x))


class GaussianModel:
    def __init__(self, params: dict, sh_degree: int = 3, spatial_lr_scale: float = 1.0, device=None):
        dev = device or params["xyz"].device
        self.p = {k: params[k].detach().to(dev, torch.float32).contiguous() for k in GROUPS}
        self.max_sh_degree = sh_degree
        self.spatial_lr_scale = spatial_lr_scale
        P = self.P
        self.exist_since_iter = torch.zeros((P,), dtype=torch.int32, device=dev)
        self.max_radii2D = torch.zeros((P,), device=dev)
        self.xyz_gradient_accum = torch.zeros((P, 1), device=dev)
        self.denom = torch.zeros((P, 1), device=dev)
        self.state = None
        self.percent_dense = 0.01

    # ------------------------------------------------------------ accessors
    @property
    def P(self) -> int:
        return self.p["xyz"].shape[0]

    @property
    def device(self):
        return self.p["xyz"].device

    def get_scaling(self):
        return torch.exp(self.p["scaling"])

    def get_rotation(self):
        return torch.nn.functional.normalize(self.p["rotation"])

    def get_opacity(self):
        return torch.sigmoid(self.p["opacity"])

    def get_features(self):
        return torch.cat([self.p["features_dc"], self.p["features_rest"]], 1)

    # --------------------------------------------------------- optimizer
    def training_setup(self, opt: OptimizationParams) -> None:
        self.percent_dense = opt.percent_dense
        self.xyz_gradient_accum = torch.zeros((self.P, 1), device=self.device)
        self.denom = torch.zeros((self.P, 1), device=self.device)
        s = self.spatial_lr_scale
        self.lr = {"xyz": opt.position_lr_init * s, "features_dc": opt.feature_lr,
                   "features_rest": opt.feature_lr / 20.0, "opacity": opt.opacity_lr, "scaling": opt.scaling_lr,
                   "rotation": opt.rotation_lr}
        self.lr_init, self.lr_final = opt.position_lr_init * s, opt.position_lr_final * s
        self.lr_delay_mult, self.lr_delay_steps, self.max_steps = opt.position_lr_delay_mult, 0, \
            opt.position_lr_max_steps
        self.state = {k: {"exp_avg": torch.zeros_like(v), "exp_avg_sq": torch.zeros_like(v), "step": 0}
                      for k, v in self.p.items()}

    def update_learning_rate(self, step: int) -> float:
        lr = float(capi.load().omnigs_expon_lr(int(step), self.lr_init, self.lr_final, self.lr_delay_steps,
                                               self.lr_delay_mult, self.max_steps))
        self.lr["xyz"] = lr
        return lr

    def step(self, grads: dict) -> None:
        """One Adam step of every group that has a gradient (grads: raw-
        parameter gradients keyed like the parameters)."""
        groups = []
        for k in GROUPS:
            g = grads.get(k)
            if g is None:
                continue
            st = self.state[k]
            st["step"] += 1
            groups.append(dict(param=self.p[k], grad=g.contiguous(), exp_avg=st["exp_avg"],
                               exp_avg_sq=st["exp_avg_sq"], lr=self.lr[k], eps=1e-15, step=st["step"]))
        if groups:
            capi.adam_step(groups)

    # ----------------------------------------------------- densification
    def add_densification_stats(self, dL_dmean2D: torch.Tensor, radii: torch.Tensor) -> None:
        lib = capi.load()
        capi._check(lib.omnigs_densification_stats(
            self.P, capi._ptr(radii.contiguous()), capi._ptr(dL_dmean2D.contiguous()), capi._ptr(self.max_radii2D),
            capi._ptr(self.xyz_gradient_accum), capi._ptr(self.denom), capi.stream_handle(self.device)),
            "omnigs_densification_stats")

    def _replace_rows(self, keep: torch.Tensor | None, new: dict | None) -> None:
        """Keeps rows where ``keep`` and appends ``new`` rows, with the
        optimizer state carried (gaussian_model.cpp:605-731)."""
        for k in GROUPS:
            v = self.p[k]
            st = self.state[k] if self.state else None
            if keep is not None:
                v = v[keep]
                if st:
                    st["exp_avg"], st["exp_avg_sq"] = st["exp_avg"][keep], st["exp_avg_sq"][keep]
            if new is not None:
                v = torch.cat([v, new[k]], 0)
                if st:
                    z = torch.zeros_like(new[k])
                    st["exp_avg"], st["exp_avg_sq"] = torch.cat([st["exp_avg"], z]), torch.cat([st["exp_avg_sq"], z])
            self.p[k] = v.contiguous()
            if st:
                st["exp_avg"], st["exp_avg_sq"] = st["exp_avg"].contiguous(), st["exp_avg_sq"].contiguous()

    def prune_points(self, mask: torch.Tensor) -> None:
        keep = ~mask
        self._replace_rows(keep, None)
        self.exist_since_iter = self.exist_since_iter[keep]
        self.xyz_gradient_accum = self.xyz_gradient_accum[keep]
        self.denom = self.denom[keep]
        self.max_radii2D = self.max_radii2D[keep]

    def densification_postfix(self, new: dict, new_exist_since_iter: torch.Tensor) -> None:
        self._replace_rows(None, new)
        self.exist_since_iter = torch.cat([self.exist_since_iter, new_exist_since_iter])
        self.xyz_gradient_accum = torch.zeros((self.P, 1), device=self.device)
        self.denom = torch.zeros((self.P, 1), device=self.device)
        self.max_radii2D = torch.zeros((self.P,), device=self.device)

    def densify_and_clone(self, grads: torch.Tensor, grad_threshold: float, scene_extent: float) -> None:
        sel = torch.linalg.norm(grads, dim=-1) >= grad_threshold
        sel = sel & (self.get_scaling().max(1).values <= self.percent_dense * scene_extent)
        self.densification_postfix({k: self.p[k][sel] for k in GROUPS}, self.exist_since_iter[sel])

    def densify_and_split(self, grads: torch.Tensor, grad_threshold: float, scene_extent: float, N: int = 2,
                          generator: torch.Generator | None = None) -> None:
        n_init = self.P
        padded = torch.zeros((n_init,), device=self.device)
        padded[:grads.shape[0]] = grads.squeeze()
        sel = (padded >= grad_threshold) & (self.get_scaling().max(1).values > self.percent_dense * scene_extent)
        stds = self.get_scaling()[sel].repeat(N, 1)
        samples = torch.normal(torch.zeros_like(stds), stds, generator=generator)
        rots = build_rotation(self.p["rotation"][sel]).repeat(N, 1, 1)
        new = {
            "xyz": torch.bmm(rots, samples.unsqueeze(-1)).squeeze(-1) + self.p["xyz"][sel].repeat(N, 1),
            "scaling": torch.log(self.get_scaling()[sel].repeat(N, 1) / (0.8 * N)),
            "rotation": self.p["rotation"][sel].repeat(N, 1),
            "features_dc": self.p["features_dc"][sel].repeat(N, 1, 1),
            "features_rest": self.p["features_rest"][sel].repeat(N, 1, 1),
            "opacity": self.p["opacity"][sel].repeat(N, 1),
        }
        self.densification_postfix(new, self.exist_since_iter[sel].repeat(N))
        prune = torch.cat([sel, torch.zeros((N * int(sel.sum()),), dtype=torch.bool, device=self.device)])
        self.prune_points(prune)

    def densify_and_prune(self, max_grad: float, min_opacity: float, extent: float, max_screen_size: int,
                          prune_by_extent: bool = True, generator: torch.Generator | None = None) -> None:
        grads = self.xyz_gradient_accum / self.denom
        grads[grads.isnan()] = 0.0
        self.densify_and_clone(grads, max_grad, extent)
        self.densify_and_split(grads, max_grad, extent, generator=generator)
        prune = (self.get_opacity() < min_opacity).squeeze(1)
        if max_screen_size:
            big_vs = self.max_radii2D > max_screen_size
            big_ws = (self.get_scaling().max(1).values > 0.1 * extent) if prune_by_extent else torch.zeros_like(big_vs)
            prune = prune | big_vs | big_ws
        self.prune_points(prune)

    def reset_opacity(self) -> None:
        """GaussianModel::resetOpacity (gaussian_model.cpp:564-573):
        opacity = inverse_sigmoid(min(sigmoid(opacity), ones_like(sigmoid(opacity) * 0.01))).
        The reference clamps against ones_like(...), a tensor of ones, so the
        activated opacities (all below 1) pass through unchanged apart from
        the sigmoid / inverse-sigmoid round trip; what the reset does there is
        restart the Adam moments (replaceTensorToOptimizer, :575-594: zeroed
        exp_avg and exp_avg_sq, step kept).  Kept as the reference behaves."""
        act = self.get_opacity()
        new = inverse_sigmoid(torch.minimum(act, torch.ones_like(act * 0.01)))
        self.p["opacity"] = new.contiguous()
        if self.state:
            st = self.state["opacity"]
            st["exp_avg"], st["exp_avg_sq"] = torch.zeros_like(new), torch.zeros_like(new)

    # ------------------------------------------------------------- files
    def save_ply(self, path) -> None:
        ply.save_ply(path, self.p, self.max_sh_degree)

    @classmethod
    def load_ply(cls, path, sh_degree: int = 3, device="cpu", spatial_lr_scale: float = 1.0) -> "GaussianModel":
        return cls(ply.load_ply(path, sh_degree, device), sh_degree, spatial_lr_scale, device)

