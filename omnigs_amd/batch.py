"""Batched multi-view forward + backward over one Gaussian set (an extension;
the reference renders one view per RasterizeGaussiansCUDA call,
src/rasterize_points.cu:49-164).

``BatchRenderer`` owns the workspace and outputs of one batch of views
(omnigs_lonlat_batch_forward / _backward).  Nothing inside a forward or a
backward synchronises with the host or allocates, so a whole training step
(forward, loss gradient, backward) can be captured once into a HIP graph and
replayed (``capture``).  The instance count lives on the device; the
capacity is sized from a first eager forward, which reports the count it
needed, and a replay whose views need more than the capacity sets the
overflow flag (``overflowed``) instead of writing out of bounds.

``autotune`` times a workload for every tile shape (16x16, 32x8, 8x32) and
views-per-launch choice, each as a captured graph, and returns the fastest.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import torch

from . import capi

TILE_SHAPES = ((16, 16), (32, 8), (8, 32))


class BatchRenderer:
    """One batch of equirectangular views (sizes [(W, H), ...], at most 32)."""

    def __init__(self, P: int, sizes, tile=(16, 16), device="cuda:0", seam_wrap: bool = False,
                 capacity: int | None = None, capacity_margin: float = 1.15):
        self.lib = capi.load()
        self.P = int(P)
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        self.V = len(self.sizes)
        if not 1 <= self.V <= 32:
            raise ValueError("a batch holds 1..32 views")
        self.tile = (int(tile[0]), int(tile[1]))
        self.device = torch.device(device)
        self.seam_wrap = bool(seam_wrap)
        self.margin = capacity_margin
        self._w = (ctypes.c_int * self.V)(*[w for w, _ in self.sizes])
        self._h = (ctypes.c_int * self.V)(*[h for _, h in self.sizes])
        npix = [w * h for w, h in self.sizes]
        self.pix_offsets = [0]
        for n in npix:
            self.pix_offsets.append(self.pix_offsets[-1] + n)
        self.total_pixels = self.pix_offsets[-1]
        self.out = torch.zeros((3 * self.total_pixels,), dtype=torch.float32, device=self.device)
        self.status = torch.zeros((2,), dtype=torch.int32, device=self.device)
        self.workspace = None
        self.capacity = 0
        self._grads = None
        self._keep = []
        self._set_capacity(capacity or 1)

    # ------------------------------------------------------------------ setup
    def _set_capacity(self, capacity: int) -> None:
        capacity = int(max(1, capacity))
        nbytes = self.lib.omnigs_lonlat_batch_workspace_bytes(self.P, self.V, self._w, self._h, self.tile[0],
                                                              self.tile[1], capacity)
        if nbytes == 0:
            raise ValueError(f"invalid batch (P={self.P}, sizes={self.sizes}, tile={self.tile})")
        if self.workspace is None or self.workspace.numel() < nbytes:
            self.workspace = None
            self.workspace = torch.empty((nbytes,), dtype=torch.uint8, device=self.device)
        self.capacity = capacity

    def bind(self, *, means3D, opacities, viewmatrices, campos, background, shs=None, colors_precomp=None,
             scales=None, rotations=None, cov3D_precomp=None, degree=3, scale_modifier=1.0, mlp=None):
        """Sets the inputs (device tensors; viewmatrices [V,16], campos [V,3]).
        mlp = (packed image, BF16 latents) takes colours from the decoder."""
        assert viewmatrices.shape[0] == self.V and campos.shape[0] == self.V
        t = [x.contiguous() if isinstance(x, torch.Tensor) else x
             for x in (means3D, opacities.reshape(-1), viewmatrices, campos, background, shs, colors_precomp, scales,
                       rotations, cov3D_precomp)]
        self._keep = t + [mlp]
        (means3D, opacities, viewmatrices, campos, background, shs, colors_precomp, scales, rotations,
         cov3D_precomp) = t
        if mlp:
            shs = colors_precomp = None
        self.has_shs, self.has_scales, self.mlp = shs is not None, scales is not None, mlp
        self.M = shs.shape[1] if shs is not None else 0
        p = capi._ptr
        self.args = capi.BatchArgs(
            self.P, int(degree), self.M, self.V, self.tile[0], self.tile[1], int(self.seam_wrap),
            ctypes.cast(self._w, ctypes.c_void_p), ctypes.cast(self._h, ctypes.c_void_p), p(viewmatrices), p(campos),
            p(background), p(means3D), p(shs), p(colors_precomp), p(opacities), p(scales), float(scale_modifier),
            p(rotations), p(cov3D_precomp), p(mlp[0]) if mlp else None, p(mlp[1]) if mlp else None, self.capacity,
            self.workspace.data_ptr(), self.workspace.numel(), self.out.data_ptr(), self.status.data_ptr())
        self._grads = None
        return self

    def _sync_args(self) -> None:
        self.args.capacity = self.capacity
        self.args.workspace = self.workspace.data_ptr()
        self.args.workspace_bytes = self.workspace.numel()

    # ----------------------------------------------------------------- passes
    def forward(self) -> None:
        """Queues the batched forward on the current stream (no host sync)."""
        self._sync_args()
        st = self.lib.omnigs_lonlat_batch_forward(ctypes.byref(self.args), capi.stream_handle(self.device))
        if st != capi.OMNIGS_OK:
            raise capi.OmnigsError(st, "omnigs_lonlat_batch_forward")

    def grads(self) -> dict:
        """Gradient tensors the backward writes (allocated once per binding)."""
        if self._grads is None:
            f = dict(dtype=torch.float32, device=self.device)
            P = self.P
            g = {"dL_dmeans3D": torch.zeros((P, 3), **f), "dL_dopacity": torch.zeros((P,), **f),
                 "dL_dcov3D": torch.zeros((P, 6), **f)}
            if self.has_shs:
                g["dL_dsh"] = torch.zeros((P, self.M, 3), **f)
            if self.has_scales:
                g["dL_dscales"] = torch.zeros((P, 3), **f)
                g["dL_drotations"] = torch.zeros((P, 4), **f)
            if self.mlp:
                g["dL_dlatent"] = torch.zeros((P, 16), **f)
                g["dL_dmlp"] = torch.zeros((self.lib.omnigs_mlp_param_count(),), **f)
                g["_mlp_ws"] = torch.empty((self.lib.omnigs_mlp_workspace_bytes(P),), dtype=torch.uint8,
                                           device=self.device)
            self._grads = g
        return self._grads

    def backward(self, dL_dpix: torch.Tensor) -> dict:
        """Queues the batched backward (dL_dpix in the layout of ``out``)."""
        g = self.grads()
        p = capi._ptr
        mws = g.get("_mlp_ws")
        gr = capi.BatchGrads(p(dL_dpix), p(g["dL_dmeans3D"]), p(g["dL_dopacity"]), p(g["dL_dcov3D"]),
                             p(g.get("dL_dsh")), p(g.get("dL_dscales")), p(g.get("dL_drotations")),
                             p(g.get("dL_dlatent")), p(g.get("dL_dmlp")), mws.data_ptr() if mws is not None else None,
                             mws.numel() if mws is not None else 0)
        self._sync_args()
        st = self.lib.omnigs_lonlat_batch_backward(ctypes.byref(self.args), ctypes.byref(gr),
                                                   capi.stream_handle(self.device))
        if st != capi.OMNIGS_OK:
            raise capi.OmnigsError(st, "omnigs_lonlat_batch_backward")
        return g

    # ------------------------------------------------------------- capacity
    def needed(self) -> tuple[int, bool]:
        """(instances the last forward needed, whether it overflowed); syncs."""
        s = self.status.cpu()
        return int(s[0]) & 0xFFFFFFFF, bool(int(s[1]))

    def overflowed(self) -> bool:
        return self.needed()[1]

    def fit_capacity(self) -> int:
        """Eager forward to learn the instance count, then sizes the workspace
        (count x margin) and runs the forward again.  Returns the capacity."""
        self.forward()
        need, over = self.needed()
        if over or need > self.capacity:
            self._set_capacity(int(need * self.margin) + 1024)
            self.forward()
            need, over = self.needed()
            assert not over, "capacity sizing failed"
        return self.capacity

    # ----------------------------------------------------------------- views
    def state(self) -> dict:
        """Torch views of the workspace's arrays (omnigs_lonlat_batch_state),
        for stage-by-stage parity tests: per-item arrays over V P items,
        point_list / tile_keys over the instances of the last forward, ranges
        per tile key; sets key_shift / key_count (the key layout)."""
        self._sync_args()
        sv = capi.BatchStateView()
        st = self.lib.omnigs_lonlat_batch_state(ctypes.byref(self.args), ctypes.byref(sv))
        if st != capi.OMNIGS_OK:
            raise capi.OmnigsError(st, "omnigs_lonlat_batch_state")
        n_items = self.V * self.P
        R = min(self.needed()[0], self.capacity)
        tiles = int(sv.key_count)
        self.key_shift, self.key_count = int(sv.key_shift), int(sv.key_count)
        spec = {  # field: (dtype, count)
            "depth_key": (torch.int32, n_items), "depth_order": (torch.int32, n_items),
            "radii": (torch.int32, n_items), "render_records": (torch.float32, 12 * n_items),
            "clamped": (torch.uint8, n_items), "tiles_touched": (torch.int32, n_items),
            "point_list": (torch.int32, R), "tile_keys": (torch.int32, R),
            "final_T": (torch.float32, self.total_pixels), "n_contrib": (torch.int32, self.total_pixels),
            "ranges": (torch.int32, 2 * tiles), "dL_dmean2D": (torch.float32, 3 * n_items),
            "dL_dconic": (torch.float32, 4 * n_items), "dL_dopacity": (torch.float32, n_items),
            "dL_dcolor": (torch.float32, 3 * n_items)}
        base = self.workspace.data_ptr()
        out = {}
        for name, (dt, n) in spec.items():
            addr = getattr(sv.s, name) if hasattr(sv.s, name) else getattr(sv, name)
            if not addr or n == 0:
                continue
            off = addr - base
            nbytes = n * torch.empty((), dtype=dt).element_size()
            assert 0 <= off and off + nbytes <= self.workspace.numel(), name
            out[name] = self.workspace[off:off + nbytes].view(dt)
        rec = out["render_records"].view(n_items, 12)
        out["means2D"] = rec[:, 0:2].reshape(-1)
        out["conic_opacity"] = rec[:, 2:6].reshape(-1)
        out["rgb"] = rec[:, 6:9].reshape(-1)
        return out

    def image(self, v: int) -> torch.Tensor:
        W, H = self.sizes[v]
        o = 3 * self.pix_offsets[v]
        return self.out[o:o + 3 * W * H].view(3, H, W)

    def images(self) -> list:
        return [self.image(v) for v in range(self.V)]


@dataclass
class TuneResult:
    tile: tuple
    views_per_launch: int
    ms_per_view: float
    views_per_sec: float
    table: list


def make_step(renderers, targets, loss_fns, dLs):
    """One training step over every launch: forward, fused L1 gradient, backward."""

    def step():
        for r, t, lf, dl in zip(renderers, targets, loss_fns, dLs):
            r.forward()
            lf(r.out.view(1, 1, -1), t.view(1, 1, -1), dL_dimage=dl)
            r.backward(dl.view(-1))

    return step


def build_launches(P, sizes, tile, vpl, inputs: dict, viewmats, campos, targets_flat, device, mlp=None):
    """Splits the views into launches of ``vpl`` views, binds and sizes each."""
    rs, ts, lfs, dls = [], [], [], []
    for s0 in range(0, len(sizes), vpl):
        sub = sizes[s0:s0 + vpl]
        r = BatchRenderer(P, sub, tile=tile, device=device)
        r.bind(viewmatrices=viewmats[s0:s0 + vpl].contiguous(), campos=campos[s0:s0 + vpl].contiguous(), mlp=mlp,
               **inputs)
        r.fit_capacity()
        rs.append(r)
        ts.append(targets_flat[3 * sum(w * h for w, h in sizes[:s0]):][:r.out.numel()])
        lfs.append(capi.PhotometricLoss(lambda_dssim=0.0))
        dls.append(torch.empty_like(r.out))
    return rs, ts, lfs, dls


def capture(step, warmup: int = 2):
    """Captures ``step`` into a CUDA/HIP graph after eager warm-up runs."""
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    torch.cuda.synchronize()
    return g


def time_graph(g, reps: int) -> float:
    """Median ms of ``reps`` replays (events on the current stream)."""
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in evs:
        a.record()
        g.replay()
        b.record()
    torch.cuda.synchronize()
    ms = sorted(a.elapsed_time(b) for a, b in evs)
    return ms[len(ms) // 2]


def autotune(P, sizes, inputs: dict, viewmats, campos, targets_flat, device="cuda:0", shapes=TILE_SHAPES,
             views_per_launch=(1, 2, 4, 8, 16, 32), reps: int = 5, mlp=None, log=None) -> TuneResult:
    """Times a full training step over all views for every (tile shape,
    views per launch) and returns the fastest configuration."""
    table = []
    best = None
    for tile in shapes:
        for vpl in views_per_launch:
            if vpl > len(sizes):
                continue
            rs, ts, lfs, dls = build_launches(P, sizes, tile, vpl, inputs, viewmats, campos, targets_flat, device, mlp)
            step = make_step(rs, ts, lfs, dls)
            g = capture(step)
            ms = time_graph(g, reps)
            if any(r.overflowed() for r in rs):
                raise RuntimeError("instance capacity overflow during timing")
            row = {"tile": f"{tile[0]}x{tile[1]}", "views_per_launch": vpl, "ms_per_step": round(ms, 3),
                   "views_per_sec": round(1000.0 * len(sizes) / ms, 1)}
            table.append(row)
            if log:
                log(row)
            if best is None or ms < best[0]:
                best = (ms, tile, vpl)
            del g, rs, ts, lfs, dls, step
            torch.cuda.synchronize()
            torch.cuda.empty_cache()
    ms, tile, vpl = best
    return TuneResult(tile, vpl, ms / len(sizes), 1000.0 * len(sizes) / ms, table)


def seconds():
    return time.perf_counter()
