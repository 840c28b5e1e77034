"""GPU-side helpers for the parity tests (call the HIP path through the C ABI)."""
from __future__ import annotations

import sys

import numpy as np
import torch

from omnigs_amd import capi


def announce(msg: str) -> None:
    """Names the case on the real stderr before a GPU step (survives capture)."""
    sys.__stderr__.write(f"[gpu] {msg}\n")
    sys.__stderr__.flush()


def to_dev(scene: dict, device="cuda:0") -> dict:
    out = {}
    for k, v in scene.items():
        if isinstance(v, np.ndarray):
            out[k] = torch.from_numpy(np.ascontiguousarray(v.astype(np.float32) if v.dtype == np.float64 else v)).to(
                device)
        else:
            out[k] = v
    return out


def run_forward(ds: dict):
    """Forward through the C ABI. Returns (R, image, radii, buffers)."""
    return capi.forward(background=ds["background"], means3D=ds["means3D"],
                        opacities=ds["opacities"].reshape(-1).contiguous(), viewmatrix=ds["viewmatrix"],
                        campos=ds["campos"], width=ds["width"], height=ds["height"], shs=ds.get("shs"),
                        colors_precomp=ds.get("colors_precomp"), scales=ds.get("scales"),
                        rotations=ds.get("rotations"), cov3D_precomp=ds.get("cov3D_precomp"),
                        degree=ds.get("degree", 3), scale_modifier=ds.get("scale_modifier", 1.0))


def run_backward(ds: dict, R, radii, bufs, dL_dpix, want_jacobian=True):
    return capi.backward(R=R, background=ds["background"], means3D=ds["means3D"], radii=radii,
                         viewmatrix=ds["viewmatrix"], campos=ds["campos"], dL_dpix=dL_dpix, buffers=bufs,
                         shs=ds.get("shs"), colors_precomp=ds.get("colors_precomp"), scales=ds.get("scales"),
                         rotations=ds.get("rotations"), cov3D_precomp=ds.get("cov3D_precomp"),
                         degree=ds.get("degree", 3), scale_modifier=ds.get("scale_modifier", 1.0),
                         want_jacobian=want_jacobian)


def gpu_state(bufs, P, R, W, H) -> dict:
    """Numpy copies of the forward's private arrays (decoded through the C ABI)."""
    st = capi.state(bufs, P, R, W, H)
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in st.items()}
    for k in ("depth_key", "tiles_touched", "depth_order", "offsets", "point_list", "tile_keys", "n_contrib",
              "ranges"):
        if k in out:
            out[k] = out[k].view(np.uint32)
    if "depth_order" in out and "depth_key" in out:
        # depth_key holds the keys in depth order after the sort; scatter back per Gaussian.
        per = np.empty(P, dtype=np.uint32)
        per[out["depth_order"]] = out["depth_key"]
        out["depth_bits"] = per
        out["depth"] = per.view(np.float32)
    return out


def np_(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


def record(case: str, **metrics) -> None:
    """Appends one JSON line of parity metrics to $OMNIGS_PARITY_RECORD (when
    set) and prints it, so a run leaves the measured errors behind."""
    import json
    import os
    line = json.dumps({"case": case, **{k: (float(v) if isinstance(v, (np.floating, np.integer)) else v)
                                        for k, v in metrics.items()}})
    print(line)
    path = os.environ.get("OMNIGS_PARITY_RECORD")
    if path:
        with open(path, "a") as f:
            f.write(line + "\n")


def fp64_errors(name: str, gpu, o64, floor_frac: float = 1e-2) -> dict:
    """Max |gpu - fp64| relative to the tensor's largest magnitude, and the
    max and 99th-percentile relative error over the entries at least
    floor_frac of that magnitude."""
    g = np.asarray(gpu, np.float64).ravel()
    r = np.asarray(o64, np.float64).ravel()
    d = np.abs(g - r)
    top = max(np.abs(r).max(), 1e-300) if r.size else 1.0
    big = np.abs(r) >= floor_frac * top
    rel = d[big] / np.abs(r[big]) if big.any() else np.zeros(1)
    return {f"{name}_abs_vs_fp64_of_max": float(d.max() / top) if d.size else 0.0,
            f"{name}_rel_vs_fp64_max": float(rel.max()), f"{name}_rel_vs_fp64_p99": float(np.percentile(rel, 99))}
