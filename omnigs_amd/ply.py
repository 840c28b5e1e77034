"""3DGS PLY files through the library's native reader / writer
(omnigs_ply_*; the layout of the reference's GaussianModel::savePly /
loadPly, src/gaussian_model.cpp:860-1069)."""
from __future__ import annotations

import ctypes

import torch

from . import capi


def _lib():
    lib = capi.load()
    if not getattr(lib, "_ply_bound", False):
        f = ctypes.c_void_p
        lib.omnigs_ply_write.argtypes = [ctypes.c_char_p, ctypes.c_longlong, ctypes.c_int] + [f] * 6
        lib.omnigs_ply_write.restype = ctypes.c_int
        lib.omnigs_ply_count.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_longlong)]
        lib.omnigs_ply_count.restype = ctypes.c_int
        lib.omnigs_ply_read.argtypes = [ctypes.c_char_p, ctypes.c_longlong, ctypes.c_int] + [f] * 6
        lib.omnigs_ply_read.restype = ctypes.c_int
        lib.omnigs_ply_last_error.restype = ctypes.c_char_p
        lib._ply_bound = True
    return lib


def _check(st: int, what: str) -> None:
    if st != capi.OMNIGS_OK:
        raise RuntimeError(f"{what}: {_lib().omnigs_ply_last_error().decode()}")


def save_ply(path: str, params: dict, sh_degree: int = 3) -> None:
    """Writes raw parameters {xyz [P,3], features_dc [P,1,3], features_rest
    [P,K,3], opacity [P,1], scaling [P,3], rotation [P,4]} (any device)."""
    lib = _lib()
    t = {k: params[k].detach().to("cpu", torch.float32).contiguous()
         for k in ("xyz", "features_dc", "features_rest", "opacity", "scaling", "rotation")}
    P = t["xyz"].shape[0]
    ptr = lambda x: x.data_ptr() if x.numel() else None  # noqa: E731
    _check(lib.omnigs_ply_write(str(path).encode(), P, sh_degree, ptr(t["xyz"]), ptr(t["features_dc"]),
                                ptr(t["features_rest"]), ptr(t["opacity"]), ptr(t["scaling"]), ptr(t["rotation"])),
           f"omnigs_ply_write({path})")


def load_ply(path: str, sh_degree: int = 3, device="cpu") -> dict:
    """Reads a 3DGS PLY into raw parameter tensors (shapes as in save_ply)."""
    lib = _lib()
    P = ctypes.c_longlong(0)
    _check(lib.omnigs_ply_count(str(path).encode(), ctypes.byref(P)), f"omnigs_ply_count({path})")
    n = P.value
    K = (sh_degree + 1) ** 2 - 1
    out = {"xyz": torch.empty((n, 3)), "features_dc": torch.empty((n, 1, 3)), "features_rest": torch.empty((n, K, 3)),
           "opacity": torch.empty((n, 1)), "scaling": torch.empty((n, 3)), "rotation": torch.empty((n, 4))}
    ptr = lambda x: x.data_ptr() if x.numel() else None  # noqa: E731
    _check(lib.omnigs_ply_read(str(path).encode(), n, sh_degree, ptr(out["xyz"]), ptr(out["features_dc"]),
                               ptr(out["features_rest"]), ptr(out["opacity"]), ptr(out["scaling"]),
                               ptr(out["rotation"])), f"omnigs_ply_read({path})")
    return {k: v.to(device) for k, v in out.items()}
