"""ctypes binding of the C ABI in include/omnigs_lonlat.h.

Device memory comes from torch (ROCm) tensors; the library itself is plain
HIP.  ``import torch`` happens before the library is loaded so that one HIP
runtime (torch's) serves both.  There is no fallback: if the library is
missing the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libomnigs_lonlat.so")

OMNIGS_OK = 0
OMNIGS_ERR_INVALID_ARGUMENT = 1
OMNIGS_ERR_NEED_COLORS = 2
OMNIGS_ERR_ALLOC = 3
OMNIGS_ERR_HIP = 4

EXPORTED_SYMBOLS = (
    "omnigs_lonlat_forward", "omnigs_lonlat_recompute", "omnigs_lonlat_backward", "omnigs_lonlat_mark_visible",
    "omnigs_lonlat_geometry_bytes", "omnigs_lonlat_binning_bytes", "omnigs_lonlat_image_bytes",
    "omnigs_lonlat_state", "omnigs_profile_enable", "omnigs_profile_read", "omnigs_stage_name",
    "omnigs_status_string", "omnigs_last_error", "omnigs_version", "omnigs_loss_workspace_bytes",
    "omnigs_photometric_loss", "omnigs_mlp_param_count", "omnigs_mlp_packed_bytes", "omnigs_mlp_pack",
    "omnigs_mlp_eval", "omnigs_mlp_workspace_bytes", "omnigs_mlp_distill_step", "omnigs_mlp_seeded_direction",
    "omnigs_adam_step", "omnigs_f32_to_bf16", "omnigs_lonlat_batch_workspace_bytes", "omnigs_lonlat_batch_forward",
    "omnigs_lonlat_batch_backward", "omnigs_ply_write", "omnigs_ply_count", "omnigs_ply_read",
    "omnigs_ply_last_error", "omnigs_densification_stats", "omnigs_expon_lr", "omnigs_lonlat_batch_state",
    "omnigs_photometric_loss_noalloc", "omnigs_lonlat_backward_checkpointed", "omnigs_lonlat_forward_checkpointed",
    "omnigs_mlp_seeded_view_centre",
)
NUM_STAGES = 11

_STATE_FIELDS = ("depth_key", "radii", "render_records", "clamped", "tiles_touched",
                 "depth_order", "offsets", "point_list", "tile_keys", "final_T", "n_contrib", "ranges", "blend")

_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
RESIZE_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t)


class Buffer(ctypes.Structure):
    _fields_ = [("resize", RESIZE_FN), ("ctx", ctypes.c_void_p)]


class ForwardArgs(ctypes.Structure):
    _fields_ = [
        ("P", ctypes.c_int), ("D", ctypes.c_int), ("M", ctypes.c_int),
        ("width", ctypes.c_int), ("height", ctypes.c_int),
        ("background", ctypes.c_void_p), ("means3D", ctypes.c_void_p), ("shs", ctypes.c_void_p),
        ("colors_precomp", ctypes.c_void_p), ("opacities", ctypes.c_void_p), ("scales", ctypes.c_void_p),
        ("scale_modifier", ctypes.c_float), ("rotations", ctypes.c_void_p), ("cov3D_precomp", ctypes.c_void_p),
        ("viewmatrix", ctypes.c_void_p), ("campos", ctypes.c_void_p), ("prefiltered", ctypes.c_int),
        ("out_color", ctypes.c_void_p), ("radii", ctypes.c_void_p), ("seam_wrap", ctypes.c_int),
        ("mlp_packed", ctypes.c_void_p), ("latents", ctypes.c_void_p),
    ]


class BackwardArgs(ctypes.Structure):
    _fields_ = [
        ("P", ctypes.c_int), ("D", ctypes.c_int), ("M", ctypes.c_int), ("R", ctypes.c_int),
        ("width", ctypes.c_int), ("height", ctypes.c_int),
        ("background", ctypes.c_void_p), ("means3D", ctypes.c_void_p), ("shs", ctypes.c_void_p),
        ("colors_precomp", ctypes.c_void_p), ("scales", ctypes.c_void_p), ("scale_modifier", ctypes.c_float),
        ("rotations", ctypes.c_void_p), ("cov3D_precomp", ctypes.c_void_p), ("viewmatrix", ctypes.c_void_p),
        ("campos", ctypes.c_void_p), ("radii", ctypes.c_void_p),
        ("geom_buffer", ctypes.c_void_p), ("binning_buffer", ctypes.c_void_p), ("image_buffer", ctypes.c_void_p),
        ("dL_dpix", ctypes.c_void_p),
        ("dL_dmean2D", ctypes.c_void_p), ("dL_dconic", ctypes.c_void_p), ("dL_dopacity", ctypes.c_void_p),
        ("dL_dcolor", ctypes.c_void_p), ("dL_dmean3D", ctypes.c_void_p), ("dL_dcov3D", ctypes.c_void_p),
        ("dL_dsh", ctypes.c_void_p), ("dL_dscale", ctypes.c_void_p), ("dL_drot", ctypes.c_void_p),
        ("dpx_dt", ctypes.c_void_p), ("dpy_dt", ctypes.c_void_p), ("seam_wrap", ctypes.c_int),
        ("mlp_packed", ctypes.c_void_p), ("latents", ctypes.c_void_p), ("dL_dlatent", ctypes.c_void_p),
        ("dL_dmlp", ctypes.c_void_p), ("mlp_workspace", ctypes.c_void_p), ("mlp_workspace_bytes", ctypes.c_size_t),
        ("binning_bytes", ctypes.c_size_t), ("image_bytes", ctypes.c_size_t),
    ]


class BatchArgs(ctypes.Structure):
    _fields_ = [
        ("P", ctypes.c_int), ("D", ctypes.c_int), ("M", ctypes.c_int), ("V", ctypes.c_int),
        ("tile_x", ctypes.c_int), ("tile_y", ctypes.c_int), ("seam_wrap", ctypes.c_int),
        ("widths", ctypes.c_void_p), ("heights", ctypes.c_void_p), ("viewmatrices", ctypes.c_void_p),
        ("campos", ctypes.c_void_p), ("background", ctypes.c_void_p), ("means3D", ctypes.c_void_p),
        ("shs", ctypes.c_void_p), ("colors_precomp", ctypes.c_void_p), ("opacities", ctypes.c_void_p),
        ("scales", ctypes.c_void_p), ("scale_modifier", ctypes.c_float), ("rotations", ctypes.c_void_p),
        ("cov3D_precomp", ctypes.c_void_p), ("mlp_packed", ctypes.c_void_p), ("latents", ctypes.c_void_p),
        ("capacity", ctypes.c_uint), ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t),
        ("out_color", ctypes.c_void_p), ("status", ctypes.c_void_p),
    ]


class BatchGrads(ctypes.Structure):
    _fields_ = [
        ("dL_dpix", ctypes.c_void_p), ("dL_dmean3D", ctypes.c_void_p), ("dL_dopacity", ctypes.c_void_p),
        ("dL_dcov3D", ctypes.c_void_p), ("dL_dsh", ctypes.c_void_p), ("dL_dscale", ctypes.c_void_p),
        ("dL_drot", ctypes.c_void_p), ("dL_dlatent", ctypes.c_void_p), ("dL_dmlp", ctypes.c_void_p),
        ("mlp_workspace", ctypes.c_void_p), ("mlp_workspace_bytes", ctypes.c_size_t),
    ]


class DistillArgs(ctypes.Structure):
    _fields_ = [
        ("P", ctypes.c_int), ("M", ctypes.c_int), ("shs", ctypes.c_void_p), ("latents", ctypes.c_void_p),
        ("mlp_packed", ctypes.c_void_p), ("seed", ctypes.c_uint), ("step", ctypes.c_int), ("num_dirs", ctypes.c_int),
        ("dL_dlatent", ctypes.c_void_p), ("dL_dmlp", ctypes.c_void_p), ("loss", ctypes.c_void_p),
        ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t),
        ("means3D", ctypes.c_void_p), ("view_box", ctypes.c_float * 6),
    ]


class AdamGroup(ctypes.Structure):
    _fields_ = [
        ("param", ctypes.c_void_p), ("grad", ctypes.c_void_p), ("exp_avg", ctypes.c_void_p),
        ("exp_avg_sq", ctypes.c_void_p), ("param_bf16", ctypes.c_void_p), ("numel", ctypes.c_size_t),
        ("lr", ctypes.c_float), ("eps", ctypes.c_float), ("step", ctypes.c_longlong),
    ]


class StateView(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in _STATE_FIELDS]


class BatchStateView(ctypes.Structure):
    _fields_ = [("s", StateView), ("dL_dmean2D", ctypes.c_void_p), ("dL_dconic", ctypes.c_void_p),
                ("dL_dopacity", ctypes.c_void_p), ("dL_dcolor", ctypes.c_void_p), ("key_shift", ctypes.c_uint),
                ("key_count", ctypes.c_uint)]


_lib = None


def load() -> ctypes.CDLL:
    """Loads libomnigs_lonlat.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m omnigs_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    lib.omnigs_lonlat_forward.argtypes = [ctypes.POINTER(ForwardArgs), Buffer, Buffer, Buffer, ctypes.c_void_p,
                                          _ip]
    lib.omnigs_lonlat_forward.restype = ctypes.c_int
    lib.omnigs_lonlat_forward_checkpointed.argtypes = lib.omnigs_lonlat_forward.argtypes
    lib.omnigs_lonlat_forward_checkpointed.restype = ctypes.c_int
    lib.omnigs_lonlat_recompute.argtypes = [ctypes.POINTER(ForwardArgs), Buffer, Buffer, ctypes.c_void_p,
                                            ctypes.c_void_p, _ip]
    lib.omnigs_lonlat_recompute.restype = ctypes.c_int
    lib.omnigs_lonlat_backward.argtypes = [ctypes.POINTER(BackwardArgs), ctypes.c_void_p]
    lib.omnigs_lonlat_backward_checkpointed.argtypes = [ctypes.POINTER(ForwardArgs), ctypes.POINTER(BackwardArgs),
                                                        Buffer, Buffer, ctypes.c_uint, ctypes.c_void_p, _ip]
    lib.omnigs_lonlat_backward_checkpointed.restype = ctypes.c_int
    lib.omnigs_lonlat_backward.restype = ctypes.c_int
    lib.omnigs_lonlat_mark_visible.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib.omnigs_lonlat_mark_visible.restype = ctypes.c_int
    for name in ("omnigs_lonlat_geometry_bytes", "omnigs_lonlat_binning_bytes"):
        getattr(lib, name).argtypes = [ctypes.c_int]
        getattr(lib, name).restype = ctypes.c_size_t
    lib.omnigs_lonlat_image_bytes.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.omnigs_lonlat_image_bytes.restype = ctypes.c_size_t
    lib.omnigs_lonlat_state.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(StateView),
                                        ctypes.c_size_t]
    lib.omnigs_lonlat_state.restype = ctypes.c_int
    lib.omnigs_profile_enable.argtypes = [ctypes.c_int]
    lib.omnigs_profile_enable.restype = ctypes.c_int
    lib.omnigs_profile_read.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
    lib.omnigs_profile_read.restype = ctypes.c_int
    lib.omnigs_stage_name.argtypes = [ctypes.c_int]
    lib.omnigs_stage_name.restype = ctypes.c_char_p
    lib.omnigs_status_string.argtypes = [ctypes.c_int]
    lib.omnigs_status_string.restype = ctypes.c_char_p
    lib.omnigs_last_error.restype = ctypes.c_char_p
    lib.omnigs_version.restype = ctypes.c_char_p
    lib.omnigs_loss_workspace_bytes.argtypes = [ctypes.c_int] * 4
    lib.omnigs_loss_workspace_bytes.restype = ctypes.c_size_t
    lib.omnigs_photometric_loss.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 4 + [
        ctypes.c_float, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.omnigs_photometric_loss.restype = ctypes.c_int
    lib.omnigs_photometric_loss_noalloc.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 4 + [
        ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.omnigs_photometric_loss_noalloc.restype = ctypes.c_int
    lib.omnigs_mlp_param_count.restype = ctypes.c_int
    lib.omnigs_mlp_packed_bytes.restype = ctypes.c_size_t
    lib.omnigs_mlp_pack.argtypes = [ctypes.c_void_p] * 3
    lib.omnigs_mlp_pack.restype = ctypes.c_int
    lib.omnigs_mlp_eval.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 5
    lib.omnigs_mlp_eval.restype = ctypes.c_int
    lib.omnigs_mlp_workspace_bytes.argtypes = [ctypes.c_int]
    lib.omnigs_mlp_workspace_bytes.restype = ctypes.c_size_t
    lib.omnigs_mlp_distill_step.argtypes = [ctypes.POINTER(DistillArgs), ctypes.c_void_p]
    lib.omnigs_mlp_distill_step.restype = ctypes.c_int
    lib.omnigs_mlp_seeded_direction.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                                ctypes.POINTER(ctypes.c_float)]
    lib.omnigs_mlp_seeded_direction.restype = None
    lib.omnigs_mlp_seeded_view_centre.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                                  ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    lib.omnigs_mlp_seeded_view_centre.restype = None
    lib.omnigs_adam_step.argtypes = [ctypes.POINTER(AdamGroup), ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_void_p]
    lib.omnigs_adam_step.restype = ctypes.c_int
    lib.omnigs_f32_to_bf16.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lib.omnigs_f32_to_bf16.restype = ctypes.c_int
    lib.omnigs_lonlat_batch_workspace_bytes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                        ctypes.c_int, ctypes.c_int, ctypes.c_uint]
    lib.omnigs_lonlat_batch_workspace_bytes.restype = ctypes.c_size_t
    lib.omnigs_lonlat_batch_forward.argtypes = [ctypes.POINTER(BatchArgs), ctypes.c_void_p]
    lib.omnigs_lonlat_batch_forward.restype = ctypes.c_int
    lib.omnigs_lonlat_batch_backward.argtypes = [ctypes.POINTER(BatchArgs), ctypes.POINTER(BatchGrads),
                                                 ctypes.c_void_p]
    lib.omnigs_lonlat_batch_backward.restype = ctypes.c_int
    lib.omnigs_lonlat_batch_state.argtypes = [ctypes.POINTER(BatchArgs), ctypes.POINTER(BatchStateView)]
    lib.omnigs_lonlat_batch_state.restype = ctypes.c_int
    lib.omnigs_densification_stats.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 6
    lib.omnigs_densification_stats.restype = ctypes.c_int
    lib.omnigs_expon_lr.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_int]
    lib.omnigs_expon_lr.restype = ctypes.c_float
    _lib = lib
    return lib


class OmnigsError(RuntimeError):
    def __init__(self, status: int, where: str):
        lib = load()
        msg = lib.omnigs_status_string(status).decode()
        detail = lib.omnigs_last_error().decode()
        super().__init__(f"{where}: {msg} ({detail})")
        self.status = status


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None or t.numel() == 0:
        return None
    assert t.is_contiguous(), "tensors passed to the C ABI must be contiguous"
    return t.data_ptr()


class TorchBuffer:
    """A uint8 device tensor grown by the library through the resize callback
    (the C form of the reference's resizeFunctional, rasterize_points.cu:41-47).
    The allocation only grows; ``requested`` is the size the library last
    asked for, which identifies the layout the forward chose (the backward's
    binning_bytes)."""

    def __init__(self, device: torch.device):
        self.device = device
        self.tensor = torch.empty(0, dtype=torch.uint8, device=device)
        self.requested = 0

        def _resize(_ctx, nbytes):
            try:
                if self.tensor.numel() < nbytes:
                    self.tensor = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
                self.requested = int(nbytes)
                return self.tensor.data_ptr() if nbytes else self.tensor.data_ptr() or None
            except Exception:  # allocation failure surfaces as OMNIGS_ERR_ALLOC
                return None

        self._cb = RESIZE_FN(_resize)
        self.c = Buffer(self._cb, None)


def stream_handle(device: torch.device | None = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def forward(*, background, means3D, opacities, viewmatrix, campos, width, height, shs=None, colors_precomp=None,
            scales=None, rotations=None, cov3D_precomp=None, degree=3, scale_modifier=1.0, prefiltered=False,
            buffers=None, seam_wrap=False, mlp=None, checkpointed=False):
    """Runs omnigs_lonlat_forward.  Returns (R, out_color [3,H,W], radii [P], (geom, binning, image)).
    mlp: (packed image, BF16 latents [P,16] as int16/uint16 tensor) to take colours from the decoder.
    checkpointed: run omnigs_lonlat_forward_checkpointed instead (no gradient
    records in the binning buffer; the image buffer serves the checkpointed
    backward)."""
    lib = load()
    dev = means3D.device
    P = means3D.shape[0]
    if mlp:
        shs = colors_precomp = None  # colours come from the decoder
    M = shs.shape[1] if (shs is not None and shs.numel()) else 0
    # every pixel and every radius is written when P > 0; P == 0 leaves zeros (reference behaviour)
    alloc = torch.empty if P > 0 else torch.zeros
    out = alloc((3, height, width), dtype=torch.float32, device=dev)
    radii = alloc((P,), dtype=torch.int32, device=dev)
    bufs = buffers or (TorchBuffer(dev), TorchBuffer(dev), TorchBuffer(dev))
    a = ForwardArgs(P, degree, M, width, height, _ptr(background), _ptr(means3D), _ptr(shs), _ptr(colors_precomp),
                    _ptr(opacities), _ptr(scales), float(scale_modifier), _ptr(rotations), _ptr(cov3D_precomp),
                    _ptr(viewmatrix), _ptr(campos), int(prefiltered), _ptr(out), _ptr(radii), int(bool(seam_wrap)),
                    _ptr(mlp[0]) if mlp else None, _ptr(mlp[1]) if mlp else None)
    R = ctypes.c_int(0)
    fn = lib.omnigs_lonlat_forward_checkpointed if checkpointed else lib.omnigs_lonlat_forward
    st = fn(ctypes.byref(a), bufs[0].c, bufs[1].c, bufs[2].c, stream_handle(dev), ctypes.byref(R))
    if st != OMNIGS_OK:
        raise OmnigsError(st, fn.__name__)
    return R.value, out, radii, bufs


def backward(*, R, background, means3D, radii, viewmatrix, campos, dL_dpix, buffers, shs=None,
             colors_precomp=None, scales=None, rotations=None, cov3D_precomp=None, degree=3, scale_modifier=1.0,
             want_jacobian=False, seam_wrap=False, mlp=None, mlp_workspace=None, checkpointed=None):
    """Runs omnigs_lonlat_backward; returns a dict of gradient tensors (plus
    dL_dlatent [P,16] and dL_dmlp [param_count] when mlp = (packed, latents)).
    checkpointed: {"opacities": [P] (activated), "band_instances": budget (0 =
    the library's default), "blend": False to withhold the image buffer's size,
    so that the bands select pairs by ellipse reach instead of the forward's
    blend bytes} runs omnigs_lonlat_backward_checkpointed instead,
    which needs only the forward's image buffer (buffers[2]); the dict then
    also holds "_bands" (bands run) and "_buffers" (the geometry and binning
    buffers it grew)."""
    lib = load()
    dev = means3D.device
    P = means3D.shape[0]
    if mlp:
        shs = colors_precomp = None  # colours come from the decoder
    H, W = dL_dpix.shape[1], dL_dpix.shape[2]
    M = shs.shape[1] if (shs is not None and shs.numel()) else 0
    f = dict(dtype=torch.float32, device=dev)
    g = {
        "dL_dmeans2D": torch.empty((P, 3), **f), "dL_dconic": torch.empty((P, 2, 2), **f),
        "dL_dopacity": torch.empty((P, 1), **f), "dL_dcolors": torch.empty((P, 3), **f),
        "dL_dmeans3D": torch.empty((P, 3), **f), "dL_dcov3D": torch.empty((P, 6), **f),
        # written in full by the kernels when their inputs are given, zeros otherwise
        "dL_dsh": (torch.empty if shs is not None else torch.zeros)((P, M, 3), **f),
        "dL_dscales": (torch.empty if scales is not None else torch.zeros)((P, 3), **f),
        "dL_drotations": (torch.empty if scales is not None else torch.zeros)((P, 4), **f),
    }
    if want_jacobian:
        g["dpx_dt"] = torch.empty((P, 3), **f)
        g["dpy_dt"] = torch.empty((P, 3), **f)
    if mlp:
        g["dL_dlatent"] = torch.empty((P, 16), **f)
        g["dL_dmlp"] = torch.empty((lib.omnigs_mlp_param_count(),), **f)
        nbytes = lib.omnigs_mlp_workspace_bytes(P)
        if mlp_workspace is None or mlp_workspace.numel() < nbytes:
            mlp_workspace = torch.empty((nbytes,), dtype=torch.uint8, device=dev)
    geom, binning, image = buffers
    geom_ptr = geom.tensor.data_ptr() if geom is not None else None
    bin_ptr = binning.tensor.data_ptr() if binning is not None and binning.tensor.numel() else None
    a = BackwardArgs(P, degree, M, R, W, H, _ptr(background), _ptr(means3D), _ptr(shs), _ptr(colors_precomp),
                     _ptr(scales), float(scale_modifier), _ptr(rotations), _ptr(cov3D_precomp), _ptr(viewmatrix),
                     _ptr(campos), _ptr(radii), geom_ptr, bin_ptr, image.tensor.data_ptr(),
                     _ptr(dL_dpix.contiguous()), _ptr(g["dL_dmeans2D"]), _ptr(g["dL_dconic"]),
                     _ptr(g["dL_dopacity"]), _ptr(g["dL_dcolors"]), _ptr(g["dL_dmeans3D"]), _ptr(g["dL_dcov3D"]),
                     _ptr(g["dL_dsh"]) if shs is not None else None,
                     _ptr(g["dL_dscales"]) if scales is not None else None,
                     _ptr(g["dL_drotations"]) if scales is not None else None,
                     _ptr(g.get("dpx_dt")), _ptr(g.get("dpy_dt")), int(bool(seam_wrap)),
                     _ptr(mlp[0]) if mlp else None, _ptr(mlp[1]) if mlp else None, _ptr(g.get("dL_dlatent")),
                     _ptr(g.get("dL_dmlp")), mlp_workspace.data_ptr() if mlp else None,
                     mlp_workspace.numel() if mlp else 0,
                     binning.requested if bin_ptr is not None else 0,
                     image.requested if checkpointed is None or checkpointed.get("blend", True) else 0)
    if checkpointed is not None:
        fa = ForwardArgs(P, degree, M, W, H, _ptr(background), _ptr(means3D), _ptr(shs), _ptr(colors_precomp),
                         _ptr(checkpointed["opacities"]), _ptr(scales), float(scale_modifier), _ptr(rotations),
                         _ptr(cov3D_precomp), _ptr(viewmatrix), _ptr(campos), 0, None, _ptr(radii),
                         int(bool(seam_wrap)), _ptr(mlp[0]) if mlp else None, _ptr(mlp[1]) if mlp else None)
        nb = ctypes.c_int(0)
        fresh = (TorchBuffer(dev), TorchBuffer(dev))
        st = lib.omnigs_lonlat_backward_checkpointed(ctypes.byref(fa), ctypes.byref(a), fresh[0].c, fresh[1].c,
                                                     int(checkpointed.get("band_instances", 0)), stream_handle(dev),
                                                     ctypes.byref(nb))
        if st != OMNIGS_OK:
            raise OmnigsError(st, "omnigs_lonlat_backward_checkpointed")
        g["_bands"] = nb.value
        g["_buffers"] = fresh
        return g
    st = lib.omnigs_lonlat_backward(ctypes.byref(a), stream_handle(dev))
    if st != OMNIGS_OK:
        raise OmnigsError(st, "omnigs_lonlat_backward")
    return g


def state(buffers, P: int, R: int, width: int, height: int) -> dict:
    """Views (torch tensors aliasing the private buffers) of the forward's
    internal arrays, decoded with omnigs_lonlat_state."""
    lib = load()
    geom, binning, image = buffers
    sv = StateView()
    bases = [b.tensor.data_ptr() if b.tensor.numel() else None for b in buffers]
    st = lib.omnigs_lonlat_state(bases[0], bases[1], bases[2], P, R, width, height, ctypes.byref(sv),
                                 binning.requested)
    if st != OMNIGS_OK:
        raise OmnigsError(st, "omnigs_lonlat_state")
    tiles = ((width + 15) // 16) * ((height + 15) // 16)
    N = width * height
    spec = {  # field: (buffer, dtype, count)
        "depth_key": (geom, torch.int32, P), "radii": (geom, torch.int32, P),
        "render_records": (geom, torch.float32, 12 * P),
        "clamped": (geom, torch.uint8, P), "tiles_touched": (geom, torch.int32, P),
        "depth_order": (geom, torch.int32, P), "offsets": (geom, torch.int32, P),
        "point_list": (binning, torch.int32, R), "tile_keys": (binning, torch.int32, R),
        "final_T": (image, torch.float32, N), "n_contrib": (image, torch.int32, N), "ranges": (image, torch.int32, 2 * tiles),
        "blend": (binning, torch.uint8, R),
    }
    out = {}
    for name, (buf, dt, n) in spec.items():
        addr = getattr(sv, name)
        if not addr or n == 0:
            continue
        off = addr - buf.tensor.data_ptr()
        nbytes = n * torch.empty((), dtype=dt).element_size()
        assert 0 <= off and off + nbytes <= buf.tensor.numel(), name
        out[name] = buf.tensor[off:off + nbytes].view(dt)
    if "render_records" in out:
        rec = out["render_records"].view(P, 12)
        out["means2D"] = rec[:, 0:2].reshape(-1)
        out["conic_opacity"] = rec[:, 2:6].reshape(-1)
        out["rgb"] = rec[:, 6:9].reshape(-1)
    return out


def mark_visible(P: int, device) -> torch.Tensor:
    lib = load()
    present = torch.zeros((P,), dtype=torch.bool, device=device)
    st = lib.omnigs_lonlat_mark_visible(P, _ptr(present), stream_handle(torch.device(device)))
    if st != OMNIGS_OK:
        raise OmnigsError(st, "omnigs_lonlat_mark_visible")
    return present


def _check(st: int, where: str) -> None:
    if st != OMNIGS_OK:
        raise OmnigsError(st, where)


def mlp_pack(theta: torch.Tensor, packed: torch.Tensor | None = None) -> torch.Tensor:
    """FP32 decoder parameters -> the kernels' BF16 operand image (uint8 tensor)."""
    lib = load()
    assert theta.dtype == torch.float32 and theta.numel() == lib.omnigs_mlp_param_count()
    if packed is None:
        packed = torch.empty((lib.omnigs_mlp_packed_bytes(),), dtype=torch.uint8, device=theta.device)
    _check(lib.omnigs_mlp_pack(_ptr(theta.contiguous()), packed.data_ptr(), stream_handle(theta.device)),
           "omnigs_mlp_pack")
    return packed


def mlp_eval(latents_bf16: torch.Tensor, dirs: torch.Tensor, packed: torch.Tensor) -> torch.Tensor:
    lib = load()
    N = dirs.shape[0]
    rgb = torch.empty((N, 3), dtype=torch.float32, device=dirs.device)
    _check(lib.omnigs_mlp_eval(N, _ptr(latents_bf16), _ptr(dirs.contiguous()), packed.data_ptr(), _ptr(rgb),
                               stream_handle(dirs.device)), "omnigs_mlp_eval")
    return rgb


def mlp_distill_step(shs, latents, packed, seed, step, num_dirs, dL_dlatent, dL_dmlp, loss, workspace,
                     means3D=None, view_box=None) -> None:
    """One distillation pass; means3D [P,3] with view_box (lo xyz, hi xyz)
    takes the directions from seeded view centres in the box."""
    lib = load()
    P, M = shs.shape[0], shs.shape[1]
    box = (ctypes.c_float * 6)(*([float(v) for v in view_box] if means3D is not None else [0.0] * 6))
    a = DistillArgs(P, M, _ptr(shs), _ptr(latents), packed.data_ptr(), int(seed) & 0xFFFFFFFF, int(step),
                    int(num_dirs), _ptr(dL_dlatent), _ptr(dL_dmlp), _ptr(loss), workspace.data_ptr(),
                    workspace.numel(), _ptr(means3D), box)
    _check(lib.omnigs_mlp_distill_step(ctypes.byref(a), stream_handle(shs.device)), "omnigs_mlp_distill_step")


def mlp_workspace(P: int, device) -> torch.Tensor:
    return torch.empty((load().omnigs_mlp_workspace_bytes(P),), dtype=torch.uint8, device=device)


def seeded_direction(seed: int, g: int, k: int):
    out = (ctypes.c_float * 3)()
    load().omnigs_mlp_seeded_direction(seed & 0xFFFFFFFF, g, k, out)
    return tuple(out)


def seeded_view_centre(seed: int, g: int, k: int, view_box):
    out = (ctypes.c_float * 3)()
    box = (ctypes.c_float * 6)(*[float(v) for v in view_box])
    load().omnigs_mlp_seeded_view_centre(seed & 0xFFFFFFFF, g, k, box, out)
    return tuple(out)


def adam_step(groups, beta1: float = 0.9, beta2: float = 0.999, device=None) -> None:
    """groups: dicts {param, grad, exp_avg, exp_avg_sq, lr, eps, step[, param_bf16]} of FP32 tensors."""
    lib = load()
    arr = (AdamGroup * len(groups))()
    for i, g in enumerate(groups):
        arr[i] = AdamGroup(_ptr(g["param"]), _ptr(g["grad"]), _ptr(g["exp_avg"]), _ptr(g["exp_avg_sq"]),
                           _ptr(g.get("param_bf16")), g["param"].numel(), float(g["lr"]), float(g.get("eps", 1e-15)),
                           int(g["step"]))
    dev = device or groups[0]["param"].device
    _check(lib.omnigs_adam_step(arr, len(groups), float(beta1), float(beta2), stream_handle(dev)), "omnigs_adam_step")


def f32_to_bf16(x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """BF16 bits (int16 tensor) of an FP32 tensor, rounded to nearest even on the GPU."""
    lib = load()
    if out is None:
        out = torch.empty(x.shape, dtype=torch.int16, device=x.device)
    _check(lib.omnigs_f32_to_bf16(_ptr(x.contiguous()), _ptr(out), x.numel(), stream_handle(x.device)),
           "omnigs_f32_to_bf16")
    return out


def skip_bottom_rows(height: int, skip_bottom_ratio: float) -> int:
    """Rows the mapper drops from the bottom: round(H * ratio)
    (reference src/gaussian_mapper.cpp:389-391; 0 when the ratio is 0)."""
    return int(round(height * skip_bottom_ratio)) if skip_bottom_ratio > 0.0 else 0


class PhotometricLoss:
    """Fused (1 - lambda) L1 + lambda (1 - SSIM) and dL/dimage over the C ABI
    (omnigs_photometric_loss), with the workspace kept between calls."""

    def __init__(self, lambda_dssim: float = 0.2, skip_bottom_ratio: float = 0.0):
        self.lambda_dssim = float(lambda_dssim)
        self.skip_bottom_ratio = float(skip_bottom_ratio)
        self._ws = None

    def __call__(self, image: torch.Tensor, gt_image: torch.Tensor, mask: torch.Tensor | None = None,
                 dL_dimage: torch.Tensor | None = None):
        """Returns (values [3] = {loss, L1, SSIM}, dL_dimage [C,H,W]), both on
        the image's device, queued on the current stream."""
        lib = load()
        if image.dim() != 3 or gt_image.shape != image.shape:
            raise ValueError(f"image {tuple(image.shape)} and gt_image {tuple(gt_image.shape)} must be equal [C,H,W]")
        if image.dtype != torch.float32 or gt_image.dtype != torch.float32:
            raise TypeError("image and gt_image must be float32")
        C, H, W = image.shape
        if mask is not None:
            mask = mask.reshape(H, W).to(torch.float32).contiguous()
        skip = skip_bottom_rows(H, self.skip_bottom_ratio)
        nbytes = lib.omnigs_loss_workspace_bytes(C, H, W, skip)
        if nbytes == 0:
            raise ValueError(f"no rows left: height {H}, skip_bottom {skip}")
        if self._ws is None or self._ws.numel() < nbytes or self._ws.device != image.device:
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=image.device)
        if dL_dimage is None:
            dL_dimage = torch.empty_like(image)
        out = torch.empty(3, dtype=torch.float32, device=image.device)
        image, gt_image = image.contiguous(), gt_image.contiguous()
        st = lib.omnigs_photometric_loss(_ptr(image), _ptr(gt_image), _ptr(mask), C, H, W, skip, self.lambda_dssim,
                                         self._ws.data_ptr(), self._ws.numel(), _ptr(dL_dimage), out.data_ptr(),
                                         stream_handle(image.device))
        if st != OMNIGS_OK:
            raise OmnigsError(st, "omnigs_photometric_loss")
        return out, dL_dimage


def scene_to_device(scene: dict, device="cuda:0") -> dict:
    """Moves the float arrays of a scene (numpy) to contiguous device tensors."""
    out = {}
    for k, v in scene.items():
        if hasattr(v, "dtype") and hasattr(v, "shape"):
            out[k] = torch.from_numpy(v).to(device).contiguous()
        else:
            out[k] = v
    return out


def forward_scene(ds: dict, **kw):
    """Convenience: forward of a device scene dict from omnigs_amd.scenes."""
    return forward(background=ds["background"], means3D=ds["means3D"], opacities=ds["opacities"].reshape(-1),
                   viewmatrix=ds["viewmatrix"], campos=ds["campos"], width=ds["width"], height=ds["height"],
                   shs=ds.get("shs"), scales=ds.get("scales"), rotations=ds.get("rotations"),
                   degree=ds.get("degree", 3), scale_modifier=ds.get("scale_modifier", 1.0), **kw)


def backward_scene(ds: dict, R, radii, buffers, dL_dpix, **kw):
    return backward(R=R, background=ds["background"], means3D=ds["means3D"], radii=radii,
                    viewmatrix=ds["viewmatrix"], campos=ds["campos"], dL_dpix=dL_dpix, buffers=buffers,
                    shs=ds.get("shs"), scales=ds.get("scales"), rotations=ds.get("rotations"),
                    degree=ds.get("degree", 3), scale_modifier=ds.get("scale_modifier", 1.0), **kw)


def profile_enable(on=True) -> None:
    """True: time every stage; False: none; a set of stage names: those."""
    if on is True:
        mask = (1 << NUM_STAGES) - 1
    elif not on:
        mask = 0
    else:
        mask = 0
        for name in on:
            mask |= 1 << stage_index(name)
    load().omnigs_profile_enable(mask)


def stage_index(name: str) -> int:
    lib = load()
    for i in range(NUM_STAGES):
        if lib.omnigs_stage_name(i).decode() == name:
            return i
    raise KeyError(name)


def profile_read() -> dict:
    """{stage name: (total ms, launches)} of the events recorded since the last read."""
    lib = load()
    ms = (ctypes.c_double * NUM_STAGES)()
    cnt = (ctypes.c_int * NUM_STAGES)()
    st = lib.omnigs_profile_read(ms, cnt)
    if st != OMNIGS_OK:
        raise OmnigsError(st, "omnigs_profile_read")
    return {lib.omnigs_stage_name(i).decode(): (ms[i], cnt[i]) for i in range(NUM_STAGES)}
