"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of bench.py, never by the ``omnigs_amd`` product path.

The oracle restates the reference's equirectangular rasterizer
(cuda_rasterizer/forward.cu, backward.cu, rasterizer_impl.cu of OmniGS) on the
CPU in FP32 or FP64; see omnigs_oracle.hpp for the per-function citations.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

STAGES = {"preprocess": 0, "binning": 1, "render": 2, "render_backward": 3, "cov2d_backward": 4,
          "preprocess_backward": 5, "cov3d": 6}

# Precision code of the running-error oracle: FP64 values, each real array
# with a first-order bound of the rounding error an FP32 evaluation of the
# same operations accumulates (get("err:<name>"), in units of 2^-24).  It
# runs the backward stages only, on state set from outside.
RUNNING_ERROR = 65

# dtype of each named state array (None = the handle's float precision)
_INT_ARRAYS = {"radii": np.int32, "clamped": np.uint8, "tiles_touched": np.uint32, "offsets": np.uint32,
               "keys": np.uint64, "point_list": np.uint32, "ranges": np.uint32, "n_contrib": np.uint32,
               "fragile": np.uint8, "gsig": np.uint64, "psig": np.uint64}


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (g++ only), and the
    reference-side tools under _ref/ where the reference sources exist."""
    if force and os.path.exists(_LIB_PATH):
        os.remove(_LIB_PATH)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)
    return _LIB_PATH


PLY_REF = os.path.join(_HERE, "_ref", "ply_ref")


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    vp = ctypes.c_void_p
    lib.oracle_create.restype = vp
    lib.oracle_create.argtypes = [ctypes.c_int] * 8 + [vp] * 6 + [ctypes.c_double] + [vp] * 4
    lib.oracle_destroy.argtypes = [vp]
    lib.oracle_run.argtypes = [vp, ctypes.c_int, vp]
    lib.oracle_run.restype = ctypes.c_int
    lib.oracle_num_rendered.argtypes = [vp]
    lib.oracle_num_rendered.restype = ctypes.c_longlong
    lib.oracle_decision_sig.argtypes = [vp]
    lib.oracle_decision_sig.restype = ctypes.c_ulonglong
    lib.oracle_size.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
    lib.oracle_size.restype = ctypes.c_longlong
    lib.oracle_get.argtypes = [vp, ctypes.c_char_p, vp, ctypes.c_longlong]
    lib.oracle_get.restype = ctypes.c_int
    lib.oracle_set.argtypes = [vp, ctypes.c_char_p, vp, ctypes.c_longlong]
    lib.oracle_set.restype = ctypes.c_int
    lib.oracle_higher_msb.argtypes = [ctypes.c_uint]
    lib.oracle_higher_msb.restype = ctypes.c_int
    lib.oracle_set_seam_wrap.argtypes = [vp, ctypes.c_int]
    lib.oracle_set_threads.argtypes = [vp, ctypes.c_int]
    lib.oracle_set_threads.restype = ctypes.c_int
    lib.oracle_time_fwd_bwd.argtypes = [vp, vp]
    lib.oracle_time_fwd_bwd.restype = ctypes.c_double
    lib.oracle_fragile_pixels.argtypes = [vp]
    lib.oracle_fragile_pixels.restype = ctypes.c_longlong
    _lib = lib
    return lib


def higher_msb(n: int) -> int:
    return _load().oracle_higher_msb(n)


class Oracle:
    """One scene in one precision (32, 64, or RUNNING_ERROR).  Arrays are
    numpy, C-contiguous."""

    def __init__(self, scene: dict, precision: int = 32, block=(16, 16)):
        lib = _load()
        self.precision = precision
        self.ftype = np.float32 if precision == 32 else np.float64
        self.scene = scene
        self._keep = []

        def arr(name):
            a = scene.get(name)
            if a is None:
                return None
            a = np.ascontiguousarray(np.asarray(a, dtype=self.ftype))
            self._keep.append(a)
            return a.ctypes.data

        P = int(np.asarray(scene["means3D"]).shape[0])
        sh = scene.get("shs")
        M = int(np.asarray(sh).shape[1]) if sh is not None else 0
        self.P, self.M = P, M
        self.W, self.H = int(scene["width"]), int(scene["height"])
        self.h = lib.oracle_create(precision, P, int(scene.get("degree", 3)), M, self.W, self.H, block[0], block[1],
                                   arr("background"), arr("means3D"), arr("shs"), arr("colors_precomp"),
                                   arr("opacities"), arr("scales"), float(scene.get("scale_modifier", 1.0)),
                                   arr("rotations"), arr("cov3D_precomp"), arr("viewmatrix"), arr("campos"))
        self._keep = []
        if scene.get("seam_wrap"):
            lib.oracle_set_seam_wrap(self.h, 1)

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.oracle_destroy(h)
            self.h = None

    def run(self, stage: str, dL_dpix=None):
        ptr = None
        if dL_dpix is not None:
            dL_dpix = np.ascontiguousarray(np.asarray(dL_dpix, dtype=self.ftype))
            ptr = dL_dpix.ctypes.data
        rc = _load().oracle_run(self.h, STAGES[stage], ptr)
        if rc:
            raise RuntimeError(f"oracle stage {stage} failed ({rc})")
        return self

    def forward(self):
        return self.run("preprocess").run("binning").run("render")

    def backward(self, dL_dpix):
        return self.run("render_backward", dL_dpix).run("cov2d_backward").run("preprocess_backward")

    @property
    def num_rendered(self) -> int:
        return int(_load().oracle_num_rendered(self.h))

    @property
    def decision_sig(self) -> int:
        return int(_load().oracle_decision_sig(self.h))

    @property
    def fragile_pixels(self) -> int:
        """Pixels with a render-backward decision within rounding of its
        threshold (running-error handles only)."""
        return int(_load().oracle_fragile_pixels(self.h))

    def get(self, name: str) -> np.ndarray:
        lib = _load()
        eb = ctypes.c_int(0)
        n = lib.oracle_size(self.h, name.encode(), ctypes.byref(eb))
        if n < 0:
            raise KeyError(name)
        dt = _INT_ARRAYS.get(name, self.ftype)
        out = np.empty(n, dtype=dt)
        assert out.itemsize == eb.value, (name, out.itemsize, eb.value)
        rc = lib.oracle_get(self.h, name.encode(), out.ctypes.data, n)
        if rc:
            raise RuntimeError(f"oracle_get({name}) -> {rc}")
        return out

    def set(self, name: str, values) -> None:
        dt = _INT_ARRAYS.get(name, self.ftype)
        a = np.ascontiguousarray(np.asarray(values).reshape(-1).astype(dt, copy=False))
        rc = _load().oracle_set(self.h, name.encode(), a.ctypes.data, a.size)
        if rc:
            raise RuntimeError(f"oracle_set({name}) -> {rc}")

    def set_threads(self, n: int) -> int:
        """Threads of the OpenMP CPU baseline (1 = serial oracle); returns the count in effect."""
        return int(_load().oracle_set_threads(self.h, int(n)))

    def time_fwd_bwd(self, dL_dpix) -> float:
        d = np.ascontiguousarray(np.asarray(dL_dpix, dtype=self.ftype))
        return float(_load().oracle_time_fwd_bwd(self.h, d.ctypes.data))
