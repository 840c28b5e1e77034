"""Builds the native parts in-tree (no JIT cache):

* ``omnigs_amd/lib/libomnigs_lonlat.so`` -- the gfx950 HIP kernels behind the
  C ABI of ``include/omnigs_lonlat.h`` (hipcc, one object per stage).
* ``omnigs_amd/lib/libomnigs_torch.so`` -- the LibTorch-ROCm boundary
  (RasterizeGaussiansCUDA / RasterizeGaussiansBackwardCUDA / markVisible and
  the autograd Function, with the reference's declarations), C++ over the C ABI.
* ``omnigs_amd/lib/_omnigs_torch*.so`` -- its Python module.
* ``tests/cpp/lib/_render_lonlat_caller*.so`` -- a test caller written against
  the reference's interface only, linked to libomnigs_torch.so.

Usage: ``python -m omnigs_amd.build [--hip-only]``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(ROOT, "_build", "obj")
HIP_LIB = os.path.join(LIBDIR, "libomnigs_lonlat.so")

HIP_SOURCES = ["preprocess.hip", "sort.hip", "render.hip", "preprocess_backward.hip", "loss.hip", "mlp.hip", "ply.hip", "train.hip", "capi.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed-FP32 VALU ops are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


# Per-source additions.  mlp.hip and preprocess.hip (whose decoder form
# evaluates the MLP inline): MFMA results in ordinary VGPRs.  Left to itself
# the compiler writes every MFMA result to accumulation registers and copies
# each back before use (150 copies per block in the training kernel, 156 per
# wave in the decoder preprocess).
HIP_FILE_FLAGS = {"mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
                  "preprocess.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"command failed ({r.returncode}): {cmd[0]} ... {cmd[-1]}")


def build_hip(jobs: int = 8) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, *HIP_FILE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


TORCH_LIB = os.path.join(LIBDIR, "libomnigs_torch.so")
CALLER_DIR = os.path.join(ROOT, "tests", "cpp")
CALLER_LIB_DIR = os.path.join(CALLER_DIR, "lib")


def _torch_flags(module_name: str | None):
    import torch
    from torch.utils import cpp_extension as ce

    torch_lib = os.path.join(os.path.dirname(torch.__file__), "lib")
    inc = ce.include_paths() + [sysconfig.get_paths()["include"], os.path.join(ROOT, "include"), "/opt/rocm/include"]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    flags = ["-O2", "-std=c++17", "-fPIC", f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-D__HIP_PLATFORM_AMD__=1",
             "-DUSE_ROCM=1", "-Wno-unused-variable", "-Wno-deprecated-declarations"]
    if module_name:
        flags += [f"-DTORCH_EXTENSION_NAME={module_name}", "-DTORCH_API_INCLUDE_EXTENSION_H"]
    return torch_lib, inc, flags


def _compile(srcs, objdir, flags, inc, jobs=4, extra_inc=()):
    cxx = os.environ.get("CXX", "g++")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src).replace(".cpp", ".o"))
        objs.append(obj)
        cmds.append([cxx, *flags, *[f"-I{p}" for p in (*extra_inc, *inc)], "-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    return objs, cxx


def build_torch_ext(jobs: int = 4) -> str | None:
    """The LibTorch boundary, C++ via g++ (no HIP device code):

    * libomnigs_torch.so -- RasterizeGaussiansCUDA / RasterizeGaussiansBackwardCUDA /
      markVisible and GaussianRasterizerFunction / GaussianRasterizer with the
      reference's declarations, plus the omnigs:: extensions; what a C++
      caller links (no Python in it);
    * _omnigs_torch*.so -- the Python module over it.
    Both link the HIP library with an $ORIGIN rpath."""
    src_dir = os.path.join(HERE, "torch_ext")
    torch_lib, inc, flags = _torch_flags(None)
    torch_libs = [f"-L{torch_lib}", "-lc10", "-ltorch", "-ltorch_cpu", "-lc10_hip", "-ltorch_hip"]
    objs, cxx = _compile([os.path.join(src_dir, f) for f in ("rasterize_points.cpp", "gaussian_rasterizer.cpp")],
                         os.path.join(OBJDIR, "torch_ext"), flags, inc, jobs)
    _run([cxx, "-shared", *objs, "-o", TORCH_LIB + ".tmp", *torch_libs, f"-L{LIBDIR}", "-lomnigs_lonlat",
          "-Wl,-soname,libomnigs_torch.so", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{torch_lib}"])
    os.replace(TORCH_LIB + ".tmp", TORCH_LIB)

    ext_suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(LIBDIR, "_omnigs_torch" + ext_suffix)
    _, _, mflags = _torch_flags("_omnigs_torch")
    mobjs, _ = _compile([os.path.join(src_dir, "bindings.cpp")], os.path.join(OBJDIR, "torch_module"), mflags, inc)
    _run([cxx, "-shared", *mobjs, "-o", out + ".tmp", *torch_libs, "-ltorch_python", f"-L{LIBDIR}", "-lomnigs_torch",
          "-lomnigs_lonlat", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{torch_lib}"])
    os.replace(out + ".tmp", out)
    return out


def build_caller_check() -> str:
    """tests/cpp/render_lonlat_caller.cpp, a caller written against the
    reference's interface alone (gaussian_rasterizer.h), as a Python module
    linked to libomnigs_torch.so, so that GPU tests can drive it.  Test
    infrastructure: the product never loads it."""
    torch_lib, inc, flags = _torch_flags("_render_lonlat_caller")
    src = os.path.join(CALLER_DIR, "render_lonlat_caller.cpp")
    os.makedirs(CALLER_LIB_DIR, exist_ok=True)
    # only the boundary headers (torch_ext/) and torch: no omnigs C ABI header on the include path
    objs, cxx = _compile([src], os.path.join(OBJDIR, "caller"), flags,
                         [p for p in inc if os.path.abspath(p) != os.path.join(ROOT, "include")],
                         extra_inc=(os.path.join(HERE, "torch_ext"),))
    ext_suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(CALLER_LIB_DIR, "_render_lonlat_caller" + ext_suffix)
    rel = os.path.relpath(LIBDIR, CALLER_LIB_DIR)
    _run([cxx, "-shared", *objs, "-o", out + ".tmp", f"-L{torch_lib}", "-lc10", "-ltorch", "-ltorch_cpu",
          "-ltorch_python", f"-L{LIBDIR}", "-lomnigs_torch", f"-Wl,-rpath,$ORIGIN/{rel}", f"-Wl,-rpath,{torch_lib}"])
    os.replace(out + ".tmp", out)
    return out


def clean(hip_only: bool = False) -> None:
    """Removes the objects and the libraries that the next build rewrites."""
    shutil.rmtree(os.path.join(ROOT, "_build"), ignore_errors=True)
    dirs = [LIBDIR] if hip_only else [LIBDIR, CALLER_LIB_DIR]
    for d in dirs:
        if os.path.isdir(d):
            for f in os.listdir(d):
                if f.endswith(".so") and (not hip_only or f == os.path.basename(HIP_LIB)):
                    os.remove(os.path.join(d, f))


def build_all(hip_only: bool = False) -> None:
    clean(hip_only)
    build_hip()
    if not hip_only:
        build_torch_ext()
        build_caller_check()


if __name__ == "__main__":
    build_all(hip_only="--hip-only" in sys.argv)
    print("built:", ", ".join(sorted(os.listdir(LIBDIR))))
