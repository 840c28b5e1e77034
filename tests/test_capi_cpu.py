"""CPU-only checks of the C ABI library: it loads without a GPU, exports every
function include/omnigs_lonlat.h declares, and its host-side validation and
size queries behave (no GPU compute is called here)."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "omnigs_lonlat.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(omnigs_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from omnigs_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail("libomnigs_lonlat.so not built (run __graft_entry__.build())")
    return capi.load()


def test_exports_every_declared_symbol(lib):
    declared = _declared_functions()
    assert len(declared) >= 10
    from omnigs_amd import capi
    assert set(declared) == set(capi.EXPORTED_SYMBOLS), set(declared) ^ set(capi.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name


def test_version_and_status_strings(lib):
    assert b"gfx950" in lib.omnigs_version()
    assert lib.omnigs_status_string(0) == b"ok"
    assert b"precomputed" in lib.omnigs_status_string(2)


def test_buffer_sizes_scale(lib):
    g1, g2 = lib.omnigs_lonlat_geometry_bytes(1000), lib.omnigs_lonlat_geometry_bytes(2000)
    assert 60 * 1000 < g1 < g2
    assert lib.omnigs_lonlat_binning_bytes(0) > 0
    b1, b2 = lib.omnigs_lonlat_binning_bytes(10000), lib.omnigs_lonlat_binning_bytes(20000)
    assert 16 * 10000 <= b1 < b2
    assert lib.omnigs_lonlat_image_bytes(64, 32) >= 64 * 32 * 8 + 8 * 4 * 2
    assert lib.omnigs_lonlat_image_bytes(0, 32) == 0


def test_argument_validation_needs_no_gpu(lib):
    from omnigs_amd import capi
    a = capi.ForwardArgs()
    a.P, a.width, a.height = -1, 64, 32
    R = ctypes.c_int(7)
    null_buf = capi.Buffer(capi.RESIZE_FN(0), None)
    st = lib.omnigs_lonlat_forward(ctypes.byref(a), null_buf, null_buf, null_buf, None, ctypes.byref(R))
    assert st == capi.OMNIGS_ERR_INVALID_ARGUMENT and R.value == 0
    assert b"P < 0" in lib.omnigs_last_error()
    a.P = 0  # the reference skips the rasterizer for P == 0
    st = lib.omnigs_lonlat_forward(ctypes.byref(a), null_buf, null_buf, null_buf, None, ctypes.byref(R))
    assert st == capi.OMNIGS_OK and R.value == 0
    a.P = 10  # required pointers missing
    st = lib.omnigs_lonlat_forward(ctypes.byref(a), null_buf, null_buf, null_buf, None, ctypes.byref(R))
    assert st == capi.OMNIGS_ERR_INVALID_ARGUMENT
    # a tile grid wider than the expansion record holds is refused before any
    # device work (the pointers are never dereferenced)
    dummy = (ctypes.c_float * 64)()
    addr = ctypes.addressof(dummy)
    for f in ("background", "means3D", "shs", "opacities", "scales", "rotations", "viewmatrix", "campos",
              "out_color"):
        setattr(a, f, addr)
    a.D, a.M, a.width, a.height = 0, 1, 16 * 32767, 32  # 32767 tile columns, one past the limit
    st = lib.omnigs_lonlat_forward(ctypes.byref(a), null_buf, null_buf, null_buf, None, ctypes.byref(R))
    assert st == capi.OMNIGS_ERR_INVALID_ARGUMENT and b"tile grid" in lib.omnigs_last_error()
    # more tiles than a scan block can sum in 32 bits
    a.width, a.height = 16 * 2048, 16 * 1024  # 2^21 tiles
    st = lib.omnigs_lonlat_forward(ctypes.byref(a), null_buf, null_buf, null_buf, None, ctypes.byref(R))
    assert st == capi.OMNIGS_ERR_INVALID_ARGUMENT and b"2^21 tiles" in lib.omnigs_last_error()
    # seam wrap needs whole tiles across the width: at W = 1000 the 63-tile ring
    # would be 1008 px wide and instances near the seam would be lost
    a.width, a.height, a.seam_wrap = 1000, 500, 1
    st = lib.omnigs_lonlat_forward(ctypes.byref(a), null_buf, null_buf, null_buf, None, ctypes.byref(R))
    assert st == capi.OMNIGS_ERR_INVALID_ARGUMENT and b"multiple of the tile width" in lib.omnigs_last_error()
    a.seam_wrap = 0
    b = capi.BackwardArgs()
    b.P, b.R, b.width, b.height = 10, 5, 64, 32
    assert lib.omnigs_lonlat_backward(ctypes.byref(b), None) == capi.OMNIGS_ERR_INVALID_ARGUMENT
    for f in ("background", "means3D", "shs", "scales", "rotations", "viewmatrix", "campos", "dL_dpix",
              "dL_dmean2D", "dL_dconic", "dL_dopacity", "dL_dcolor", "dL_dmean3D", "dL_dcov3D", "dL_dsh",
              "dL_dscale", "dL_drot"):
        setattr(b, f, addr)
    b.geom_buffer = b.binning_buffer = b.image_buffer = addr
    b.D, b.M, b.width, b.height, b.seam_wrap = 0, 1, 1000, 500, 1
    assert lib.omnigs_lonlat_backward(ctypes.byref(b), None) == capi.OMNIGS_ERR_INVALID_ARGUMENT
    assert b"multiple of the tile width" in lib.omnigs_last_error()
    # the binning buffer's size names the forward's layout: required, refused
    # before any device work when it names none
    b.seam_wrap = 0
    for nb in (0, lib.omnigs_lonlat_binning_bytes(5) + 4):
        b.binning_bytes = nb
        assert lib.omnigs_lonlat_backward(ctypes.byref(b), None) == capi.OMNIGS_ERR_INVALID_ARGUMENT
        assert b"binning_bytes" in lib.omnigs_last_error()
    assert lib.omnigs_lonlat_mark_visible(-1, None, None) == capi.OMNIGS_ERR_INVALID_ARGUMENT
    assert lib.omnigs_lonlat_mark_visible(0, None, None) == capi.OMNIGS_OK


def test_state_view_offsets_are_inside_buffers(lib):
    from omnigs_amd import capi
    P, R, W, H = 1000, 5000, 200, 90
    base_g, base_b, base_i = 0x100000, 0x900000, 0x2000000
    sv = capi.StateView()
    assert lib.omnigs_lonlat_state(base_g, base_b, base_i, P, R, W, H, ctypes.byref(sv), 0) == 0
    gsz = lib.omnigs_lonlat_geometry_bytes(P)
    bsz = lib.omnigs_lonlat_binning_bytes(R)
    # a binning buffer laid out for a larger capacity (a forward's speculative
    # binning that covered R) is recognised by its size; other sizes refused
    assert lib.omnigs_lonlat_state(base_g, base_b, base_i, P, R, W, H, ctypes.byref(sv), bsz + 4) == \
        capi.OMNIGS_ERR_INVALID_ARGUMENT
    assert lib.omnigs_lonlat_state(base_g, base_b, base_i, P, R, W, H, ctypes.byref(sv), bsz) == 0
    isz = lib.omnigs_lonlat_image_bytes(W, H)
    for f in ("depth_key", "radii", "render_records", "clamped", "tiles_touched",
              "depth_order", "offsets"):
        v = getattr(sv, f)
        assert base_g <= v < base_g + gsz and v % 256 == 0, f
    for f in ("point_list", "tile_keys"):
        assert base_b <= getattr(sv, f) < base_b + bsz, f
    for f in ("final_T", "n_contrib", "ranges"):
        assert base_i <= getattr(sv, f) < base_i + isz, f


def test_ctypes_structs_match_header(tmp_path):
    """The ctypes mirrors in omnigs_amd.capi have the C header's sizes and
    field offsets (compiled with the host C compiler)."""
    import shutil
    import subprocess

    from omnigs_amd import capi
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    structs = {"omnigs_lonlat_forward_args": capi.ForwardArgs, "omnigs_lonlat_backward_args": capi.BackwardArgs,
               "omnigs_mlp_distill_args": capi.DistillArgs, "omnigs_adam_group": capi.AdamGroup,
               "omnigs_lonlat_state_view": capi.StateView, "omnigs_lonlat_batch_args": capi.BatchArgs,
               "omnigs_lonlat_batch_grads": capi.BatchGrads, "omnigs_lonlat_batch_state_view": capi.BatchStateView}
    lines = []
    for cname, ct in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in ct._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    src = tmp_path / "abi.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"omnigs_lonlat.h\"\nint main(void) {\n"
                   + "\n".join(lines) + "\nreturn 0;\n}\n")
    exe = tmp_path / "abi"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        name, field, val = line.split()
        got[(name, field)] = int(val)
    for cname, ct in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(ct), cname
        for fname, _ in ct._fields_:
            assert got[(cname, fname)] == getattr(ct, fname).offset, (cname, fname)
