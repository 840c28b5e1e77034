"""3DGS PLY reader / writer (the layout of the reference's savePly / loadPly,
src/gaussian_model.cpp:860-1069).

Pins: a golden header; exact round trips; byte-identical files and exact
values against the reference's own tinyply (third_party/tinyply) driven the
way savePly / loadPly drive it (oracle/_ref/ply_ref, built where the
reference tree exists; those cases skip elsewhere); and a reader that takes
other property orders, types, byte orders and ASCII.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest
import torch

from omnigs_amd import ply

GOLDEN_HEADER_D3_P2 = (
    "ply\nformat binary_little_endian 1.0\nelement vertex 2\n"
    + "".join(f"property float {n}\n" for n in ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
              + [f"f_rest_{i}" for i in range(45)] + ["opacity", "scale_0", "scale_1", "scale_2",
                                                       "rot_0", "rot_1", "rot_2", "rot_3"])
    + "end_header\n")


def _params(P, deg, seed=0):
    g = torch.Generator().manual_seed(seed)
    K = (deg + 1) ** 2 - 1
    return {"xyz": torch.randn(P, 3, generator=g), "features_dc": torch.randn(P, 1, 3, generator=g),
            "features_rest": torch.randn(P, K, 3, generator=g), "opacity": torch.randn(P, 1, generator=g),
            "scaling": torch.randn(P, 3, generator=g), "rotation": torch.randn(P, 4, generator=g)}


def _rows(p, deg):
    """Rows of the layout from the parameter dict (channel-major f_rest)."""
    P = p["xyz"].shape[0]
    K = (deg + 1) ** 2 - 1
    rest = p["features_rest"].numpy().transpose(0, 2, 1).reshape(P, 3 * K)
    return np.concatenate([p["xyz"].numpy(), np.zeros((P, 3), np.float32), p["features_dc"].numpy().reshape(P, 3),
                           rest, p["opacity"].numpy(), p["scaling"].numpy(), p["rotation"].numpy()], 1)


def test_golden_header_and_body(tmp_path):
    p = _params(2, 3)
    f = tmp_path / "a.ply"
    ply.save_ply(f, p, 3)
    data = f.read_bytes()
    assert data.startswith(GOLDEN_HEADER_D3_P2.encode())
    body = np.frombuffer(data[len(GOLDEN_HEADER_D3_P2):], dtype="<f4").reshape(2, 62)
    np.testing.assert_array_equal(body, _rows(p, 3))


@pytest.mark.parametrize("P,deg", [(1000, 3), (1, 1), (0, 3), (37, 0)])
def test_round_trip_exact(tmp_path, P, deg):
    p = _params(P, deg, seed=P)
    f = tmp_path / "r.ply"
    ply.save_ply(f, p, deg)
    q = ply.load_ply(f, deg)
    for k in p:
        assert torch.equal(p[k].reshape(q[k].shape), q[k]), k


def _ref_tool():
    from oracle import PLY_REF, build
    build()
    if not os.path.exists(PLY_REF):
        pytest.skip("reference tinyply tool not built here (no reference tree)")
    return PLY_REF


def _raw(p, deg, path):
    P = p["xyz"].shape[0]
    K = (deg + 1) ** 2 - 1
    rows = _rows(p, deg)
    with open(path, "wb") as fh:
        fh.write(np.int64(P).tobytes() + np.int32(K).tobytes())
        for a in (rows[:, 0:3], rows[:, 6:9], rows[:, 9:9 + 3 * K], rows[:, 9 + 3 * K:10 + 3 * K],
                  rows[:, 10 + 3 * K:13 + 3 * K], rows[:, 13 + 3 * K:]):
            fh.write(np.ascontiguousarray(a, dtype="<f4").tobytes())


@pytest.mark.parametrize("P,deg", [(777, 3), (5, 1), (1, 0)])
def test_matches_reference_tinyply(tmp_path, P, deg):
    tool = _ref_tool()
    p = _params(P, deg, seed=11)
    raw = tmp_path / "in.raw"
    _raw(p, deg, raw)
    ref_file, my_file = tmp_path / "ref.ply", tmp_path / "mine.ply"
    subprocess.run([tool, "write", str(raw), str(ref_file)], check=True)
    ply.save_ply(my_file, p, deg)
    assert my_file.read_bytes() == ref_file.read_bytes()
    q = ply.load_ply(ref_file, deg)  # our reader on tinyply's file
    for k in p:
        assert torch.equal(p[k].reshape(q[k].shape), q[k]), k
    back = tmp_path / "back.raw"  # tinyply's reader on our file
    subprocess.run([tool, "read", str(my_file), str((deg + 1) ** 2 - 1), str(back)], check=True)
    assert back.read_bytes() == raw.read_bytes()


def _write_custom(path, rows, names, fmt, types):
    P = rows.shape[0]
    hdr = f"ply\nformat {fmt} 1.0\ncomment made by a test\nelement vertex {P}\n"
    hdr += "".join(f"property {t} {n}\n" for t, n in zip(types, names))
    hdr += "end_header\n"
    with open(path, "wb") as fh:
        fh.write(hdr.encode())
        if fmt == "ascii":
            for r in rows:
                fh.write((" ".join(repr(float(v)) for v in r) + "\n").encode())
        else:
            end = "<" if fmt == "binary_little_endian" else ">"
            dt = np.dtype([(n, end + ("f8" if t == "double" else "f4")) for t, n in zip(types, names)])
            rec = np.empty(P, dtype=dt)
            for i, n in enumerate(names):
                rec[n] = rows[:, i]
            fh.write(rec.tobytes())


@pytest.mark.parametrize("fmt", ["binary_little_endian", "binary_big_endian", "ascii"])
def test_reader_takes_other_orders_types_and_formats(tmp_path, fmt):
    deg = 1
    p = _params(9, deg, seed=3)
    rows = _rows(p, deg).astype(np.float64)
    names = (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"] + [f"f_rest_{i}" for i in range(9)]
             + ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])
    perm = np.random.default_rng(0).permutation(len(names))
    types = ["double" if i % 3 == 0 else "float" for i in range(len(names))]
    f = tmp_path / "c.ply"
    _write_custom(f, rows[:, perm], [names[i] for i in perm], fmt, types)
    q = ply.load_ply(f, deg)
    for k in p:
        np.testing.assert_allclose(q[k].numpy(), p[k].reshape(q[k].shape).numpy(), rtol=1e-7)


def test_reader_errors(tmp_path):
    with pytest.raises(RuntimeError, match="f_rest_"):
        # a degree-0 file (no f_rest properties) read as degree 3
        rows = _rows(_params(4, 0), 0)
        g = tmp_path / "d0.ply"
        _write_custom(g, rows, ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0",
                                "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"], "binary_little_endian",
                      ["float"] * 17)
        ply.load_ply(g, 3)
    with pytest.raises(RuntimeError, match="cannot open"):
        ply.load_ply(tmp_path / "missing.ply", 3)


@pytest.mark.gpu
def test_gpu_render_of_a_ply_in_the_reference_layout(gpu, tmp_path):
    """A scene written in the reference's layout (raw parameters: log scales,
    unnormalised quaternions, opacity logits, channel-major f_rest) is read
    back onto the GPU and rendered through renderLonlat's contract (model
    activations, autograd Function); image and gradients match the CPU oracle
    run on the original parameters, and the gradients reach the loaded
    tensors (end-to-end bars of the GPU parity tests)."""
    from oracle import Oracle
    from omnigs_amd import scenes
    from omnigs_amd.rasterizer import render_lonlat
    from tests.gpu_util import announce, np_
    from tests.helpers import activate64, chain_activations, tiny_scene
    from tests.test_gpu_parity import FLIP_GRAD_FRACTION, FLIP_MAX_ABS, FLIP_PIXEL_FRACTION, GRAD_RTOL, \
        GRAD_SCALE_ATOL, IMG_ATOL, _frac_outside
    announce("PLY in the reference layout -> GPU render vs oracle, 2000 Gaussians 256x128")
    sc = tiny_scene(P=2000, W=256, H=128, seed=12, log_scale=(-3.6, 0.3), radius=(1.5, 4.0))
    raw = {"xyz": torch.tensor(sc["means3D"], dtype=torch.float32),
           "features_dc": torch.tensor(sc["shs"][:, :1], dtype=torch.float32),
           "features_rest": torch.tensor(sc["shs"][:, 1:], dtype=torch.float32),
           "opacity": torch.tensor(sc["logits"], dtype=torch.float32),
           "scaling": torch.tensor(sc["log_scales"], dtype=torch.float32),
           "rotation": torch.tensor(sc["raw_rotations"], dtype=torch.float32)}
    f = tmp_path / "scene.ply"
    ply.save_ply(f, raw, 3)
    params = {k: v.requires_grad_(True) for k, v in ply.load_ply(f, 3, device=gpu).items()}
    for k in raw:
        assert torch.equal(params[k].detach().cpu(), raw[k]), k
    vm = torch.tensor(sc["viewmatrix"], dtype=torch.float32, device=gpu)
    cp = torch.tensor(sc["campos"], dtype=torch.float32, device=gpu)
    bg = torch.tensor(sc["background"], dtype=torch.float32, device=gpu)
    img, vs, vis, radii = render_lonlat(params, vm, cp, 128, 256, bg, 3)
    (img * torch.tensor(sc["dL_dpix"], dtype=torch.float32, device=gpu)).sum().backward()
    s32 = {k: v for k, v in activate64(sc).items() if k not in ("log_scales", "raw_rotations", "logits")}
    o = Oracle(s32, 32).forward().backward(sc["dL_dpix"])
    d = np.abs(np_(img).reshape(-1) - o.get("out_color"))
    assert np.mean(d > IMG_ATOL) <= FLIP_PIXEL_FRACTION and d.max() <= FLIP_MAX_ABS
    ch = chain_activations(sc, {k: o.get(k) for k in ("dL_dmeans3D", "dL_dscales", "dL_drot", "dL_dopacity",
                                                      "dL_dsh")})
    for mine, ref in (("xyz", "means3D"), ("scaling", "log_scales"), ("rotation", "raw_rotations"),
                      ("opacity", "logits")):
        f_out = _frac_outside(np_(params[mine].grad).reshape(-1), ch[ref].reshape(-1), GRAD_RTOL, GRAD_SCALE_ATOL)
        assert f_out <= FLIP_GRAD_FRACTION, (mine, f_out)
    sh = torch.cat([params["features_dc"].grad, params["features_rest"].grad], 1)
    assert _frac_outside(np_(sh).reshape(-1), ch["shs"].reshape(-1), GRAD_RTOL, GRAD_SCALE_ATOL) <= FLIP_GRAD_FRACTION
