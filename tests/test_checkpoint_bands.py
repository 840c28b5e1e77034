"""Checkpointed backward in tile-row bands (omnigs_lonlat_backward_checkpointed)
against the stored backward (omnigs_lonlat_backward) of the same forward.

The banded form re-runs preprocess and the depth sort, then bins and runs the
render backward one band of tile rows at a time, adding each band's
per-Gaussian sums.  A tile's instance list is the same in both forms (every
Gaussian over the tile, in (depth, id) order), so the per-(tile, Gaussian)
terms agree; only the order in which a Gaussian's terms are summed across
bands (and the last bits of the render backward's in-tile LDS sums) differ,
hence the FP32 tolerance.  Band budgets from one band (the default at this
size) down to one tile row per band, clamp and seam-wrap modes, an all-culled
scene and a wrong R are covered.
"""
from __future__ import annotations

import numpy as np
import pytest

from omnigs_amd import scenes

pytestmark = pytest.mark.gpu

_GRADS = ("dL_dmeans2D", "dL_dconic", "dL_dopacity", "dL_dcolors", "dL_dmeans3D", "dL_dcov3D", "dL_dsh",
          "dL_dscales", "dL_drotations")


def _banded(ds, R, radii, image_buffer, dL, budget, wrap=False, blend=True):
    from omnigs_amd import capi
    return capi.backward(R=R, background=ds["background"], means3D=ds["means3D"], radii=radii,
                         viewmatrix=ds["viewmatrix"], campos=ds["campos"], dL_dpix=dL,
                         buffers=(None, None, image_buffer), shs=ds.get("shs"), scales=ds.get("scales"),
                         rotations=ds.get("rotations"), degree=ds.get("degree", 3), seam_wrap=wrap,
                         checkpointed={"opacities": ds["opacities"].reshape(-1).contiguous(),
                                       "band_instances": budget, "blend": blend})


def _expected_bands(tile_keys: np.ndarray, gx: int, gy: int, budget: int, R: int) -> int:
    """Greedy partition of the tile rows (whole rows, at most `budget`
    instances per band unless one row holds more), rows from the forward's
    own tile keys."""
    rows = np.bincount(tile_keys.astype(np.int64) // gx, minlength=gy)
    b = budget if budget else max(1 << 20, (R + 3) // 4)
    n_bands, y = 0, 0
    while y < gy:
        n = int(rows[y])
        y += 1
        while y < gy and n + int(rows[y]) <= b:
            n += int(rows[y])
            y += 1
        n_bands += n > 0
    return n_bands


@pytest.mark.parametrize("budget,wrap", [(0, False), (3000, False), (1, False), (3000, True)])
def test_banded_equals_stored(gpu, budget, wrap):
    import torch

    from omnigs_amd import capi
    from tests.gpu_util import announce, gpu_state, to_dev
    announce(f"banded checkpointed backward, budget {budget}, wrap {wrap}, 20k @ 512x256")
    sc = scenes.small_scene(P=20000, width=512, height=256, seed=7)
    ds = to_dev(sc)
    W, H = sc["width"], sc["height"]
    R, img, radii, bufs = capi.forward_scene(ds, seam_wrap=wrap)
    dL = torch.from_numpy(sc["dL_dpix"]).to(gpu)
    keys = gpu_state(bufs, 20000, R, W, H)["tile_keys"]
    ref = capi.backward_scene(ds, R, radii, bufs, dL, seam_wrap=wrap)
    got = _banded(ds, R, radii, bufs[2], dL, budget, wrap)
    torch.cuda.synchronize()
    gx, gy = (W + 15) // 16, (H + 15) // 16
    assert got["_bands"] == _expected_bands(keys, gx, gy, budget, R)
    if budget == 3000:
        assert got["_bands"] > 4
    # the largest band sizes the binning buffer: below the whole R when banded
    if got["_bands"] > 1:
        from omnigs_amd.capi import load
        assert got["_buffers"][1].tensor.numel() < load().omnigs_lonlat_binning_bytes(R)
    for k in _GRADS:
        a, b = ref[k], got[k]
        assert torch.allclose(a, b, rtol=1e-5, atol=1e-6 * float(a.abs().max())), (k, float((a - b).abs().max()))


def test_banded_twice_and_image_state_kept(gpu):
    """Two banded backwards of one forward agree with the stored backward run
    before them: the image buffer's per-pixel state is only read, and its
    per-tile ranges are rebuilt for every band."""
    import torch

    from omnigs_amd import capi
    from tests.gpu_util import announce, to_dev
    announce("stored backward, then banded twice, 20k @ 512x256")
    sc = scenes.small_scene(P=20000, width=512, height=256, seed=8)
    ds = to_dev(sc)
    R, img, radii, bufs = capi.forward_scene(ds)
    dL = torch.from_numpy(sc["dL_dpix"]).to(gpu)
    ref = capi.backward_scene(ds, R, radii, bufs, dL)
    g1 = _banded(ds, R, radii, bufs[2], dL, 2500)
    g2 = _banded(ds, R, radii, bufs[2], dL, 2500)
    torch.cuda.synchronize()
    for k in _GRADS:
        a = ref[k]
        for b in (g1[k], g2[k]):
            assert torch.allclose(a, b, rtol=1e-5, atol=1e-6 * float(a.abs().max())), k


def test_banded_all_culled_gives_zeros(gpu):
    import torch

    from omnigs_amd import capi
    from tests.gpu_util import announce, to_dev
    announce("banded backward, all Gaussians culled")
    sc = scenes.small_scene(P=300, width=64, height=32, seed=3)
    sc["means3D"] = (sc["means3D"] / np.linalg.norm(sc["means3D"], axis=1, keepdims=True) * 0.1).astype(np.float32)
    ds = to_dev(sc)
    R, img, radii, bufs = capi.forward_scene(ds)
    assert R == 0
    got = _banded(ds, R, radii, bufs[2], torch.ones((3, 32, 64), device=gpu), 0)
    torch.cuda.synchronize()
    assert got["_bands"] == 0
    for k in _GRADS:
        assert float(got[k].abs().sum()) == 0.0, k


def test_banded_rejects_wrong_R(gpu):
    import torch

    from omnigs_amd import capi
    from tests.gpu_util import announce, to_dev
    announce("banded backward, wrong R refused")
    sc = scenes.small_scene(P=2000, width=128, height=64, seed=5)
    ds = to_dev(sc)
    R, img, radii, bufs = capi.forward_scene(ds)
    with pytest.raises(capi.OmnigsError, match="differs"):
        _banded(ds, R + 1, radii, bufs[2], torch.from_numpy(sc["dL_dpix"]).to(gpu), 0)


def test_banded_with_decoder_colours(gpu):
    """Decoder colours (omnigs_mlp_*): the banded backward recomputes the
    preprocess with the decoder and ends with the decoder's backward, equal to
    the stored backward's (2D and covariance gradients to FP32 rounding, the
    decoder-dependent ones in norm)."""
    import torch

    from omnigs_amd import capi
    from tests.gpu_util import announce, to_dev
    from tests.test_mlp import _bf16_bits, om
    announce("banded backward with decoder colours, 20k @ 512x256")
    P = 20000
    sc = scenes.small_scene(P=P, width=512, height=256, seed=2)
    ds = to_dev(sc)
    rng = np.random.default_rng(3)
    lat_bits = torch.from_numpy(_bf16_bits(rng.normal(0, 0.7, size=(P, 16)).astype(np.float32))).to(gpu)
    packed = capi.mlp_pack(torch.from_numpy(om.init_theta(8)).to(gpu))
    common = dict(background=ds["background"], means3D=ds["means3D"], opacities=ds["opacities"].reshape(-1),
                  viewmatrix=ds["viewmatrix"], campos=ds["campos"], width=512, height=256, scales=ds["scales"],
                  rotations=ds["rotations"])
    R, img, radii, bufs = capi.forward(mlp=(packed, lat_bits), **common)
    dL = torch.from_numpy(sc["dL_dpix"]).to(gpu)
    bc = dict(background=ds["background"], means3D=ds["means3D"], viewmatrix=ds["viewmatrix"], campos=ds["campos"],
              dL_dpix=dL, scales=ds["scales"], rotations=ds["rotations"], mlp=(packed, lat_bits))
    ref = capi.backward(R=R, radii=radii, buffers=bufs, **bc)
    got = capi.backward(R=R, radii=radii, buffers=(None, None, bufs[2]),
                        checkpointed={"opacities": common["opacities"].contiguous(), "band_instances": 3000}, **bc)
    torch.cuda.synchronize()
    assert got["_bands"] > 4
    for k in ("dL_dmeans2D", "dL_dconic", "dL_dopacity", "dL_dcolors", "dL_dcov3D", "dL_dscales", "dL_drotations"):
        a, b = ref[k], got[k]
        assert torch.allclose(a, b, rtol=1e-5, atol=1e-6 * float(a.abs().max())), (k, float((a - b).abs().max()))
    # the decoder's backward rounds dL/dcolor-derived adjoints to BF16, so a
    # last-bit difference in a summed dL/dcolor can move one Gaussian's
    # contribution by a BF16 step: compared in norm
    for k in ("dL_dmeans3D", "dL_dlatent", "dL_dmlp"):
        a, b = ref[k], got[k]
        assert float((a - b).norm()) <= 1e-3 * float(a.norm()), (k, float((a - b).norm() / a.norm()))


@pytest.mark.parametrize("wrap", [False, True])
def test_forward_checkpointed_matches_forward(gpu, wrap):
    """omnigs_lonlat_forward_checkpointed: the same R, image, radii and
    per-pixel state as omnigs_lonlat_forward, tile lists and blend bytes
    bit-exact, with a binning buffer that leaves out the gradient records (17
    of 53 bytes per instance); its image buffer feeds the banded backward,
    which then equals the stored backward of the plain forward, with and
    without the blend bytes."""
    import torch

    from omnigs_amd import capi
    from tests.gpu_util import announce, gpu_state, to_dev
    announce(f"forward without gradient records, wrap {wrap}, 20k @ 512x256")
    P = 20000
    sc = scenes.small_scene(P=P, width=512, height=256, seed=11)
    ds = to_dev(sc)
    W, H = sc["width"], sc["height"]
    R, img, radii, bufs = capi.forward_scene(ds, seam_wrap=wrap)
    Rc, imgc, radiic, bufsc = capi.forward_scene(ds, seam_wrap=wrap, checkpointed=True)
    torch.cuda.synchronize()
    assert Rc == R
    assert torch.equal(img, imgc) and torch.equal(radii, radiic)
    st, stc = gpu_state(bufs, P, R, W, H), gpu_state(bufsc, P, R, W, H)
    for k in ("final_T", "n_contrib", "ranges", "point_list", "tile_keys"):
        assert np.array_equal(st[k], stc[k]), k
    full = capi.load().omnigs_lonlat_binning_bytes(R)
    got_bytes = bufsc[1].requested
    assert got_bytes < 0.4 * full, (got_bytes, full)
    # the render's blend bytes sit at the image buffer's end, equal to the
    # plain forward's (which keeps them in its binning buffer)
    off = capi.load().omnigs_lonlat_image_bytes(W, H) + 256
    assert bufsc[2].requested >= off + R
    tail = bufsc[2].tensor[off:off + R].cpu().numpy()
    assert np.array_equal(tail & 0xF, st["blend"][:R] & 0xF)
    dL = torch.from_numpy(sc["dL_dpix"]).to(gpu)
    ref = capi.backward_scene(ds, R, radii, bufs, dL, seam_wrap=wrap)
    # bands on the forward's blend bytes (twice: the image buffer is only
    # read), and on the ellipses' reach with the image size withheld
    runs = [_banded(ds, R, radii, bufsc[2], dL, 3000, wrap), _banded(ds, R, radii, bufsc[2], dL, 3000, wrap),
            _banded(ds, R, radii, bufsc[2], dL, 3000, wrap, blend=False)]
    torch.cuda.synchronize()
    for got in runs:
        assert got["_bands"] > 4
        for k in _GRADS:
            a, b = ref[k], got[k]
            assert torch.allclose(a, b, rtol=1e-5, atol=1e-6 * float(a.abs().max())), (k, float((a - b).abs().max()))
