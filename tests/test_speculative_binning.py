"""The forward bins against the previous call's capacity (with headroom) and
keeps that binning whenever it covers R; the backward recovers the layout
from the binning buffer's size.  A sequence of poses whose R rises and falls
(so that some forwards keep a larger layout and some bin again) must give the
same R, image, point list and gradients as a forward that bins exactly for R
(a fresh host thread has no previous R and always does)."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from omnigs_amd import scenes

pytestmark = pytest.mark.gpu

_GRADS = ("dL_dmeans2D", "dL_dconic", "dL_dopacity", "dL_dcolors", "dL_dmeans3D", "dL_dcov3D", "dL_dsh",
          "dL_dscales", "dL_drotations")


def _run(ds, centre, dL):
    import torch

    from omnigs_amd import capi
    from tests.gpu_util import gpu_state
    vm, cc = scenes.camera_at(np.asarray(centre, dtype=np.float64))
    d = dict(ds)
    d["viewmatrix"] = torch.from_numpy(vm).to(ds["means3D"].device)
    d["campos"] = torch.from_numpy(cc).to(ds["means3D"].device)
    R, img, radii, bufs = capi.forward_scene(d)
    g = capi.backward_scene(d, R, radii, bufs, dL)
    torch.cuda.synchronize()
    st = gpu_state(bufs, d["means3D"].shape[0], R, d["width"], d["height"])
    return R, img, st["point_list"], g, bufs[1].requested


def test_kept_speculative_binning_equals_exact(gpu):
    import torch

    from omnigs_amd import capi
    from tests.gpu_util import announce, to_dev
    announce("speculative binning kept across poses vs exact binning, 20k @ 512x256")
    sc = scenes.small_scene(P=20000, width=512, height=256, seed=21)
    ds = to_dev(sc)
    dL = torch.from_numpy(sc["dL_dpix"]).to(gpu)
    # camera centres moving towards and away from the Gaussians (R up and down)
    centres = [(0.0, 0.0, 0.0), (0.4, 0.0, 0.0), (0.0, 0.0, 0.0), (-0.3, 0.2, 0.1), (0.35, -0.1, 0.3),
               (0.05, 0.0, 0.0), (0.0, 0.0, 0.0)]
    seq = [_run(ds, c, dL) for c in centres]
    kept = sum(1 for R, _, _, _, nb in seq if nb != capi.load().omnigs_lonlat_binning_bytes(R))
    assert kept >= 2, kept  # the sequence exercises layouts larger than R's own
    for c, (R, img, pl, g, _) in zip(centres, seq):
        out = {}
        t = threading.Thread(target=lambda: out.update(r=_run(ds, c, dL)))
        t.start()
        t.join()
        R0, img0, pl0, g0, nb0 = out["r"]
        assert nb0 == capi.load().omnigs_lonlat_binning_bytes(R0)  # a fresh thread bins exactly
        assert R == R0
        assert torch.equal(img, img0)
        assert np.array_equal(pl, pl0)
        for k in _GRADS:
            a, b = g0[k], g[k]
            assert torch.allclose(a, b, rtol=1e-5, atol=1e-6 * float(a.abs().max())), (k, float((a - b).abs().max()))


def test_backward_refuses_a_foreign_binning_size(gpu):
    import torch

    from omnigs_amd import capi
    from tests.gpu_util import announce, to_dev
    announce("backward with a binning size that matches no layout")
    sc = scenes.small_scene(P=2000, width=128, height=64, seed=4)
    ds = to_dev(sc)
    R, img, radii, bufs = capi.forward_scene(ds)
    torch.cuda.synchronize()
    bufs[1].requested += 256  # a size no forward carved
    with pytest.raises(capi.OmnigsError, match="binning_bytes"):
        capi.backward_scene(ds, R, radii, bufs, torch.from_numpy(sc["dL_dpix"]).to(gpu))


def test_no_speculation_across_scenes(gpu):
    """A forward of another scene (another Gaussian count or image size) bins
    exactly for its R: its binning buffer is never grown for the previous
    scene's capacity."""
    import torch

    from omnigs_amd import capi
    from tests.gpu_util import announce, to_dev
    announce("no speculative binning across scenes: 200k @ 1024x512 then 2k @ 128x64")
    big = to_dev(scenes.small_scene(P=200_000, width=1024, height=512, seed=6))
    small = to_dev(scenes.small_scene(P=2000, width=128, height=64, seed=6))
    capi.forward_scene(big)
    capi.forward_scene(big)
    R, _, _, bufs = capi.forward_scene(small)
    torch.cuda.synchronize()
    exact = capi.load().omnigs_lonlat_binning_bytes(R)
    assert bufs[1].requested == exact
    assert bufs[1].tensor.numel() == exact  # the only size this fresh buffer was ever grown to
    # the same scene again keeps a speculative layout that covers R
    R2, _, _, bufs2 = capi.forward_scene(small)
    torch.cuda.synchronize()
    assert R2 == R and bufs2[1].requested >= exact
