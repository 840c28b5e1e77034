"""Fused training loss: (1 - lambda) L1 + lambda (1 - SSIM) with bottom crop
and optional undistortion mask, and its image gradient (reference
src/gaussian_mapper.cpp:384-407, include/loss_utils.h:31-129).

CPU tests pin the FP64 oracle (oracle/loss.py):
  * the window against its closed form,
  * the grouped-conv SSIM against explicit zero-padded windows,
  * SSIM(x, x) = 1 with a vanishing gradient there,
  * the L1 term against its closed form,
  * the autograd gradient against central finite differences,
  * the crop against invariance to the dropped rows.
GPU tests compare omnigs_photometric_loss with the oracle.  The tolerances
are 1e-5 relative on the loss values, and on dL/dimage 1e-3 relative plus
1e-4 of the largest entry, leaving out pixels within 1e-5 of the L1 kink
|masked - gt| = 0.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

from oracle import loss as oloss


def _pair(C, H, W, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.0, 1.0, size=(C, H, W))
    b = np.clip(a + rng.normal(0.0, 0.15, size=a.shape), 0.0, 1.0)
    return a, b


def test_window_closed_form():
    g = oloss.gaussian_window_1d().numpy()
    z = sum(math.exp(-(t * t) / 4.5) for t in range(-5, 6))
    assert abs(g.sum() - 1.0) < 1e-15
    np.testing.assert_allclose(g, g[::-1], rtol=0, atol=0)
    assert abs(g[5] - 1.0 / z) < 1e-15
    assert abs(g[0] - math.exp(-25 / 4.5) / z) < 1e-15


@pytest.mark.parametrize("shape", [(2, 7, 9), (1, 4, 3), (3, 13, 12)])
def test_ssim_matches_explicit_windows(shape):
    a, b = _pair(*shape, seed=sum(shape))
    import torch
    got = float(oloss.ssim(torch.tensor(a), torch.tensor(b)))
    ref = oloss.ssim_brute_force(a, b)
    assert abs(got - ref) < 1e-12, (got, ref)


def test_ssim_of_identical_images_is_one_with_zero_gradient():
    a, _ = _pair(3, 16, 20, 4)
    loss, l1, s, g = oloss.photometric_loss(a, a, lambda_dssim=1.0)
    assert abs(s - 1.0) < 1e-12 and abs(loss) < 1e-12 and l1 == 0.0
    assert np.abs(g).max() < 1e-9  # SSIM is maximal at x = y and L1 has weight 0


def test_l1_closed_form():
    a, _ = _pair(3, 10, 12, 5)
    b = a - 0.25
    loss, l1, s, g = oloss.photometric_loss(a, b, lambda_dssim=0.0)
    assert abs(l1 - 0.25) < 1e-12 and loss == l1
    np.testing.assert_allclose(g, np.full_like(a, 1.0 / a.size), rtol=1e-12)


@pytest.mark.parametrize("lam,skip,masked", [(0.2, 0, False), (0.2, 3, True), (1.0, 2, False)])
def test_oracle_gradient_matches_finite_differences(lam, skip, masked):
    C, H, W = 2, 10, 9
    a, b = _pair(C, H, W, 7)
    mask = np.random.default_rng(8).uniform(0.5, 1.0, size=(H, W)) if masked else None
    _, _, _, g = oloss.photometric_loss(a, b, lam, skip, mask)
    rng = np.random.default_rng(9)
    h = 1e-6
    for _ in range(25):
        c, i, j = rng.integers(C), rng.integers(H), rng.integers(W)
        xm = a[c, i, j] * (mask[i, j] if masked else 1.0)
        if abs(xm - b[c, i, j]) < 1e-4:
            continue  # stay off the L1 kink
        ap, am = a.copy(), a.copy()
        ap[c, i, j] += h
        am[c, i, j] -= h
        fd = (oloss.photometric_loss(ap, b, lam, skip, mask)[0] - oloss.photometric_loss(am, b, lam, skip, mask)[0]) / (2 * h)
        assert abs(fd - g[c, i, j]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (c, i, j, fd, g[c, i, j])


def test_crop_ignores_dropped_rows():
    a, b = _pair(3, 20, 16, 11)
    skip = oloss.skip_bottom_rows(20, 0.2)
    assert skip == 4
    ref = oloss.photometric_loss(a, b, 0.2, skip)
    a2, b2 = a.copy(), b.copy()
    a2[:, -skip:, :] = 0.9
    b2[:, -skip:, :] = 0.1
    got = oloss.photometric_loss(a2, b2, 0.2, skip)
    assert got[0] == pytest.approx(ref[0], rel=1e-14)
    assert np.all(got[3][:, -skip:, :] == 0.0)
    # The crop removes rows from the SSIM support too: the kept image alone gives the same loss.
    alone = oloss.photometric_loss(a[:, :-skip], b[:, :-skip], 0.2, 0)
    assert alone[0] == pytest.approx(ref[0], rel=1e-14)


def test_capi_loss_validation_without_gpu():
    from omnigs_amd import capi
    lib = capi.load()
    assert lib.omnigs_loss_workspace_bytes(3, 10, 10, 10) == 0
    assert lib.omnigs_loss_workspace_bytes(3, 10, 10, -1) == 0
    assert lib.omnigs_loss_workspace_bytes(3, 960, 1920, 0) >= 3 * 3 * 960 * 1920 * 4
    fake = ctypes.c_void_p(16)
    # every argument check returns before any device work
    assert lib.omnigs_photometric_loss(None, fake, None, 3, 8, 8, 0, 0.2, fake, 1 << 20, fake, fake, None) \
        == capi.OMNIGS_ERR_INVALID_ARGUMENT
    assert lib.omnigs_photometric_loss(fake, fake, None, 3, 8, 8, 8, 0.2, fake, 1 << 20, fake, fake, None) \
        == capi.OMNIGS_ERR_INVALID_ARGUMENT
    assert lib.omnigs_photometric_loss(fake, fake, None, 3, 8, 8, 0, 1.5, fake, 1 << 20, fake, fake, None) \
        == capi.OMNIGS_ERR_INVALID_ARGUMENT
    assert lib.omnigs_photometric_loss(fake, fake, None, 3, 8, 8, 0, 0.2, fake, 16, fake, fake, None) \
        == capi.OMNIGS_ERR_INVALID_ARGUMENT
    assert capi.skip_bottom_rows(960, 0.1) == 96 and capi.skip_bottom_rows(960, 0.0) == 0


def _check_gpu(a, b, lam, ratio, mask, gpu):
    import torch
    from omnigs_amd import capi
    H = a.shape[1]
    skip = capi.skip_bottom_rows(H, ratio)
    ref = oloss.photometric_loss(a, b, lam, skip, mask)
    f = lambda v: torch.from_numpy(np.ascontiguousarray(v, dtype=np.float32)).to(gpu)  # noqa: E731
    lf = capi.PhotometricLoss(lam, ratio)
    # The oracle sees the same FP32-rounded inputs the kernel sees.
    a32, b32 = a.astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64)
    m32 = mask.astype(np.float32).astype(np.float64) if mask is not None else None
    ref = oloss.photometric_loss(a32, b32, lam, skip, m32)
    for _ in range(2):  # the second call reuses the workspace
        vals, g = lf(f(a), f(b), f(mask) if mask is not None else None)
        torch.cuda.synchronize()
        vals = vals.cpu().numpy()
        g = g.cpu().numpy()
        assert abs(vals[0] - ref[0]) <= 1e-5 * abs(ref[0]), (vals, ref[:3])
        assert abs(vals[1] - ref[1]) <= 1e-5 * abs(ref[1]), (vals, ref[:3])
        if lam:
            assert abs(vals[2] - ref[2]) <= 1e-5 * abs(ref[2]), (vals, ref[:3])
        xm = a32 * (m32 if m32 is not None else 1.0)
        far = np.abs(xm - b32) > 1e-5
        if skip:
            far[:, H - skip:, :] = True
            assert np.all(g[:, H - skip:, :] == 0.0)
        err = np.abs(g - ref[3])
        tol = 1e-3 * np.abs(ref[3]) + 1e-4 * np.abs(ref[3]).max()
        assert np.all(err[far] <= tol[far]), (err[far].max(), np.abs(ref[3]).max())


@pytest.mark.gpu
@pytest.mark.parametrize("shape,lam,ratio,masked", [
    ((3, 240, 480), 0.2, 0.0, False),
    ((3, 240, 480), 0.2, 0.1, True),
    ((3, 240, 480), 0.0, 0.0, False),
    ((3, 37, 53), 0.2, 0.05, False),
    ((1, 5, 4), 0.2, 0.0, True),
    ((3, 960, 1920), 0.2, 0.1, False),
])
def test_gpu_loss_matches_oracle(gpu, shape, lam, ratio, masked):
    from tests.gpu_util import announce
    announce(f"fused loss {shape} lambda={lam} skip_ratio={ratio} mask={masked}")
    a, b = _pair(*shape, seed=shape[1] + shape[2])
    mask = np.random.default_rng(3).uniform(0.0, 1.0, size=shape[1:]) if masked else None
    _check_gpu(a, b, lam, ratio, mask, gpu)


@pytest.mark.gpu
def test_gpu_loss_identical_images(gpu):
    import torch
    from omnigs_amd import capi
    from tests.gpu_util import announce
    announce("fused loss of identical images")
    a, _ = _pair(3, 64, 128, 2)
    x = torch.from_numpy(a.astype(np.float32)).to(gpu)
    vals, g = capi.PhotometricLoss(0.2)(x, x.clone())
    torch.cuda.synchronize()
    vals = vals.cpu().numpy()
    assert abs(vals[2] - 1.0) < 1e-5 and abs(vals[0]) < 1e-5 and vals[1] == 0.0
    assert float(g.abs().max()) < 1e-9
