"""The allocation-free loss entry point (omnigs_photometric_loss_noalloc)
against the workspace form (omnigs_photometric_loss), through the C ABI.

Both run the same kernels with deterministic reductions, so the loss and
dL/dimage are bit-identical.  Two calls in a row, the second on a smaller
image, check that nothing left in the shared device array from the first call
leaks into the second; a call that would need more than the array holds is
refused before any device work.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

LOSS_NOALLOC_MAX_BYTES = 80 << 20


def _lib():
    from omnigs_amd import capi
    return capi.load()


def test_noalloc_refuses_images_past_its_array():
    """No GPU needed: the size check comes before any HIP call."""
    from omnigs_amd import capi
    lib = _lib()
    C, H, W = 3, 2048, 4096
    assert lib.omnigs_loss_workspace_bytes(C, H, W, 0) > LOSS_NOALLOC_MAX_BYTES
    dummy = (ctypes.c_float * 4)()
    a = ctypes.addressof(dummy)
    st = lib.omnigs_photometric_loss_noalloc(a, a, None, C, H, W, 0, 0.2, a, a, None)
    assert st == capi.OMNIGS_ERR_INVALID_ARGUMENT
    assert b"too large" in lib.omnigs_last_error()
    # the bench's image size fits, with SSIM
    assert lib.omnigs_loss_workspace_bytes(3, 960, 1920, 0) <= LOSS_NOALLOC_MAX_BYTES


@pytest.mark.gpu
def test_noalloc_matches_workspace_form(gpu):
    import torch

    from omnigs_amd import capi
    from tests.gpu_util import announce
    lib = _lib()
    stream = capi.stream_handle(torch.device(gpu))
    gen = torch.Generator(device=gpu).manual_seed(7)
    cases = [  # (C, H, W, skip, lambda, mask): the second call is smaller than the first
        (3, 960, 1920, 48, 0.2, True),
        (3, 127, 255, 5, 0.2, False),
        (3, 960, 1920, 0, 0.0, False),
        (3, 64, 96, 0, 0.0, True),
    ]
    for C, H, W, skip, lam, use_mask in cases:
        announce(f"no-alloc loss {C}x{H}x{W} skip {skip} lambda {lam} mask {use_mask}")
        x = torch.rand((C, H, W), generator=gen, device=gpu)
        y = torch.rand((C, H, W), generator=gen, device=gpu)
        mask = (torch.rand((H, W), generator=gen, device=gpu) > 0.1).float() if use_mask else None
        p = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
        ws_bytes = lib.omnigs_loss_workspace_bytes(C, H, W, skip)
        ws = torch.empty((ws_bytes,), dtype=torch.uint8, device=gpu)
        d_ref = torch.empty_like(x)
        l_ref = torch.empty(3, device=gpu)
        st = lib.omnigs_photometric_loss(p(x), p(y), p(mask), C, H, W, skip, lam, p(ws), ws_bytes, p(d_ref),
                                         p(l_ref), stream)
        assert st == capi.OMNIGS_OK, lib.omnigs_last_error()
        torch.cuda.synchronize()
        d_na = torch.full_like(x, float("nan"))
        l_na = torch.full((3,), float("nan"), device=gpu)
        torch.cuda.reset_peak_memory_stats(gpu)
        before = torch.cuda.memory_allocated(gpu)
        st = lib.omnigs_photometric_loss_noalloc(p(x), p(y), p(mask), C, H, W, skip, lam, p(d_na), p(l_na), stream)
        assert st == capi.OMNIGS_OK, lib.omnigs_last_error()
        torch.cuda.synchronize()
        # nothing new through torch's allocator (the count may fall if the collector frees earlier tensors)
        assert torch.cuda.max_memory_allocated(gpu) <= before
        assert torch.equal(l_na, l_ref), (l_na.tolist(), l_ref.tolist())
        assert torch.equal(d_na, d_ref)
        assert np.isfinite(l_na.cpu().numpy()).all()
