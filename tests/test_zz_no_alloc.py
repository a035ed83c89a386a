"""The loss entry point that allocates nothing (gsparc_loss_l1_ssim_noalloc,
scratch in a library-owned device array) against the workspace form
(gsparc_loss_l1_ssim), through the C ABI.  Both run the same kernels in the
same fixed order, so loss and gradient must be bit-identical.  Two calls in a
row, the second on a smaller input, so that the second starts on scratch that
still holds the first call's contents."""
import ctypes as C

import pytest
import torch

import gsparc_hip as gh

pytestmark = pytest.mark.gpu

C1, C2, LAM = 0.01 ** 2, 0.03 ** 2, 0.2


def _inputs(R, H, W, seed):
    g = torch.Generator().manual_seed(seed)
    img = (torch.rand((R, H, W), generator=g) * 0.4).cuda()
    tgt = (torch.rand((R, H, W), generator=g) * 0.4).cuda()
    return img, tgt


def _workspace_form(img, tgt, grad):
    R, H, W = img.shape
    L = gh.lib()
    work = torch.empty(L.gsparc_loss_work_size(R, H, W), dtype=torch.uint8, device="cuda")
    loss = torch.empty(1, device="cuda")
    g = torch.empty_like(img) if grad else None
    gh.check(L.gsparc_loss_l1_ssim(R, H, W, LAM, C1, C2, gh.ptr(img), gh.ptr(tgt), gh.ptr(work), gh.ptr(loss),
                                   gh.ptr(g), gh.stream_handle()), "loss")
    return loss, g


def _noalloc_form(img, tgt, grad):
    R, H, W = img.shape
    loss = torch.empty(1, device="cuda")
    g = torch.empty_like(img) if grad else None
    gh.check(gh.lib().gsparc_loss_l1_ssim_noalloc(R, H, W, LAM, C1, C2, gh.ptr(img), gh.ptr(tgt), gh.ptr(loss),
                                                  gh.ptr(g), gh.stream_handle()), "noalloc loss")
    return loss, g


def test_noalloc_loss_equals_workspace_form():
    for i, (R, H, W) in enumerate([(64, 90, 360), (3, 37, 50)]):
        print("noalloc loss", R, H, W, flush=True)
        img, tgt = _inputs(R, H, W, 100 + i)
        l1, g1 = _workspace_form(img, tgt, True)
        l2, g2 = _noalloc_form(img, tgt, True)
        torch.cuda.synchronize()
        assert torch.equal(l1, l2), (R, H, W)
        assert torch.equal(g1, g2), (R, H, W)
        assert bool(torch.isfinite(g2).all())
    # loss only (no gradient maps written)
    l3, _ = _noalloc_form(img, tgt, False)
    torch.cuda.synchronize()
    assert torch.equal(l3, l1)


def test_noalloc_loss_rejects_sizes_beyond_its_scratch():
    L = gh.lib()
    dummy = C.c_void_p(256)
    # 87 spectra of 90 x 360 need more than the 32 MiB scratch: refused before any device work
    assert L.gsparc_loss_l1_ssim_noalloc(87, 90, 360, LAM, C1, C2, dummy, dummy, dummy, dummy, None) != 0
    assert b"bytes of loss scratch" in L.gsparc_last_error()
    assert L.gsparc_loss_l1_ssim_noalloc(86, 90, 360, LAM, C1, C2, None, dummy, dummy, dummy, None) != 0
    assert b"null pointer" in L.gsparc_last_error()
