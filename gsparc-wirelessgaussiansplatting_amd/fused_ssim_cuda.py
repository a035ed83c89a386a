"""Native-level SSIM API: drop-in for the reference's `fused_ssim_cuda`
extension (fused-ssim/ext.cpp:4-7, ssim.h:7-26).

    fusedssim(C1, C2, img1, img2, train) -> (ssim_map, dm_dmu1, dm_dsigma1_sq, dm_dsigma12)
    fusedssim_backward(C1, C2, img1, img2, dL_dmap, dm_dmu1, dm_dsigma1_sq, dm_dsigma12) -> dL_dimg1

Images are [B, CH, H, W] float32 on a HIP device (ssim.cu:368-444); with
train=False the three derivative maps are empty tensors, as in the reference.
"""
from __future__ import annotations

import torch

import gsparc_hip as gh


def _planes(img):
    if img.dim() != 4:
        raise RuntimeError("fused SSIM expects [B, CH, H, W] images")
    B, CH, H, W = img.shape
    return B * CH, H, W


def fusedssim(C1, C2, img1, img2, train=True):
    planes, H, W = _planes(img1)
    a = gh.f32(img1)
    b = gh.f32(img2)
    gh.require_cuda(a, b)
    with torch.cuda.device(a.device):
        ssim_map = torch.empty_like(a)
        if train:
            d1, d2, d3 = torch.empty_like(a), torch.empty_like(a), torch.empty_like(a)
        else:
            d1 = d2 = d3 = torch.empty(0, device=a.device)
        gh.check(gh.lib().gsparc_ssim_forward(planes, H, W, float(C1), float(C2), gh.ptr(a), gh.ptr(b),
                                              gh.ptr(ssim_map), gh.ptr(d1) if train else None,
                                              gh.ptr(d2) if train else None, gh.ptr(d3) if train else None,
                                              gh.stream_handle(a.device)), "fusedssim")
    return ssim_map, d1, d2, d3


def fusedssim_backward(C1, C2, img1, img2, dL_dmap, dm_dmu1, dm_dsigma1_sq, dm_dsigma12):
    planes, H, W = _planes(img1)
    a, b = gh.f32(img1), gh.f32(img2)
    g = gh.f32(dL_dmap)
    d1, d2, d3 = gh.f32(dm_dmu1), gh.f32(dm_dsigma1_sq), gh.f32(dm_dsigma12)
    gh.require_cuda(a, b, g, d1, d2, d3)
    with torch.cuda.device(a.device):
        out = torch.empty_like(a)
        gh.check(gh.lib().gsparc_ssim_backward(planes, H, W, float(C1), float(C2), gh.ptr(a), gh.ptr(b), gh.ptr(g),
                                               gh.ptr(d1), gh.ptr(d2), gh.ptr(d3), gh.ptr(out),
                                               gh.stream_handle(a.device)), "fusedssim_backward")
    return out
