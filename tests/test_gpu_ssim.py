"""GPU parity of the fused SSIM against the CPU oracle and against the
textbook conv2d SSIM (the reference's own test pattern,
fused-ssim/tests/test.py:14-53, 82-91)."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

import oracle as O
from fused_ssim import fused_ssim
import fused_ssim_cuda as fc

pytestmark = pytest.mark.gpu


def textbook_ssim(img1, img2):
    # window built as in fused-ssim/tests/test.py:14-16 (float32 samples / float32 sum)
    g = torch.Tensor([np.exp(-(x - 5) ** 2 / (2 * 1.5 ** 2)) for x in range(11)])
    g = (g / g.sum()).to(img1.dtype)
    w = (g[:, None] @ g[None, :]).expand(img1.size(1), 1, 11, 11).to(img1)
    ch = img1.size(1)
    mu1 = F.conv2d(img1, w, padding=5, groups=ch)
    mu2 = F.conv2d(img2, w, padding=5, groups=ch)
    s1 = F.conv2d(img1 * img1, w, padding=5, groups=ch) - mu1 ** 2
    s2 = F.conv2d(img2 * img2, w, padding=5, groups=ch) - mu2 ** 2
    s12 = F.conv2d(img1 * img2, w, padding=5, groups=ch) - mu1 * mu2
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    return ((2 * mu1 * mu2 + C1) * (2 * s12 + C2)) / ((mu1 ** 2 + mu2 ** 2 + C1) * (s1 + s2 + C2))


@pytest.mark.parametrize("shape", [(1, 1, 90, 360), (3, 2, 37, 53), (2, 1, 5, 7)])
def test_ssim_maps_match_oracle(shape):
    rng = np.random.default_rng(1)
    a = rng.uniform(size=shape).astype(np.float32)
    b = rng.uniform(size=shape).astype(np.float32)
    m, d1, d2, d3 = fc.fusedssim(O.C1, O.C2, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), True)
    om, o1, o2, o3 = O.ssim_fwd(a, b, dtype=np.float32)
    for x, y in ((m, om), (d1, o1), (d2, o2), (d3, o3)):
        np.testing.assert_allclose(x.cpu().numpy(), y, rtol=1e-5, atol=1e-6)
    g = rng.normal(size=shape).astype(np.float32)
    dimg = fc.fusedssim_backward(O.C1, O.C2, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(),
                                 torch.from_numpy(g).cuda(), d1, d2, d3)
    od = O.ssim_bwd(a, b, g, o1, o2, o3, dtype=np.float32)
    np.testing.assert_allclose(dimg.cpu().numpy(), od, rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("padding", ["same", "valid"])
def test_ssim_vs_textbook(padding):
    torch.manual_seed(0)
    a = torch.rand(2, 3, 90, 360, dtype=torch.float32)
    b = torch.rand(2, 3, 90, 360, dtype=torch.float32)
    x = a.cuda().requires_grad_(True)
    v = fused_ssim(x, b.cuda(), padding=padding)
    v.backward()
    ar = a.double().requires_grad_(True)
    ref_map = textbook_ssim(ar, b.double())
    if padding == "valid":
        ref_map = ref_map[:, :, 5:-5, 5:-5]
    ref = ref_map.mean()
    ref.backward()
    assert abs(float(v) - float(ref)) < 1e-5
    np.testing.assert_allclose(x.grad.cpu().numpy(), ar.grad.numpy(), rtol=1e-3, atol=1e-8)


def test_ssim_identical_images():
    a = torch.rand(1, 1, 90, 360).cuda().requires_grad_(True)
    v = fused_ssim(a, a.detach().clone())
    v.backward()
    assert abs(float(v) - 1.0) < 1e-5
    assert float(a.grad.abs().max()) < 1e-5
