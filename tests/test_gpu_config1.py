"""Config 1 on the GPU: the rasterizer's emission input and the full
single-receiver training step (20k Gaussians, shared 68->128->128->1 MLP,
0.8*L1 + 0.2*(1-SSIM)) against the CPU chain of the oracle's FP32 mirror (same
discrete decisions as the GPU) and a float64 MLP, with the MLP on the
library's kernels; and those kernels against the torch FP32 form of the same
MLP (tests/mlp_reference.py) on ragged row counts."""
import numpy as np
import pytest
import torch

import cuda_rasterizer as cr
import gsparc_emission as GE
import mlp_reference as MR
import gsparc_scenes as S
import oracle as O
import wireless_rasterizer as wr
from fused_ssim import fused_ssim
from helpers import to_dev
from test_config1 import _gradients, _mlp64

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("mode", ["exact", "reference"])
def test_emission_input_parity(mode):
    sc = S.small_scene(71, 3000, R=3, walls=True)
    d = to_dev(sc)
    rng = np.random.default_rng(71)
    E = rng.uniform(0.0, 1.0, (sc.R, sc.P, 2)).astype(np.float32)
    Et = torch.from_numpy(E).cuda()
    K, ore, oim, radii, geom, binbuf, img = wr.forward_batched(
        d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"], d["tx"], d["rx"],
        emission=Et)
    dre = rng.normal(size=(sc.R, sc.H, sc.W)).astype(np.float32)
    dim = rng.normal(size=(sc.R, sc.H, sc.W)).astype(np.float32)
    g = wr.backward_batched(d["means"], None, radii, d["opac"], d["scales"], d["rots"], 1.0, d["views"],
                            torch.from_numpy(dre).cuda(), torch.from_numpy(dim).cuda(), d["tx"], d["rx"], geom, K,
                            binbuf, img, grad_mode=mode, emission=Et)
    torch.cuda.synchronize()
    m = O.GRAD_EXACT if mode == "exact" else O.GRAD_REFERENCE
    for r in range(sc.R):
        pre = O.preprocess(sc, r)
        pre.sig_re[:] = E[r, :, 0]
        pre.sig_im[:] = E[r, :, 1]
        keys, vals, ranges = O.bin_gaussians(pre, sc.W, sc.H)
        f = O.render_fwd(pre, vals, ranges, sc.W, sc.H)
        peak = np.abs(f["out_re"]).max()
        assert np.abs(ore[r].cpu().numpy() - f["out_re"]).max() <= 1e-5 * peak
        assert np.abs(oim[r].cpu().numpy() - f["out_im"]).max() <= 1e-5 * np.abs(f["out_im"]).max()
        g8 = O.render_bwd(pre, vals, ranges, f, dre[r], dim[r], sc.W, sc.H, m)
        dem = g[7][r].cpu().numpy()
        vis = pre.radii > 0
        assert _rel(dem[vis, 0], g8[vis, 6]) < 1e-4
        assert _rel(dem[vis, 1], g8[vis, 7]) < 1e-4


def test_config1_training_step():
    sc = S.config1()
    tgt = S.config1(seed=2)
    d = to_dev(sc)
    dt = to_dev(tgt)
    mlp64 = _mlp64(3)
    mlp = GE.SharedEmissionMLP().cuda()
    with torch.no_grad():
        for p32, p64 in zip(mlp.parameters(), mlp64.parameters()):
            p32.copy_(p64.float())
        Et = mlp(dt["means"], dt["rx"], dt["tx"])
        target = wr.forward_batched(dt["means"], None, dt["opac"], dt["scales"], dt["rots"], tgt.H, tgt.W, 1.0,
                                    dt["views"], dt["tx"], dt["rx"], emission=Et)[1].clone()
    means = d["means"].clone().requires_grad_(True)
    opac = d["opac"].clone().requires_grad_(True)
    settings = cr.BatchedSetting(sc.H, sc.W, 1.0, d["views"], d["tx"], d["rx"], "exact")
    E = mlp(means, d["rx"], d["tx"])
    ore, oim, radii = cr.rasterize_gaussians_batched(means, None, opac, d["scales"], d["rots"], settings, emission=E)
    loss = 0.8 * (ore - target).abs().mean() + 0.2 * (1.0 - fused_ssim(ore.unsqueeze(1), target.unsqueeze(1)))
    loss.backward()
    torch.cuda.synchronize()
    ref = _gradients(sc, mlp64, target.cpu().numpy()[None].astype(np.float64), dtype=np.float32)
    # 1e-4 relative to the largest magnitude of each gradient (the MLP weight
    # gradients are FP32 GEMM reductions over the 20k rows).
    errs = {name: _rel(p.grad.cpu().numpy(), ref[name]) for name, p in mlp.named_parameters()}
    errs["means"] = _rel(means.grad.cpu().numpy(), ref["means"])
    errs["opacities"] = _rel(opac.grad.cpu().numpy(), ref["opacities"])
    print("relative errors", errs)
    for name, e in errs.items():
        assert e < 1e-4, (name, errs)


@pytest.mark.parametrize("R", [1, 4])
def test_graphed_shared_mlp_step_matches_autograd(R):
    """The graph-captured step (MLP forward, rasterizer forward with a
    capacity, fused loss, exact backward, MLP backward adding to the means
    gradient) equals the autograd step on every replay, with the means and
    the MLP weights changed in place between replays."""
    import gsparc_train as T
    sc = S.config1() if R == 1 else S.small_scene(72, 6000, R=R, walls=True)
    tgt = S.config1(seed=2) if R == 1 else S.small_scene(73, 6000, R=R, walls=True)
    d, dt = to_dev(sc), to_dev(tgt)
    mlp = GE.SharedEmissionMLP(generator=torch.Generator().manual_seed(5)).cuda()
    with torch.no_grad():
        Et = mlp(dt["means"], d["rx"], d["tx"])
        target = wr.forward_batched(dt["means"], None, dt["opac"], dt["scales"], dt["rots"], sc.H, sc.W, 1.0,
                                    d["views"], d["tx"], d["rx"], emission=Et)[1].clone()
        K = wr.forward_batched(d["means"], None, d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                               d["tx"], d["rx"], emission=mlp(d["means"], d["rx"], d["tx"]))[0]
    live = {k: d[k].clone() for k in ("means", "scales", "rots", "opac")}
    weights = [p.detach().clone().contiguous() for p in mlp.parameters()]
    print("graphed shared-MLP step R", R, "K", K, flush=True)
    step = T.GraphedSharedMlpStep(live["means"], live["scales"], live["rots"], live["opac"], weights, d["views"],
                                  d["tx"], d["rx"], target, sc.H, sc.W, capacity=int(1.25 * K) + 64)
    rng = np.random.default_rng(72)
    settings = cr.BatchedSetting(sc.H, sc.W, 1.0, d["views"], d["tx"], d["rx"], "exact")
    for rep in range(3):
        with torch.no_grad():
            live["means"].copy_(d["means"] + torch.from_numpy(
                rng.normal(scale=1e-3, size=tuple(d["means"].shape))).float().cuda())
            for w in weights:
                w.mul_(1.0 + 1e-3 * rep)
        step.replay()
        torch.cuda.synchronize()
        assert not step.overflowed()
        ref = GE.SharedEmissionMLP().cuda()
        with torch.no_grad():
            for p_, w in zip(ref.parameters(), weights):
                p_.copy_(w)
        means = live["means"].clone().requires_grad_(True)
        opac = live["opac"].clone().requires_grad_(True)
        scales = live["scales"].clone().requires_grad_(True)
        rots = live["rots"].clone().requires_grad_(True)
        E = ref(means, d["rx"], d["tx"])
        ore, oim, radii = cr.rasterize_gaussians_batched(means, None, opac, scales, rots, settings, emission=E)
        loss = 0.8 * (ore - target).abs().mean() + 0.2 * (1.0 - fused_ssim(ore.unsqueeze(1), target.unsqueeze(1)))
        loss.backward()
        torch.cuda.synchronize()
        assert torch.equal(step.out_real, ore.detach()), rep
        assert abs(float(step.loss) - float(loss.detach())) <= 1e-5 * abs(float(loss.detach())), rep
        for name, t_ in (("means", means), ("opacities", opac), ("scales", scales), ("rotations", rots)):
            assert _rel(step.grads[name].reshape(t_.shape).cpu().numpy(), t_.grad.cpu().numpy()) < 1e-4, (rep, name)
        gm = torch.cat([p_.grad.reshape(-1) for p_ in ref.parameters()])
        assert _rel(step.grads["mlp"].cpu().numpy(), gm.cpu().numpy()) < 1e-4, rep


@pytest.mark.parametrize("P,R", [(5000, 3), (37, 1), (300, 7), (130, 64), (1, 1)])
def test_mfma_mlp_matches_torch(P, R):
    """Amplitudes within 1e-5 and every gradient (weights, biases, means)
    within 1e-4 of the torch FP32 form; row counts not multiples of the
    kernels' 128 / 64-row tiles."""
    print("mfma mlp", P, R, flush=True)
    g = torch.Generator().manual_seed(P + R)
    means = (torch.rand((P, 3), generator=g) * torch.tensor([6.0, 6.0, 3.0])).cuda()
    rx = (0.5 + torch.rand((R, 3), generator=g) * 2.0).cuda()
    tx = torch.tensor([0.0, 0.5, 3.0]).cuda()
    ref = MR.TorchSharedMLP(generator=torch.Generator().manual_seed(5)).cuda()
    fast = GE.SharedEmissionMLP(generator=torch.Generator().manual_seed(5)).cuda()
    w = torch.randn((R, P), generator=g).cuda()
    out = {}
    for name, m in (("ref", ref), ("fast", fast)):
        mu = means.clone().requires_grad_(True)
        a = m.amplitude(mu, rx, tx)
        (a * w).sum().backward()
        out[name] = (a.detach(), mu.grad, {n: p.grad for n, p in m.named_parameters()})
    torch.cuda.synchronize()
    a_ref, gm_ref, gp_ref = out["ref"]
    a_fast, gm_fast, gp_fast = out["fast"]
    assert _rel(a_fast.cpu(), a_ref.cpu()) < 1e-5
    assert _rel(gm_fast.cpu(), gm_ref.cpu()) < 1e-4
    for n in gp_ref:
        assert _rel(gp_fast[n].cpu(), gp_ref[n].cpu()) < 1e-4, n
