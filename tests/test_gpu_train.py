"""The fused training loss and the HIP-graph-captured training step.

Loss: (1 - 0.2) * mean|img - target| + 0.2 * (1 - mean SSIM) and its gradient
against the CPU oracle's SSIM forward/backward (FP64 twin) plus the L1 term,
and against the same loss written with torch ops and fused_ssim.

Graph: every replay is checked against the eager torch-autograd step on the
same parameters (loss within 1e-5 relative, gradients within 1e-4 of the
largest magnitude), with eager work and synchronisations between replays and
parameters changed in place between them."""
import numpy as np
import pytest
import torch

import cuda_rasterizer as cr
import gsparc_scenes as S
import gsparc_train as T
import oracle as O
import wireless_rasterizer as wr
from fused_ssim import fused_ssim
from helpers import to_dev

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(b).double().cpu()
    return float((a - b).abs().max() / max(float(b.abs().max()), 1e-30))


def _torch_loss(img, tgt):
    l1 = (img - tgt).abs().mean()
    return 0.8 * l1 + 0.2 * (1.0 - fused_ssim(img.unsqueeze(1), tgt.unsqueeze(1)))


@pytest.mark.parametrize("R,H,W", [(1, 90, 360), (3, 37, 50), (2, 16, 32)])
def test_loss_value_and_gradient(R, H, W):
    print("loss", R, H, W, flush=True)
    g = torch.Generator().manual_seed(R * 1000 + H)
    img = torch.rand((R, H, W), generator=g) * 0.4
    tgt = torch.rand((R, H, W), generator=g) * 0.4
    tgt[0, :3, :5] = img[0, :3, :5]          # ties: sign(0) = 0
    loss, grad = T.l1_ssim_loss(img.cuda(), tgt.cuda())
    torch.cuda.synchronize()
    # oracle: FP64 SSIM forward/backward per plane plus the L1 term
    x = img.numpy().astype(np.float64)[:, None]
    y = tgt.numpy().astype(np.float64)[:, None]
    N = x.size
    m, d1, d2, d3 = O.ssim_fwd(x, y, dtype=np.float64)
    ref_loss = 0.8 * np.abs(x - y).sum() / N + 0.2 * (1.0 - m.sum() / N)
    ref_grad = 0.8 * np.sign(x - y) / N + O.ssim_bwd(x, y, np.full_like(x, -0.2 / N), d1, d2, d3, dtype=np.float64)
    assert abs(float(loss) - ref_loss) <= 1e-6 * abs(ref_loss)
    assert _rel(grad, torch.from_numpy(ref_grad[:, 0])) < 1e-5
    # same loss written with torch ops
    ic = img.cuda().requires_grad_(True)
    tl = _torch_loss(ic, tgt.cuda())
    tl.backward()
    assert abs(float(loss) - float(tl)) <= 1e-5 * abs(float(tl))
    assert _rel(grad, ic.grad) < 1e-5
    # autograd form
    ic2 = img.cuda().requires_grad_(True)
    l2 = T.L1SSIMLoss.apply(ic2, tgt.cuda())
    (3.0 * l2).backward()
    assert float(l2) == float(loss)
    assert _rel(ic2.grad, 3.0 * grad) < 1e-6


def _step_setup(seed=61, P=2000, R=4):
    sc = S.small_scene(seed, P, R=R, walls=True)
    d = to_dev(sc)
    tgt = S.small_scene(seed + 1, P, R=R, walls=True)
    dt = to_dev(tgt)
    with torch.no_grad():
        target = wr.forward_batched(dt["means"], dt["mlp"], dt["opac"], dt["scales"], dt["rots"], sc.H, sc.W, 1.0,
                                    d["views"], d["tx"], d["rx"])[1].clone()
    return sc, d, target


def _eager_reference(sc, d, vals, target):
    names = ("means", "scales", "rots", "opac", "mlp")
    params = [v.clone().requires_grad_(True) for v in vals]
    st = cr.BatchedSetting(sc.H, sc.W, 1.0, d["views"], d["tx"], d["rx"], "reference")
    m, s_, q, o, w = params
    ore, oim, radii = cr.rasterize_gaussians_batched(m, w, o, s_, q, st)
    loss = _torch_loss(ore, target)
    loss.backward()
    torch.cuda.synchronize()
    return float(loss), ore.detach(), dict(zip(names, [p_.grad for p_ in params]))


def test_graphed_training_step_matches_eager():
    sc, d, target = _step_setup()
    K = wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                           d["tx"], d["rx"])[0]
    live = dict(means=d["means"].clone(), scales=d["scales"].clone(), rots=d["rots"].clone(),
                opac=d["opac"].clone(), mlp=d["mlp"].clone())
    print("capturing graphed step, K", K, flush=True)
    step = T.GraphedTrainingStep(live["means"], live["scales"], live["rots"], live["opac"], live["mlp"], d["views"],
                                 d["tx"], d["rx"], target, sc.H, sc.W, capacity=2 * K)
    rng = np.random.default_rng(5)
    sets = []
    for i in range(3):
        noise = torch.from_numpy(rng.normal(scale=2e-3, size=(sc.P, 3))).float().cuda()
        sets.append([d["means"] + noise, d["scales"], d["rots"], d["opac"] * (1.0 - 0.05 * i), d["mlp"]])
    refs = [_eager_reference(sc, d, vals, target) for vals in sets]
    gname = dict(means="means", scales="scales", rots="rotations", opac="opacities", mlp="emission_mlps")
    for rep in range(2):
        for i, vals in enumerate(sets):
            with torch.no_grad():
                for k, v in zip(("means", "scales", "rots", "opac", "mlp"), vals):
                    live[k].copy_(v)
            print("replay", rep, i, flush=True)
            step.replay()
            torch.cuda.synchronize()
            assert not step.overflowed()
            ref_loss, ref_out, ref_grads = refs[i]
            assert torch.equal(step.out_real, ref_out), (rep, i)
            assert abs(float(step.loss) - ref_loss) <= 1e-5 * abs(ref_loss), (rep, i)
            for k, gref in ref_grads.items():
                assert _rel(step.grads[gname[k]].reshape(gref.shape), gref) < 1e-4, (rep, i, k)
            # unrelated eager work between replays
            wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0,
                               d["views"], d["tx"], d["rx"])
            torch.cuda.synchronize()
    # the eager launch sequence gives the same spectra and loss as the replay
    step.run_eager()
    torch.cuda.synchronize()
    assert torch.equal(step.out_real, refs[2][1])
    step.close()


def test_graphed_step_reports_overflow():
    sc, d, target = _step_setup(seed=63, P=1500, R=3)
    K = wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                           d["tx"], d["rx"])[0]
    step = T.GraphedTrainingStep(d["means"], d["scales"], d["rots"], d["opac"], d["mlp"], d["views"], d["tx"],
                                 d["rx"], target, sc.H, sc.W, capacity=K // 3)
    step.replay()
    torch.cuda.synchronize()
    assert step.overflowed()
    assert step.instance_count() == K
    for g in step.grads.values():
        assert float(g.abs().max()) == 0.0
    step.close()


def _conf_inputs(R, H, W, seed):
    g = torch.Generator().manual_seed(seed)
    img = torch.rand((R, H, W), generator=g) * 0.4
    tgt = torch.rand((R, H, W), generator=g) * 0.4
    rx = torch.rand((R, 3), generator=g) * torch.tensor([6.0, 6.0, 3.0])
    torch.manual_seed(seed)
    net = T.ConfidenceMLP(32)
    with torch.no_grad():
        net.fc2.bias.fill_(0.3)          # away from C = 2 exactly
    return img, tgt, rx, net


@pytest.mark.parametrize("R,H,W", [(1, 90, 360), (5, 37, 50), (300, 12, 20)])
def test_confidence_loss_matches_torch_and_fp64(R, H, W):
    """mean_r (C_r L_r - alpha log C_r) (train.py:418-436) against (a) the
    same loss written with torch ops (L1, fused_ssim per spectrum and the
    ConfidenceMLP in torch) through autograd, and (b) an FP64 evaluation
    from the oracle's SSIM and the network's definition."""
    print("confidence loss", R, H, W, flush=True)
    img, tgt, rx, net = _conf_inputs(R, H, W, 300 + R)
    netg = T.ConfidenceMLP(32, device="cuda")
    netg.load_state_dict(net.state_dict())
    ic = img.cuda().requires_grad_(True)
    loss = T.confidence_loss(ic, tgt.cuda(), rx.cuda(), netg, alpha=0.3)
    loss.backward()
    torch.cuda.synchronize()
    # (a) torch composition
    net_t = T.ConfidenceMLP(32, device="cuda")
    net_t.load_state_dict(net.state_dict())
    it = img.cuda().requires_grad_(True)
    C = net_t(rx.cuda())
    terms = []
    for r in range(R):
        Lr = 0.8 * (it[r] - tgt[r].cuda()).abs().mean() + 0.2 * (1.0 - fused_ssim(it[r][None, None],
                                                                               tgt[r].cuda()[None, None]))
        terms.append(C[r] * Lr - 0.3 * torch.log(C[r]))
    lt = torch.stack(terms).mean()
    lt.backward()
    assert abs(float(loss) - float(lt)) <= 1e-5 * abs(float(lt))
    assert _rel(ic.grad, it.grad) < 1e-5
    for name in ("fc1.weight", "fc1.bias", "fc2.weight", "fc2.bias"):
        a = dict(netg.named_parameters())[name].grad
        b = dict(net_t.named_parameters())[name].grad
        assert _rel(a, b) < 1e-5, name
    # (b) FP64
    x = img.numpy().astype(np.float64)[:, None]
    y = tgt.numpy().astype(np.float64)[:, None]
    m, d1, d2, d3 = O.ssim_fwd(x, y, dtype=np.float64)
    W1 = net.fc1.weight.detach().double().numpy()
    b1 = net.fc1.bias.detach().double().numpy()
    w2 = net.fc2.weight.detach().double().numpy()[0]
    b2 = float(net.fc2.bias.detach()[0])
    rxd = rx.double().numpy()
    Cd = 1.0 + np.exp(np.maximum(rxd @ W1.T + b1, 0.0) @ w2 + b2)
    Ld = 0.8 * np.abs(x - y).mean(axis=(1, 2, 3)) + 0.2 * (1.0 - m.mean(axis=(1, 2, 3)))
    ref = float(np.mean(Cd * Ld - 0.3 * np.log(Cd)))
    assert abs(float(loss) - ref) <= 1e-6 * abs(ref)
    n = H * W
    gd = Cd[:, None, None] / R * (0.8 * np.sign(x - y)[:, 0] / n +
                                   O.ssim_bwd(x, y, np.full_like(x, -0.2 / n), d1, d2, d3, dtype=np.float64)[:, 0])
    assert _rel(ic.grad, torch.from_numpy(gd)) < 1e-5


def test_confidence_loss_without_network_is_the_plain_loss():
    img, tgt, rx, _ = _conf_inputs(4, 90, 360, 311)
    loss, dimg, dmlp, rec, conf = T.confidence_l1_ssim_loss(img.cuda(), tgt.cuda(), rx.cuda(), None)
    l0, g0 = T.l1_ssim_loss(img.cuda(), tgt.cuda())
    torch.cuda.synchronize()
    assert dmlp is None and bool((conf == 1.0).all())
    assert abs(float(loss) - float(l0)) <= 1e-6 * abs(float(l0))
    assert torch.equal(dimg, g0)
    # the per-spectrum losses average to the batch loss
    assert abs(float(rec.double().mean()) - float(l0)) <= 1e-6 * abs(float(l0))


def test_graphed_step_with_confidence_matches_eager():
    sc, d, target = _step_setup(seed=65, P=1500, R=3)
    K = wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                           d["tx"], d["rx"])[0]
    torch.manual_seed(5)
    net = T.ConfidenceMLP(32, device="cuda")
    packed = net.packed().detach().clone().contiguous()
    live = dict(means=d["means"].clone(), scales=d["scales"].clone(), rots=d["rots"].clone(),
                opac=d["opac"].clone(), mlp=d["mlp"].clone())
    step = T.GraphedTrainingStep(live["means"], live["scales"], live["rots"], live["opac"], live["mlp"], d["views"],
                                 d["tx"], d["rx"], target, sc.H, sc.W, capacity=2 * K, confidence=packed)
    # eager: autograd rasterizer + confidence loss
    params = [d[k].clone().requires_grad_(True) for k in ("means", "scales", "rots", "opac", "mlp")]
    m, s_, q, o, w = params
    st = cr.BatchedSetting(sc.H, sc.W, 1.0, d["views"], d["tx"], d["rx"], "reference")
    ore, oim, radii = cr.rasterize_gaussians_batched(m, w, o, s_, q, st)
    loss = T.confidence_loss(ore, target, d["rx"], net)
    loss.backward()
    for rep in range(2):
        step.replay()
        torch.cuda.synchronize()
        assert abs(float(step.loss) - float(loss)) <= 1e-5 * abs(float(loss)), rep
        for k, p_ in zip(("means", "scales", "rotations", "opacities", "emission_mlps"), params):
            assert _rel(step.grads[k].reshape(p_.shape), p_.grad) < 1e-4, (rep, k)
        assert _rel(step.grads["confidence"], torch.cat([net.fc1.weight.grad.reshape(-1), net.fc1.bias.grad,
                                                          net.fc2.weight.grad.reshape(-1), net.fc2.bias.grad])) < 1e-5
    step.close()
