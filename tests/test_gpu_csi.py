"""GPU parity of the multi-frequency (CSI) path against the CPU oracle.

Bar: radii, contributor counts and final T bit-exact; the 52-subcarrier maps
within 1e-5 of the peak (they are bit-exact in practice: same phasor code,
same operation order); gradients within 1e-4 of the largest magnitude of each
tensor; the loss within 1e-6 relative and its gradient within 1e-5."""
import numpy as np
import pytest
import torch

import cuda_rasterizer as cr
import gsparc_scenes as S
import gsparc_train as T
import oracle as O
import wireless_rasterizer as wr
from helpers import to_dev

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _fwd(sc, d, **kw):
    out = wr.forward_csi(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                         d["tx"], d["rx"], **kw)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("ppl", [0, 1, 2], ids=["mfma", "quarter_tile", "half_tile"])
@pytest.mark.parametrize("seed,P,R", [(101, 300, 1), (102, 1500, 3)])
def test_csi_forward_matches_oracle(seed, P, R, ppl):
    import gsparc_hip as gh
    sc = S.small_scene(seed, P, R=R, walls=True)
    sc.mlp *= 3.0
    d = to_dev(sc)
    print("csi forward", seed, P, R, ppl, flush=True)
    gh.set_option("csi_fwd_ppl", ppl)
    try:
        K, out, radii, geom, binbuf, img = _fwd(sc, d)
    finally:
        gh.set_option("csi_fwd_ppl", 0)
    out = out.cpu().numpy()
    for r in range(R):
        pre, plen, vals, ranges, fwd = O.render_csi(sc, r)
        np.testing.assert_array_equal(radii[r].cpu().numpy(), pre.radii)
        peak = max(np.abs(fwd["out"]).max(), 1e-30)
        assert np.abs(out[r] - fwd["out"]).max() <= 1e-5 * peak, r


def test_csi_zero_frequency_equals_single_channel():
    sc = S.small_scene(103, 1200, R=2)
    d = to_dev(sc)
    _, out, *_ = _fwd(sc, d, f0=0.0, df=0.0)
    one = wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                             d["tx"], d["rx"])
    torch.cuda.synchronize()
    for k in (0, 13, 51):
        assert torch.equal(out[..., k, 0], one[1])
        assert torch.equal(out[..., k, 1], one[2])


def test_csi_capacity_forward_equals_exact():
    sc = S.small_scene(104, 800, R=2)
    d = to_dev(sc)
    K, out, *_ = _fwd(sc, d)
    Kc, out2, *_ = _fwd(sc, d, capacity=K + 64)
    assert Kc == K + 64
    assert torch.equal(out, out2)


@pytest.mark.parametrize("seed,P,R", [(111, 400, 1), (112, 1500, 2)])
def test_csi_backward_matches_oracle(seed, P, R):
    sc = S.small_scene(seed, P, R=R, walls=True)
    sc.mlp *= 3.0
    d = to_dev(sc)
    print("csi backward", seed, P, R, flush=True)
    K, out, radii, geom, binbuf, img = _fwd(sc, d)
    rng = np.random.default_rng(seed)
    dout = rng.normal(size=(R, sc.H, sc.W, wr.CSI_NC, 2)).astype(np.float32)
    g = wr.backward_csi(d["means"], d["mlp"], radii, d["opac"], d["scales"], d["rots"], 1.0, d["views"],
                        torch.from_numpy(dout).cuda(), d["tx"], d["rx"], geom, K, binbuf, img)
    torch.cuda.synchronize()
    d2, d3, dmlp, dop, dcov, dsc, drot, _ = [None if x is None else x.cpu().numpy() for x in g]
    acc = O.Grads(sc.P)
    g2 = []
    for r in range(R):
        pre, plen, vals, ranges, fwd = O.render_csi(sc, r)
        _, g9 = O.backward_csi(sc, pre, plen, vals, ranges, fwd, dout[r], r, grads=acc)
        g2.append(np.stack([g9[:, 0] * (0.5 * sc.W), g9[:, 1] * (0.5 * sc.H), 0 * g9[:, 0]], 1))
    assert _rel(d2, np.stack(g2)) < 1e-4
    assert _rel(d3, acc.means) < 1e-4
    assert _rel(dmlp, acc.mlp) < 1e-4
    assert _rel(dop, acc.opacities) < 1e-4
    assert _rel(dcov, acc.cov3) < 1e-4
    assert _rel(dsc, acc.scales) < 1e-4
    assert _rel(drot, acc.rotations) < 1e-4


def test_csi_loss_matches_oracle():
    rng = np.random.default_rng(7)
    R, H, W = 2, 37, 50
    out = rng.normal(size=(R, H, W, 52, 2)).astype(np.float32) * 0.1
    tgt = rng.normal(size=(R, H, W, 52, 2)).astype(np.float32) * 0.1
    tgt[0, :2, :3] = out[0, :2, :3]        # equal values: sign(0) = 0
    loss, grad = T.csi_loss(torch.from_numpy(out).cuda(), torch.from_numpy(tgt).cuda())
    torch.cuda.synchronize()
    x = out.astype(np.float64)
    y = tgt.astype(np.float64)
    ma = np.sqrt((x ** 2).sum(-1)).transpose(0, 3, 1, 2)      # [R,52,H,W]
    mb = np.sqrt((y ** 2).sum(-1)).transpose(0, 3, 1, 2)
    n1, n2 = x.size, ma.size
    m, d1, d2, d3 = O.ssim_fwd(ma, mb, dtype=np.float64)
    ref = 0.8 * np.abs(x - y).sum() / n1 + 0.2 * (1.0 - m.sum() / n2)
    assert abs(float(loss) - ref) <= 1e-6 * abs(ref)
    dmag = O.ssim_bwd(ma, mb, np.full_like(ma, -0.2 / n2), d1, d2, d3, dtype=np.float64).transpose(0, 2, 3, 1)
    mag = ma.transpose(0, 2, 3, 1)[..., None]
    ref_g = 0.8 * np.sign(x - y) / n1 + np.where(mag > 0, dmag[..., None] * x / np.where(mag > 0, mag, 1), 0)
    assert _rel(grad.cpu().numpy(), ref_g) < 1e-5


def test_csi_autograd_step():
    """rasterize_csi + CSILoss through autograd equal backward_csi driven by
    the loss gradient."""
    sc = S.small_scene(121, 900, R=2)
    d = to_dev(sc)
    tgt = torch.randn((2, sc.H, sc.W, 52, 2), device="cuda") * 0.05
    means = d["means"].clone().requires_grad_(True)
    mlp = d["mlp"].clone().requires_grad_(True)
    opac = d["opac"].clone().requires_grad_(True)
    st = cr.CsiSetting(sc.H, sc.W, 1.0, d["views"], d["tx"], d["rx"])
    out, radii = cr.rasterize_csi(means, mlp, opac, d["scales"], d["rots"], st)
    loss = T.CSILoss.apply(out, tgt)
    loss.backward()
    _, gout = T.csi_loss(out.detach(), tgt)
    K, out2, radii2, geom, binbuf, img = _fwd(sc, d)
    assert torch.equal(out.detach(), out2)
    g = wr.backward_csi(d["means"], d["mlp"], radii2, d["opac"], d["scales"], d["rots"], 1.0, d["views"], gout,
                        d["tx"], d["rx"], geom, K, binbuf, img)
    torch.cuda.synchronize()
    assert _rel(means.grad.cpu(), g[1].cpu()) < 1e-4
    assert _rel(mlp.grad.cpu(), g[2].cpu()) < 1e-4
    assert _rel(opac.grad.cpu(), g[3].cpu()) < 1e-4


def test_graphed_csi_step_matches_eager():
    """The graph-captured CSI step (forward with a capacity, CSI loss, exact
    adjoint) equals the eager autograd step on every replay, with the
    parameters changed in place and eager work between replays."""
    R = 3
    sc = S.small_scene(131, 1200, R=R, walls=True)
    d = to_dev(sc)
    tg = S.small_scene(132, 1200, R=R, walls=True)
    dt = to_dev(tg)
    with torch.no_grad():
        target = _fwd(tg, dt)[1].clone()
    K = _fwd(sc, d)[0]
    live = {k: d[k].clone() for k in ("means", "scales", "rots", "opac", "mlp")}
    print("graphed csi step, K", K, flush=True)
    step = T.GraphedCsiStep(live["means"], live["scales"], live["rots"], live["opac"], live["mlp"], d["views"],
                            d["tx"], d["rx"], target, sc.H, sc.W, capacity=int(1.25 * K) + 64)
    rng = np.random.default_rng(131)
    for rep in range(3):
        with torch.no_grad():
            live["means"].copy_(d["means"] + torch.from_numpy(
                rng.normal(scale=2e-3, size=tuple(d["means"].shape))).float().cuda())
        step.replay()
        torch.cuda.synchronize()
        assert not step.overflowed()
        params = [live[k].clone().requires_grad_(True) for k in ("means", "scales", "rots", "opac", "mlp")]
        m, s_, q, o, w = params
        st = cr.CsiSetting(sc.H, sc.W, 1.0, d["views"], d["tx"], d["rx"])
        out, radii = cr.rasterize_csi(m, w, o, s_, q, st)
        loss = T.CSILoss.apply(out, target)
        loss.backward()
        torch.cuda.synchronize()
        assert torch.equal(step.out, out.detach()), rep
        assert torch.equal(step.radii, radii), rep
        assert abs(float(step.loss) - float(loss.detach())) <= 1e-6 * abs(float(loss.detach())), rep
        for k, p_ in zip(("means", "scales", "rotations", "opacities", "emission_mlps"), params):
            assert _rel(step.grads[k].reshape(p_.shape).cpu(), p_.grad.cpu()) < 1e-4, (rep, k)
        _fwd(tg, dt)   # unrelated eager work between replays
