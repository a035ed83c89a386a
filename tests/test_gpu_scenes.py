"""Uneven scenes in one launch sequence: a packed batch of S scenes, each with
its own Gaussian count, receiver count and transmitter, must give exactly what
each scene gives when rendered alone through the uniform batched path.

Bar: radii and spectra bit-exact (same float32 operation order per record);
gradients within 1e-4 of the largest magnitude (the blend backward adds with
atomics, whose order differs run to run)."""
import numpy as np
import pytest
import torch

import gsparc_hip as gh
import gsparc_scenes as S
from helpers import gpu_forward, to_dev
import wireless_rasterizer as wr

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-4
SHAPES = [(500, 3), (3000, 9), (1200, 1), (64, 2)]   # (P_s, R_s): R_s not a multiple of 8, uneven P_s


def _scenes():
    out = []
    for i, (P, R) in enumerate(SHAPES):
        sc = S.small_scene(70 + i, P, R=R, walls=(i % 2 == 1))
        sc.tx_pos = np.array([0.3 * i, 0.5 + 0.7 * i, 3.0 - 0.4 * i], np.float32)
        out.append(sc)
    return out


def _packed(scs, dev="cuda"):
    ds = [to_dev(sc, dev) for sc in scs]
    cat = lambda k: torch.cat([d[k] for d in ds])
    return dict(means=cat("means"), scales=cat("scales"), rots=cat("rots"), opac=cat("opac"),
                mlp=wr.pack_mlps([d["mlp"] for d in ds]), tx=torch.stack([d["tx"] for d in ds]),
                rx=cat("rx"), views=cat("views")), ds


def _rel(a, b):
    a = a.double().cpu()
    b = b.double().cpu()
    return float((a - b).abs().max() / max(float(b.abs().max()), 1e-30))


def _forward_scenes(scs, p, batch, emission=None, capacity=None):
    H, W = scs[0].H, scs[0].W
    out = wr.forward_scenes(p["means"], p["mlp"], p["opac"], p["scales"], p["rots"], H, W, 1.0, p["views"],
                            p["tx"], p["rx"], batch, emission=emission, capacity=capacity)
    torch.cuda.synchronize()
    return out


def _check_forward(scs, out, alone):
    _, ore, oim, radii, *_ = out
    r0 = rec0 = 0
    for sc, a in zip(scs, alone):
        _, are, aim, arad, *_ = a
        assert torch.equal(ore[r0:r0 + sc.R], are)
        assert torch.equal(oim[r0:r0 + sc.R], aim)
        assert torch.equal(radii[rec0:rec0 + sc.R * sc.P], arad.reshape(-1))
        r0 += sc.R
        rec0 += sc.R * sc.P


@pytest.mark.parametrize("order", [1, 2, 3], ids=["global", "xcd_mixed", "xcd_receiver"])
def test_scene_batch_forward_equals_scenes_alone(order):
    gh.set_option("render_order", order)
    try:
        scs = _scenes()
        print("scenes", SHAPES, "order", order, flush=True)
        alone = [gpu_forward(sc)[1] for sc in scs]
        p, _ = _packed(scs)
        batch = wr.SceneBatch([sc.P for sc in scs], [sc.R for sc in scs], "cuda")
        out = _forward_scenes(scs, p, batch)
        assert out[0] == sum(a[0] for a in alone)
        _check_forward(scs, out, alone)
        # the single-call forward with a tile-list capacity gives the same
        outc = _forward_scenes(scs, p, batch, capacity=out[0] + 100)
        _check_forward(scs, outc, alone)
    finally:
        gh.set_option("render_order", 0)


def test_scene_batch_with_far_gaussians():
    """One scene holds Gaussians 15 km to 1000 km away, beside room-sized
    scenes, in one launch sequence: every scene renders exactly as alone."""
    from test_gpu_rasterizer import far_gaussians
    scs = _scenes()
    far_gaussians(scs[1], 77)
    print("scenes with far Gaussians in scene 1", flush=True)
    alone = [gpu_forward(sc)[1] for sc in scs]
    p, _ = _packed(scs)
    batch = wr.SceneBatch([sc.P for sc in scs], [sc.R for sc in scs], "cuda")
    out = _forward_scenes(scs, p, batch)
    assert out[0] == sum(a[0] for a in alone)
    _check_forward(scs, out, alone)


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_scene_batch_backward_equals_scenes_alone(mode):
    scs = _scenes()
    print("scenes", SHAPES, "mode", mode, flush=True)
    p, ds = _packed(scs)
    batch = wr.SceneBatch([sc.P for sc in scs], [sc.R for sc in scs], "cuda")
    out = _forward_scenes(scs, p, batch)
    Rtot = batch.R
    H, W = scs[0].H, scs[0].W
    g = torch.Generator().manual_seed(5)
    dre = torch.randn((Rtot, H, W), generator=g).cuda()
    dim = torch.randn((Rtot, H, W), generator=g).cuda()
    K, ore, oim, radii, geom, binbuf, img = out
    gs = wr.backward_scenes(p["means"], p["mlp"], radii, p["opac"], p["scales"], p["rots"], 1.0, p["views"], dre,
                            dim, p["tx"], p["rx"], batch, geom, K, binbuf, img, grad_mode=mode)
    torch.cuda.synchronize()
    d2, d3, dmlp, dop, dcov, dsc, drot, _ = gs
    dmlp_s = wr.unpack_mlps(dmlp, batch.scene_P)
    r0 = rec0 = g0 = 0
    for s, (sc, d) in enumerate(zip(scs, ds)):
        _, a = gpu_forward(sc)
        aK, _, _, arad, ageom, abin, aimg = a
        ga = wr.backward_batched(d["means"], d["mlp"], arad, d["opac"], d["scales"], d["rots"], 1.0, d["views"],
                                 dre[r0:r0 + sc.R].contiguous(), dim[r0:r0 + sc.R].contiguous(), d["tx"], d["rx"],
                                 ageom, aK, abin, aimg, grad_mode=mode)
        torch.cuda.synchronize()
        a2, a3, amlp, aop, acov, asc, arot, _ = ga
        P = sc.P
        assert _rel(d2[rec0:rec0 + sc.R * P], a2.reshape(-1, 3)) < GRAD_TOL, s
        assert _rel(d3[g0:g0 + P], a3) < GRAD_TOL, s
        assert _rel(dmlp_s[s], amlp) < GRAD_TOL, s
        assert _rel(dop[g0:g0 + P], aop) < GRAD_TOL, s
        assert _rel(dcov[g0:g0 + P], acov) < GRAD_TOL, s
        assert _rel(dsc[g0:g0 + P], asc) < GRAD_TOL, s
        assert _rel(drot[g0:g0 + P], arot) < GRAD_TOL, s
        r0 += sc.R
        rec0 += sc.R * P
        g0 += P


def test_scene_batch_emission_input_and_autograd():
    """emission= input [Nrec,2] per record, through the autograd function."""
    import cuda_rasterizer as cr
    scs = _scenes()
    p, ds = _packed(scs)
    batch = wr.SceneBatch([sc.P for sc in scs], [sc.R for sc in scs], "cuda")
    g = torch.Generator().manual_seed(9)
    em = (torch.rand((batch.Nrec, 2), generator=g) - 0.3).cuda().requires_grad_(True)
    means = p["means"].clone().requires_grad_(True)
    opac = p["opac"].clone().requires_grad_(True)
    st = cr.SceneSetting(scs[0].H, scs[0].W, 1.0, batch, p["views"], p["tx"], p["rx"])
    re, im, radii = cr.rasterize_scenes(means, None, opac, p["scales"], p["rots"], st, emission=em)
    (re.sum() + 0.25 * im.sum()).backward()
    r0 = rec0 = g0 = 0
    for sc, d in zip(scs, ds):
        n = sc.R * sc.P
        e = em.detach()[rec0:rec0 + n].reshape(sc.R, sc.P, 2).contiguous()
        _, a = gpu_forward(sc, emission=e)
        assert torch.equal(re.detach()[r0:r0 + sc.R], a[1])
        assert torch.equal(im.detach()[r0:r0 + sc.R], a[2])
        aK, _, _, arad, ageom, abin, aimg = a
        ga = wr.backward_batched(d["means"], None, arad, d["opac"], d["scales"], d["rots"], 1.0, d["views"],
                                 torch.ones_like(a[1]), torch.full_like(a[2], 0.25), d["tx"], d["rx"], ageom, aK,
                                 abin, aimg, grad_mode="reference", emission=e)
        torch.cuda.synchronize()
        assert _rel(means.grad[g0:g0 + sc.P], ga[1]) < GRAD_TOL
        assert _rel(opac.grad[g0:g0 + sc.P], ga[3]) < GRAD_TOL
        assert _rel(em.grad[rec0:rec0 + n], ga[7].reshape(-1, 2)) < GRAD_TOL
        r0 += sc.R
        rec0 += n
        g0 += sc.P


def test_scene_batch_with_empty_scenes():
    """A scene without Gaussians renders zero spectra; a scene without
    receivers takes no records and leaves its Gaussians' gradients zero."""
    scs = _scenes()[:2]
    empty_g = S.small_scene(80, 0, R=2)
    no_rx = S.small_scene(81, 300, R=1)
    no_rx.rx_pos = no_rx.rx_pos[:0]
    no_rx.views = no_rx.views[:0]
    allsc = [scs[0], empty_g, no_rx, scs[1]]
    p, ds = _packed(allsc)
    batch = wr.SceneBatch([sc.P for sc in allsc], [sc.R for sc in allsc], "cuda")
    out = _forward_scenes(allsc, p, batch)
    _, ore, oim, radii, geom, binbuf, img = out
    a0 = gpu_forward(scs[0])[1]
    a1 = gpu_forward(scs[1])[1]
    R0 = scs[0].R
    assert torch.equal(ore[:R0], a0[1]) and torch.equal(oim[:R0], a0[2])
    assert not ore[R0:R0 + 2].any() and not oim[R0:R0 + 2].any()
    assert torch.equal(ore[R0 + 2:], a1[1]) and torch.equal(oim[R0 + 2:], a1[2])
    dre = torch.ones_like(ore)
    gs = wr.backward_scenes(p["means"], p["mlp"], radii, p["opac"], p["scales"], p["rots"], 1.0, p["views"], dre,
                            dre, p["tx"], p["rx"], batch, geom, out[0], binbuf, img)
    torch.cuda.synchronize()
    P0 = scs[0].P
    assert not gs[1][P0:P0 + 300].any() and not gs[3][P0:P0 + 300].any()
    assert gs[1][:P0].abs().max() > 0 and gs[1][P0 + 300:].abs().max() > 0


def test_scene_table_rejects_bad_input():
    with pytest.raises(RuntimeError):
        wr.SceneBatch([10, 20], [1], "cuda")
    with pytest.raises(RuntimeError):
        wr.SceneBatch([10, -1], [1, 1], "cuda")
    with pytest.raises(RuntimeError):
        wr.SceneBatch([], [], "cuda")


def test_graphed_scene_step_matches_autograd():
    """The captured scene step (forward, fused L1+SSIM loss, backward) equals
    the torch-autograd scene step on the same parameters, replay after replay,
    with parameters changed in place and eager work between replays."""
    import cuda_rasterizer as cr
    import gsparc_train as T
    from fused_ssim import fused_ssim
    scs = _scenes()
    p, _ = _packed(scs)
    batch = wr.SceneBatch([sc.P for sc in scs], [sc.R for sc in scs], "cuda")
    H, W = scs[0].H, scs[0].W
    g = torch.Generator().manual_seed(17)
    target = (torch.rand((batch.R, H, W), generator=g) * 0.3).cuda()
    K = _forward_scenes(scs, p, batch)[0]
    names = ("means", "scales", "rots", "opac", "mlp")
    live = {k: p[k].clone() for k in names}
    print("capturing graphed scene step, K", K, flush=True)
    step = T.GraphedSceneStep(live["means"], live["scales"], live["rots"], live["opac"], live["mlp"], p["views"],
                              p["tx"], p["rx"], target, H, W, batch, capacity=2 * K)
    st = cr.SceneSetting(H, W, 1.0, batch, p["views"], p["tx"], p["rx"], "reference")

    def reference(vals):
        prm = [v.clone().requires_grad_(True) for v in vals]
        m, s_, q, o, w = prm
        ore, _, _ = cr.rasterize_scenes(m, w, o, s_, q, st)
        loss = 0.8 * (ore - target).abs().mean() + 0.2 * (1.0 - fused_ssim(ore.unsqueeze(1), target.unsqueeze(1)))
        loss.backward()
        torch.cuda.synchronize()
        return float(loss), ore.detach(), [x.grad for x in prm]

    rng = np.random.default_rng(3)
    sets = []
    for i in range(3):
        noise = torch.from_numpy(rng.normal(scale=2e-3, size=(batch.Ptot, 3))).float().cuda()
        sets.append([p["means"] + noise, p["scales"], p["rots"], p["opac"] * (1.0 - 0.05 * i), p["mlp"]])
    refs = [reference(v) for v in sets]
    gname = dict(means="means", scales="scales", rots="rotations", opac="opacities", mlp="emission_mlps")
    for rep in range(2):
        for i, vals in enumerate(sets):
            with torch.no_grad():
                for k, v in zip(names, vals):
                    live[k].copy_(v)
            print("replay", rep, i, flush=True)
            step.replay()
            torch.cuda.synchronize()
            assert not step.overflowed()
            ref_loss, ref_out, ref_grads = refs[i]
            assert torch.equal(step.out_real, ref_out), (rep, i)
            assert abs(float(step.loss) - ref_loss) <= 1e-5 * abs(ref_loss), (rep, i)
            for k, gref in zip(names, ref_grads):
                assert _rel(step.grads[gname[k]].reshape(gref.shape), gref) < GRAD_TOL, (rep, i, k)
            _forward_scenes(scs, p, batch)            # unrelated eager work between replays
    step.run_eager()
    torch.cuda.synchronize()
    assert torch.equal(step.out_real, refs[2][1])
    step.close()
