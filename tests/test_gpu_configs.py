"""The configurations at their full workload on the GPU.

Batched receivers (50k wall-clustered Gaussians): backward against the oracle
on two receivers in both gradient modes; at 64 receivers the graph-captured
step against the eager autograd step, and the batched forward and backward
against the sum of one-receiver calls.
Multi-frequency maps (50k Gaussians, 52 subcarriers): forward and backward
against the oracle on two receivers; at 32 receivers the zero-frequency
reduction (bit-exact) and the capacity forward against the exact one.
Uneven scenes (16 scenes of 5k-200k Gaussians, 4-64 receivers each): the
batch against every scene rendered alone (spectra and radii bit-exact,
gradients 1e-4), and the smallest scene against the oracle.

Bars as everywhere: spectra 1e-5 of the peak, gradients 1e-4 of the largest
magnitude, discrete outputs bit-exact."""
import numpy as np
import pytest
import torch

import cuda_rasterizer as cr
import gsparc_scenes as S
import gsparc_train as T
import oracle as O
import wireless_rasterizer as wr
from fused_ssim import fused_ssim
from helpers import gpu_forward, read_state, to_dev

pytestmark = pytest.mark.gpu

SPEC_TOL = 1e-5
GRAD_TOL = 1e-4


def _rel(a, b):
    a = np.asarray(torch.as_tensor(a).double().cpu() if isinstance(a, torch.Tensor) else a, np.float64)
    b = np.asarray(torch.as_tensor(b).double().cpu() if isinstance(b, torch.Tensor) else b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _check_receiver(sc, st, r):
    pre, keys, vals, ranges, fwd = O.render(sc, r)
    np.testing.assert_array_equal(st["radii"][r], pre.radii)
    np.testing.assert_array_equal(st["n_contrib"][r], fwd["n_contrib"])
    peak = max(np.abs(fwd["out_re"]).max(), 1e-30)
    assert np.abs(st["out_re"][r] - fwd["out_re"]).max() <= SPEC_TOL * peak
    assert np.abs(st["final_T"][r] - fwd["final_T"]).max() <= 1e-6
    return pre, vals, ranges, fwd


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_config2_backward_matches_oracle(mode):
    sc = S.config2(R=2)
    print("config2 R=2 backward", mode, flush=True)
    d, out = gpu_forward(sc)
    st = read_state(sc, out)
    rng = np.random.default_rng(202)
    dre = rng.normal(size=(sc.R, sc.H, sc.W)).astype(np.float32)
    dim = rng.normal(size=(sc.R, sc.H, sc.W)).astype(np.float32)
    K, ore, oim, radii, geom, binbuf, img = out
    g = wr.backward_batched(d["means"], d["mlp"], radii, d["opac"], d["scales"], d["rots"], 1.0, d["views"],
                            torch.from_numpy(dre).cuda(), torch.from_numpy(dim).cuda(), d["tx"], d["rx"], geom, K,
                            binbuf, img, grad_mode=mode)
    torch.cuda.synchronize()
    m = O.GRAD_EXACT if mode == "exact" else O.GRAD_REFERENCE
    acc = O.Grads(sc.P)
    d2 = []
    for r in range(sc.R):
        pre, vals, ranges, fwd = _check_receiver(sc, st, r)
        _, g8 = O.backward(sc, pre, vals, ranges, fwd, dre[r], dim[r], m, r=r, grads=acc)
        d2.append(np.stack([g8[:, 0] * (0.5 * sc.W), g8[:, 1] * (0.5 * sc.H), 0 * g8[:, 0]], 1))
    dmeans2D, dmeans3D, dmlp, dop, dcov, dsc, drot, _ = [None if x is None else x.cpu().numpy() for x in g]
    assert _rel(dmeans2D, np.stack(d2)) < GRAD_TOL
    assert _rel(dmeans3D, acc.means) < GRAD_TOL
    assert _rel(dmlp, acc.mlp) < GRAD_TOL
    assert _rel(dop, acc.opacities) < GRAD_TOL
    assert _rel(dcov, acc.cov3) < GRAD_TOL
    assert _rel(dsc, acc.scales) < GRAD_TOL
    assert _rel(drot, acc.rotations) < GRAD_TOL


def _torch_loss(img, tgt):
    l1 = (img - tgt).abs().mean()
    return 0.8 * l1 + 0.2 * (1.0 - fused_ssim(img.unsqueeze(1), tgt.unsqueeze(1)))


def test_config2_full_size_step():
    """50k Gaussians x 64 receivers, fwd + loss + bwd: the graph-captured
    step equals the eager autograd step; the batch equals its receivers one
    at a time (spectra bit-exact, gradients summed over receivers)."""
    sc = S.config2()
    d = to_dev(sc)
    tg = S.config2(seed=3)
    dt = to_dev(tg)
    print("config2 50k x 64: target", flush=True)
    with torch.no_grad():
        target = wr.forward_batched(dt["means"], dt["mlp"], dt["opac"], dt["scales"], dt["rots"], sc.H, sc.W, 1.0,
                                    d["views"], d["tx"], d["rx"])[1].clone()
    K = wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                           d["tx"], d["rx"])[0]
    # eager autograd step
    params = [d[k].clone().requires_grad_(True) for k in ("means", "scales", "rots", "opac", "mlp")]
    m, s_, q, o, w = params
    st = cr.BatchedSetting(sc.H, sc.W, 1.0, d["views"], d["tx"], d["rx"], "reference")
    ore, oim, radii = cr.rasterize_gaussians_batched(m, w, o, s_, q, st)
    loss = _torch_loss(ore, target)
    loss.backward()
    torch.cuda.synchronize()
    print("config2 50k x 64: graphed step, K", K, flush=True)
    live = {k: d[k].clone() for k in ("means", "scales", "rots", "opac", "mlp")}
    step = T.GraphedTrainingStep(live["means"], live["scales"], live["rots"], live["opac"], live["mlp"], d["views"],
                                 d["tx"], d["rx"], target, sc.H, sc.W, capacity=int(1.25 * K))
    for rep in range(2):
        step.replay()
        torch.cuda.synchronize()
        assert not step.overflowed()
        assert torch.equal(step.out_real, ore.detach()), rep
        assert abs(float(step.loss) - float(loss)) <= 1e-5 * abs(float(loss)), rep
        for k, p_ in zip(("means", "scales", "rotations", "opacities", "emission_mlps"), params):
            assert _rel(step.grads[k].reshape(p_.shape), p_.grad) < GRAD_TOL, (rep, k)
    # the same loss gradient through the batched backward and through 64 one-receiver calls
    dimg = step.dimg.clone()
    zero = torch.zeros_like(dimg)
    Kb, bre, bim, brad, bgeom, bbin, bimg = wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"],
                                                               d["rots"], sc.H, sc.W, 1.0, d["views"], d["tx"],
                                                               d["rx"])
    gb = wr.backward_batched(d["means"], d["mlp"], brad, d["opac"], d["scales"], d["rots"], 1.0, d["views"], dimg,
                             zero, d["tx"], d["rx"], bgeom, Kb, bbin, bimg)
    print("config2 50k x 64: one receiver at a time", flush=True)
    acc = None
    for r in range(sc.R):
        K1, re1, im1, rad1, geom1, bin1, img1 = wr.forward_batched(
            d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"][r:r + 1],
            d["tx"], d["rx"][r:r + 1])
        assert torch.equal(re1[0], bre[r]) and torch.equal(im1[0], bim[r]) and torch.equal(rad1[0], brad[r]), r
        g1 = wr.backward_batched(d["means"], d["mlp"], rad1, d["opac"], d["scales"], d["rots"], 1.0,
                                 d["views"][r:r + 1], dimg[r:r + 1].contiguous(), zero[r:r + 1].contiguous(), d["tx"],
                                 d["rx"][r:r + 1], geom1, K1, bin1, img1)
        assert _rel(g1[0][0], gb[0][r]) < GRAD_TOL, r          # screen-space means of receiver r
        part = [g1[i].double() for i in (1, 2, 3, 5, 6)]
        acc = part if acc is None else [a + b for a, b in zip(acc, part)]
    torch.cuda.synchronize()
    for a, i in zip(acc, (1, 2, 3, 5, 6)):
        assert _rel(gb[i], a) < GRAD_TOL, i
    step.close()


def _csi_fwd(sc, d, **kw):
    out = wr.forward_csi(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                         d["tx"], d["rx"], **kw)
    torch.cuda.synchronize()
    return out


def test_csi_50k_matches_oracle():
    sc = S.config2(R=2)
    sc.mlp *= 3.0
    d = to_dev(sc)
    print("csi 50k x 2 forward/backward", flush=True)
    K, out, radii, geom, binbuf, img = _csi_fwd(sc, d)
    rng = np.random.default_rng(303)
    dout = rng.normal(size=(sc.R, sc.H, sc.W, wr.CSI_NC, 2)).astype(np.float32)
    g = wr.backward_csi(d["means"], d["mlp"], radii, d["opac"], d["scales"], d["rots"], 1.0, d["views"],
                        torch.from_numpy(dout).cuda(), d["tx"], d["rx"], geom, K, binbuf, img)
    torch.cuda.synchronize()
    outn = out.cpu().numpy()
    acc = O.Grads(sc.P)
    for r in range(sc.R):
        pre, plen, vals, ranges, fwd = O.render_csi(sc, r)
        np.testing.assert_array_equal(radii[r].cpu().numpy(), pre.radii)
        peak = max(np.abs(fwd["out"]).max(), 1e-30)
        assert np.abs(outn[r] - fwd["out"]).max() <= SPEC_TOL * peak, r
        O.backward_csi(sc, pre, plen, vals, ranges, fwd, dout[r], r, grads=acc)
    d2, d3, dmlp, dop, dcov, dsc, drot, _ = [None if x is None else x.cpu().numpy() for x in g]
    assert _rel(d3, acc.means) < GRAD_TOL
    assert _rel(dmlp, acc.mlp) < GRAD_TOL
    assert _rel(dop, acc.opacities) < GRAD_TOL
    assert _rel(dcov, acc.cov3) < GRAD_TOL
    assert _rel(dsc, acc.scales) < GRAD_TOL
    assert _rel(drot, acc.rotations) < GRAD_TOL


def test_csi_50k_32rx_properties():
    """At the multi-frequency workload's size: with all frequencies zero each
    subcarrier is the single-channel render bit for bit, and the capacity
    forward equals the exact one."""
    sc = S.config2(R=32)
    d = to_dev(sc)
    print("csi 50k x 32 properties", flush=True)
    _, z, *_ = _csi_fwd(sc, d, f0=0.0, df=0.0)
    one = wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                             d["tx"], d["rx"])
    torch.cuda.synchronize()
    for k in (0, 25, 51):
        assert torch.equal(z[..., k, 0], one[1]) and torch.equal(z[..., k, 1], one[2]), k
    del z
    K, out, *_ = _csi_fwd(sc, d)
    Kc, out2, *_ = _csi_fwd(sc, d, capacity=K + 1000)
    assert torch.equal(out, out2)
    assert bool(torch.isfinite(out).all())


def _pack(scs):
    ds = [to_dev(sc) for sc in scs]
    cat = lambda k: torch.cat([x[k] for x in ds])
    return dict(means=cat("means"), scales=cat("scales"), rots=cat("rots"), opac=cat("opac"),
                mlp=wr.pack_mlps([x["mlp"] for x in ds]), tx=torch.stack([x["tx"] for x in ds]), rx=cat("rx"),
                views=cat("views")), ds


def test_config4_batch_equals_scenes_alone_and_oracle():
    scs = S.config4()
    print("config4", [(sc.P, sc.R) for sc in scs], flush=True)
    p, ds = _pack(scs)
    batch = wr.SceneBatch([sc.P for sc in scs], [sc.R for sc in scs], "cuda")
    H, W = scs[0].H, scs[0].W
    out = wr.forward_scenes(p["means"], p["mlp"], p["opac"], p["scales"], p["rots"], H, W, 1.0, p["views"], p["tx"],
                            p["rx"], batch)
    torch.cuda.synchronize()
    K, ore, oim, radii, geom, binbuf, img = out
    g = torch.Generator().manual_seed(404)
    dre = torch.randn((batch.R, H, W), generator=g).cuda()
    dim = torch.randn((batch.R, H, W), generator=g).cuda()
    gs = wr.backward_scenes(p["means"], p["mlp"], radii, p["opac"], p["scales"], p["rots"], 1.0, p["views"], dre,
                            dim, p["tx"], p["rx"], batch, geom, K, binbuf, img)
    torch.cuda.synchronize()
    d2, d3, dmlp, dop, dcov, dsc, drot, _ = gs
    dmlp_s = wr.unpack_mlps(dmlp, batch.scene_P)
    r0 = rec0 = g0 = 0
    Ksum = 0
    for s, (sc, dd) in enumerate(zip(scs, ds)):
        print("config4 scene", s, sc.P, sc.R, flush=True)
        aK, are, aim, arad, ageom, abin, aimg = wr.forward_batched(
            dd["means"], dd["mlp"], dd["opac"], dd["scales"], dd["rots"], H, W, 1.0, dd["views"], dd["tx"], dd["rx"])
        Ksum += aK
        assert torch.equal(ore[r0:r0 + sc.R], are) and torch.equal(oim[r0:r0 + sc.R], aim), s
        assert torch.equal(radii[rec0:rec0 + sc.R * sc.P], arad.reshape(-1)), s
        ga = wr.backward_batched(dd["means"], dd["mlp"], arad, dd["opac"], dd["scales"], dd["rots"], 1.0,
                                 dd["views"], dre[r0:r0 + sc.R].contiguous(), dim[r0:r0 + sc.R].contiguous(),
                                 dd["tx"], dd["rx"], ageom, aK, abin, aimg)
        torch.cuda.synchronize()
        P = sc.P
        assert _rel(d3[g0:g0 + P], ga[1]) < GRAD_TOL, s
        assert _rel(dmlp_s[s], ga[2]) < GRAD_TOL, s
        assert _rel(dop[g0:g0 + P], ga[3]) < GRAD_TOL, s
        assert _rel(dsc[g0:g0 + P], ga[5]) < GRAD_TOL, s
        assert _rel(drot[g0:g0 + P], ga[6]) < GRAD_TOL, s
        r0 += sc.R
        rec0 += sc.R * P
        g0 += P
    assert K == Ksum
    # the smallest scene against the oracle, forward and backward
    s = int(np.argmin([sc.P for sc in scs]))
    sc = scs[s]
    r0 = sum(x.R for x in scs[:s])
    g0 = sum(x.P for x in scs[:s])
    print("config4 oracle: scene", s, sc.P, sc.R, flush=True)
    ore_s = ore[r0:r0 + sc.R].cpu().numpy()
    dre_s = dre[r0:r0 + sc.R].cpu().numpy()
    dim_s = dim[r0:r0 + sc.R].cpu().numpy()
    acc = O.Grads(sc.P)
    for r in range(sc.R):
        pre, keys, vals, ranges, fwd = O.render(sc, r)
        peak = max(np.abs(fwd["out_re"]).max(), 1e-30)
        assert np.abs(ore_s[r] - fwd["out_re"]).max() <= SPEC_TOL * peak, r
        O.backward(sc, pre, vals, ranges, fwd, dre_s[r], dim_s[r], O.GRAD_REFERENCE, r=r, grads=acc)
    assert _rel(d3[g0:g0 + sc.P].cpu().numpy(), acc.means) < GRAD_TOL
    assert _rel(dop[g0:g0 + sc.P].cpu().numpy(), acc.opacities) < GRAD_TOL
    assert _rel(dmlp_s[s].cpu().numpy(), acc.mlp) < GRAD_TOL
    assert _rel(dsc[g0:g0 + sc.P].cpu().numpy(), acc.scales) < GRAD_TOL
    assert _rel(drot[g0:g0 + sc.P].cpu().numpy(), acc.rotations) < GRAD_TOL
