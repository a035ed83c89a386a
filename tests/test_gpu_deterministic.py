"""Deterministic gradients (GSPARC_GRAD_DETERMINISTIC): the blend adjoint
stores each tile-list instance's sums to its own slot and every record's
instances are summed in a fixed order, so the same inputs give bit-identical
gradients; they agree with the oracle (and with the atomic form) to the usual
1e-4 of the largest magnitude."""
import numpy as np
import pytest
import torch

import gsparc_scenes as S
import gsparc_train as T
import oracle as O
import wireless_rasterizer as wr
from helpers import gpu_forward, to_dev

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-4
NAMES = ("dmeans2D", "dmeans3D", "dmlp", "dopacity", "dcov3D", "dscales", "drotations", "demission")


def _rel(a, b):
    a = a.double().cpu() if isinstance(a, torch.Tensor) else torch.from_numpy(np.asarray(a, np.float64))
    b = b.double().cpu() if isinstance(b, torch.Tensor) else torch.from_numpy(np.asarray(b, np.float64))
    return float((a - b).abs().max() / max(float(b.abs().max()), 1e-30))


def _bwd(d, out, dre, dim, mode, deterministic):
    K, ore, oim, radii, geom, binbuf, img = out
    g = wr.backward_batched(d["means"], d["mlp"], radii, d["opac"], d["scales"], d["rots"], 1.0, d["views"], dre,
                            dim, d["tx"], d["rx"], geom, K, binbuf, img, grad_mode=mode, deterministic=deterministic)
    torch.cuda.synchronize()
    return g


def _same(g1, g2):
    for name, a, b in zip(NAMES, g1, g2):
        if a is not None:
            assert torch.equal(a, b), name


@pytest.mark.parametrize("mode", ["reference", "exact"])
def test_deterministic_backward_matches_oracle_and_repeats(mode):
    R = 3
    sc = S.small_scene(61, 3000, R=R, walls=True)
    print("deterministic backward", mode, flush=True)
    d, out = gpu_forward(sc)
    rng = np.random.default_rng(61)
    dre = rng.normal(size=(R, sc.H, sc.W)).astype(np.float32)
    dim = rng.normal(size=(R, sc.H, sc.W)).astype(np.float32)
    tre, tim = torch.from_numpy(dre).cuda(), torch.from_numpy(dim).cuda()
    g1 = _bwd(d, out, tre, tim, mode, True)
    g2 = _bwd(d, out, tre, tim, mode, True)
    _same(g1, g2)
    m = O.GRAD_EXACT if mode == "exact" else O.GRAD_REFERENCE
    acc = O.Grads(sc.P)
    d2 = []
    for r in range(R):
        pre, keys, vals, ranges, fwd = O.render(sc, r)
        _, g8 = O.backward(sc, pre, vals, ranges, fwd, dre[r], dim[r], m, r=r, grads=acc)
        d2.append(np.stack([g8[:, 0] * (0.5 * sc.W), g8[:, 1] * (0.5 * sc.H), 0 * g8[:, 0]], 1))
    dmeans2D, dmeans3D, dmlp, dop, dcov, dsc, drot, _ = g1
    assert _rel(dmeans2D, np.stack(d2)) < GRAD_TOL
    assert _rel(dmeans3D, acc.means) < GRAD_TOL
    assert _rel(dmlp, acc.mlp) < GRAD_TOL
    assert _rel(dop, acc.opacities) < GRAD_TOL
    assert _rel(dcov, acc.cov3) < GRAD_TOL
    assert _rel(dsc, acc.scales) < GRAD_TOL
    assert _rel(drot, acc.rotations) < GRAD_TOL
    # the atomic form gives the same gradients up to summation order
    ga = _bwd(d, out, tre, tim, mode, False)
    for name, a, b in zip(NAMES, g1, ga):
        if a is not None:
            assert _rel(a, b) < GRAD_TOL, name


def test_deterministic_full_size_and_capacity():
    """50k wall-clustered Gaussians x 64 receivers: two deterministic
    backwards are bit-identical, also through a capacity-sized forward, and
    agree with the atomic backward; an overflowing capacity leaves them zero."""
    sc = S.config2()
    d = to_dev(sc)
    print("deterministic 50k x 64", flush=True)
    out = wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                             d["tx"], d["rx"])
    g = torch.Generator(device="cuda").manual_seed(5)
    dre = torch.randn((sc.R, sc.H, sc.W), generator=g, device="cuda")
    dim = torch.randn((sc.R, sc.H, sc.W), generator=g, device="cuda")
    g1 = _bwd(d, out, dre, dim, "reference", True)
    g2 = _bwd(d, out, dre, dim, "reference", True)
    _same(g1, g2)
    ga = _bwd(d, out, dre, dim, "reference", False)
    for name, a, b in zip(NAMES, g1, ga):
        if a is not None:
            assert _rel(a, b) < GRAD_TOL, name
    K = out[0]
    print("deterministic 50k x 64: capacity", int(1.25 * K), flush=True)
    oc = wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                            d["tx"], d["rx"], capacity=int(1.25 * K))
    assert torch.equal(oc[1], out[1])
    _same(_bwd(d, oc, dre, dim, "reference", True), g1)
    print("deterministic 50k x 64: overflow", flush=True)
    ov = wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                            d["tx"], d["rx"], capacity=K // 2)
    gz = _bwd(d, ov, dre, dim, "reference", True)
    assert float(gz[1].abs().max()) == 0.0 and float(gz[3].abs().max()) == 0.0


def test_deterministic_scene_batch_repeats():
    import test_gpu_scenes as TS
    scs = [S.small_scene(71 + i, P, R=R, walls=True) for i, (P, R) in enumerate([(900, 2), (2500, 5), (0, 3)])]
    p, _ = TS._packed(scs)
    batch = wr.SceneBatch([sc.P for sc in scs], [sc.R for sc in scs], "cuda")
    print("deterministic scenes", [(sc.P, sc.R) for sc in scs], flush=True)
    out = TS._forward_scenes(scs, p, batch)
    K, ore, oim, radii, geom, binbuf, img = out
    g = torch.Generator(device="cuda").manual_seed(7)
    dre = torch.randn(ore.shape, generator=g, device="cuda")
    dim = torch.randn(ore.shape, generator=g, device="cuda")

    def bwd(det):
        r = wr.backward_scenes(p["means"], p["mlp"], radii, p["opac"], p["scales"], p["rots"], 1.0, p["views"], dre,
                               dim, p["tx"], p["rx"], batch, geom, K, binbuf, img, deterministic=det)
        torch.cuda.synchronize()
        return r

    g1, g2, ga = bwd(True), bwd(True), bwd(False)
    _same(g1, g2)
    for name, a, b in zip(NAMES, g1, ga):
        if a is not None:
            assert _rel(a, b) < GRAD_TOL, name


def test_deterministic_graphed_step_equals_eager():
    """The graph-captured step with deterministic gradients: every replay is
    bit-identical to the others and to the eager deterministic backward of
    the same loss gradient."""
    sc = S.small_scene(81, 4000, R=6, walls=True)
    d = to_dev(sc)
    tg = S.small_scene(82, 4000, R=6, walls=True)
    dt = to_dev(tg)
    with torch.no_grad():
        target = wr.forward_batched(dt["means"], dt["mlp"], dt["opac"], dt["scales"], dt["rots"], sc.H, sc.W, 1.0,
                                    d["views"], d["tx"], d["rx"])[1].clone()
    out = wr.forward_batched(d["means"], d["mlp"], d["opac"], d["scales"], d["rots"], sc.H, sc.W, 1.0, d["views"],
                             d["tx"], d["rx"])
    K = out[0]
    print("deterministic graphed step, K", K, flush=True)
    live = {k: d[k].clone() for k in ("means", "scales", "rots", "opac", "mlp")}
    step = T.GraphedTrainingStep(live["means"], live["scales"], live["rots"], live["opac"], live["mlp"], d["views"],
                                 d["tx"], d["rx"], target, sc.H, sc.W, capacity=int(1.25 * K), deterministic=True)
    first = None
    for rep in range(3):
        step.replay()
        torch.cuda.synchronize()
        assert not step.overflowed()
        got = {k: v.clone() for k, v in step.grads.items()}
        if first is None:
            first = got
        else:
            for k in got:
                assert torch.equal(got[k], first[k]), (rep, k)
        torch.zeros(1 << 20, device="cuda").add_(1.0)   # eager work between replays
    ge = _bwd(d, out, step.dimg.clone(), torch.zeros_like(step.dimg), "reference", True)
    assert torch.equal(first["means"], ge[1])
    assert torch.equal(first["emission_mlps"], ge[2])
    assert torch.equal(first["opacities"].reshape(-1), ge[3].reshape(-1))
