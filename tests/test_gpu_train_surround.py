"""GPU parity of the training surround (k_train.hip through the C ABI).

Adam: the library's step against torch.optim.Adam(eps=1e-15) itself (the
reference's optimizer, gaussian.py:197) on the same gradients, through the
activations (exp / normalize / sigmoid by torch autograd), with the
reference's per-iteration rules: NaN centre gradients zeroed, the centre
rate from the schedule, an opacity reset whose parameter does not step, an
iteration where only the confidence network steps, the clamp.  Within 4
float ulps of torch, or 1e-5 of one step (the group's rate) for values an
update carried through zero.

Densify / prune: the library against the CPU restatement
(oracle/train_oracle.py, pinned in tests/test_train_oracle.py) on the same
injected noise: counts equal, copied arrays bit-exact, split children within
2e-6; twice in a row, through a capacity refusal and regrowth; the pruning
of the post-densification phase; the statistics.

Trainer: the schedule's events and the per-group step counts; the graphed
iterations against eager ones; the first iteration's gradients against the
autograd path and its update against torch.optim.Adam on those gradients."""
import math

import numpy as np
import pytest
import torch

import gsparc_gaussian as G
import gsparc_hip as gh
import train_oracle as TO

pytestmark = pytest.mark.gpu

LO = np.array([-11.0, -8.0, -4.0], np.float32)
HI = np.array([11.0, 10.0, 8.0], np.float32)
EXTENT = float(np.linalg.norm(np.array([7.0, 6.0, 4.0]) - np.array([-7.0, -4.0, 0.0])))


def _close(a, b, ulps=4, atol=1e-9, what=""):
    a = np.asarray(a, np.float32).reshape(-1)
    b = np.asarray(b, np.float32).reshape(-1)
    tol = ulps * np.spacing(np.abs(b)) + atol
    bad = ~(np.abs(a.astype(np.float64) - b) <= tol)
    assert not bad.any(), (what, int(bad.sum()), a[bad][:5], b[bad][:5])


def _raw(P, seed):
    rng = np.random.default_rng(seed)
    centre = rng.uniform(-6, 6, size=(P, 3)).astype(np.float32)
    ls = np.log(rng.uniform(0.02, 0.6, size=(P, 3))).astype(np.float32)
    rot = rng.normal(size=(P, 4)).astype(np.float32)
    logit = rng.normal(0, 3, size=P).astype(np.float32)
    mlp = rng.normal(size=P * 194).astype(np.float32)
    return centre, ls, rot, logit, mlp


def _gpu_model(raw, headroom=1.5):
    m = G.GaussianModel(headroom=headroom)
    m.load_raw(*[torch.from_numpy(x).cuda() for x in raw])
    return m


@pytest.mark.parametrize("P", [3000, 1001])
def test_adam_steps_match_torch_adam(P):
    """P = 3000: the MLP moments go through the float4 kernel; P = 1001
    (194 P not a multiple of 4): through the scalar one."""
    rng = np.random.default_rng(11)
    raw = _raw(P, 3)
    conf0 = (rng.normal(size=161) * 0.3).astype(np.float32)
    conf_gpu = torch.from_numpy(conf0).cuda()
    m = _gpu_model(raw)
    opt_p = G.OptimizationParams()
    m.training_setup(opt_p, confidence=conf_gpu, scene_min=LO, scene_max=HI, first_iteration=100)
    # the reference's parameters and optimizer, on the CPU
    centre, ls, rot, logit, mlp = raw
    P_ = torch.nn.Parameter
    tp = {"centre": P_(torch.from_numpy(centre.copy())), "scaling": P_(torch.from_numpy(ls.copy())),
          "rotation": P_(torch.from_numpy(rot.copy())), "opacity": P_(torch.from_numpy(logit.copy()).reshape(P, 1))}
    o = 0
    for name, w, shape in (("fc1_weights", 96, (P, 32, 3)), ("fc1_bias", 32, (P, 32)), ("fc2_weights", 64, (P, 2, 32)),
                           ("fc2_bias", 2, (P, 2))):
        tp[name] = P_(torch.from_numpy(mlp[o: o + w * P].copy()).reshape(shape))
        o += w * P
    tp["conf"] = P_(torch.from_numpy(conf0.copy()))
    lrs = {"centre": 0.0, "scaling": opt_p.scaling_lr, "rotation": opt_p.rotation_lr, "opacity": opt_p.opacity_lr,
           "conf": opt_p.confidence_lr}
    groups = [{"params": [p], "lr": lrs.get(n, opt_p.emission_mlp_lr), "name": n} for n, p in tp.items()]
    opt = torch.optim.Adam(groups, lr=0.0, eps=1e-15)
    per_gauss = [n for n in tp if n != "conf"]
    ALL = gh.GROUPS_ALL
    schedule = [ALL, ALL, ALL & ~gh.GROUP_OPACITY, ALL, gh.GROUP_CONFIDENCE, ALL, ALL]
    for k, mask in enumerate(schedule):
        it = 100 + k
        print("adam step", k, "groups", mask, flush=True)
        gc = (rng.normal(size=(P, 3)) * 1e-3).astype(np.float32)
        gc[rng.integers(0, P, 20), rng.integers(0, 3, 20)] = np.nan
        gs = rng.normal(size=(P, 3)).astype(np.float32)
        gr = rng.normal(size=(P, 4)).astype(np.float32)
        go = rng.normal(size=P).astype(np.float32)
        gm = rng.normal(size=P * 194).astype(np.float32)
        gcf = rng.normal(size=161).astype(np.float32)
        reset = mask == ALL & ~gh.GROUP_OPACITY
        # --- torch
        if reset:
            with torch.no_grad():
                z = torch.log(torch.min(torch.sigmoid(tp["opacity"]), torch.ones_like(tp["opacity"]) * 0.01) /
                              (1 - torch.min(torch.sigmoid(tp["opacity"]), torch.ones_like(tp["opacity"]) * 0.01)))
                tp["opacity"].copy_(torch.nan_to_num(z, nan=0.0, posinf=1.0, neginf=0.0))
                st = opt.state[tp["opacity"]]
                st["exp_avg"].zero_()
                st["exp_avg_sq"].zero_()
        for p in tp.values():
            p.grad = None
        act = (torch.from_numpy(gs) * torch.exp(tp["scaling"])).sum() + \
              (torch.from_numpy(gr) * torch.nn.functional.normalize(tp["rotation"], dim=-1)).sum() + \
              (torch.from_numpy(go).reshape(P, 1) * torch.sigmoid(tp["opacity"])).sum()
        act.backward()
        g = torch.from_numpy(gc.copy())
        g[torch.isnan(g)] = 0.0
        tp["centre"].grad = g
        o = 0
        for name, w in (("fc1_weights", 96), ("fc1_bias", 32), ("fc2_weights", 64), ("fc2_bias", 2)):
            tp[name].grad = torch.from_numpy(gm[o: o + w * P].copy()).reshape(tp[name].shape)
            o += w * P
        tp["conf"].grad = torch.from_numpy(gcf.copy())
        if mask == gh.GROUP_CONFIDENCE:
            for n in per_gauss:
                tp[n].grad = None
        if reset:
            tp["opacity"].grad = None
        groups[0]["lr"] = TO.lr_schedule(it, opt_p.position_lr_init, opt_p.position_lr_final,
                                         opt_p.position_lr_max_steps)
        opt.step()
        with torch.no_grad():
            tp["centre"].clamp_(min=torch.from_numpy(LO), max=torch.from_numpy(HI))
        # --- library
        if reset:
            m.reset_opacity()
        m.update_learning_rate(it)
        m.conf_grad.copy_(torch.from_numpy(gcf))
        t = lambda x: torch.from_numpy(x).cuda()
        m.optimizer_step(t(gc), t(gs), t(gr), t(go), t(gm), groups=mask)
        torch.cuda.synchronize()
        # --- compare: within 4 ulps, or 1e-5 of one step of the group (a
        # value that an update carried through zero)
        step_tol = lambda lr: 1e-5 * lr
        _close(m.get_centre.cpu(), tp["centre"].detach(), atol=step_tol(groups[0]["lr"]), what=("centre", k))
        _close(m._scaling.cpu(), tp["scaling"].detach(), atol=step_tol(opt_p.scaling_lr), what=("scaling", k))
        _close(m._rotation.cpu(), tp["rotation"].detach(), atol=step_tol(opt_p.rotation_lr), what=("rotation", k))
        _close(m._opacity.cpu(), tp["opacity"].detach(), atol=step_tol(opt_p.opacity_lr), what=("opacity", k))
        flat = torch.cat([tp[n].detach().reshape(-1) for n in ("fc1_weights", "fc1_bias", "fc2_weights", "fc2_bias")])
        _close(m.get_emission_mlps.cpu(), flat, atol=step_tol(opt_p.emission_mlp_lr), what=("mlp", k))
        _close(m.conf.cpu(), tp["conf"].detach(), atol=step_tol(opt_p.confidence_lr), what=("conf", k))
        # activations, of the library's own raw values (the raw values are
        # within 4 ulps of torch's; exp and sigmoid would scale that up)
        _close(m.get_scaling.cpu(), torch.exp(m._scaling.cpu()), what=("scale act", k))
        _close(m.get_rotation.cpu(), torch.nn.functional.normalize(m._rotation.cpu(), dim=-1), ulps=8,
               what=("rotation act", k))
        _close(m.get_opacity.cpu(), torch.sigmoid(m._opacity.cpu()), what=("opacity act", k))
    # the step counts are torch's state["step"]
    c = m.counters()
    assert c["iteration"] == 100 + len(schedule)
    assert c["steps"]["centre"] == int(opt.state[tp["centre"]]["step"]) == 6
    assert c["steps"]["opacity"] == int(opt.state[tp["opacity"]]["step"]) == 5
    assert c["steps"]["confidence"] == int(opt.state[tp["conf"]]["step"]) == 7
    assert not c["nan_centre"]


def test_infinite_centre_gradient_raises_the_nan_flag():
    """An infinite centre gradient is not zeroed (train.py:441-444 masks NaN
    only): Adam turns it into a NaN centre, which the step flags."""
    raw = _raw(500, 4)
    m = _gpu_model(raw)
    m.training_setup(G.OptimizationParams(), scene_min=LO, scene_max=HI)
    z = lambda n: torch.zeros(n, device="cuda")
    gc = torch.zeros(1500, device="cuda")
    gc[7] = float("inf")
    m.optimizer_step(gc, z(1500), z(2000), z(500), z(500 * 194))
    torch.cuda.synchronize()
    assert m.counters()["nan_centre"]
    assert torch.isnan(m.get_centre[2, 1]).item()
    assert int(torch.isnan(m.get_centre).sum()) == 1


def _densify_inputs(P, seed):
    rng = np.random.default_rng(seed)
    raw = list(_raw(P, seed))
    out = rng.uniform(size=P) < 0.03
    raw[0][out] += np.float32(30.0)                  # outside the bounds
    accum = rng.uniform(0, 0.004, size=P).astype(np.float32)
    denom = rng.integers(0, 20, size=P).astype(np.float32)
    accum[denom == 0] = 0.0
    maxr = rng.integers(0, 60, size=P).astype(np.float32)
    geo = rng.normal(size=22 * P).astype(np.float32)
    mst = rng.normal(size=388 * P).astype(np.float32)
    return raw, accum, denom, maxr, geo, mst


def _load(m, accum, denom, maxr, geo, mst):
    t = lambda x: torch.from_numpy(x).cuda()
    m.array("accum").copy_(t(accum))
    m.array("denom").copy_(t(denom))
    m.array("max_radius").copy_(t(maxr))
    m.array("state").copy_(t(geo))
    m.array("state_mlp").copy_(t(mst))


def _compare_to_oracle(m, o, what):
    assert m.P == o.P, (what, m.P, o.P)
    _close(m.get_centre.cpu(), o.p["centre"], ulps=16, atol=2e-6, what=(what, "centre"))
    _close(m._scaling.cpu(), o.p["scaling"], ulps=4, atol=1e-6, what=(what, "log scale"))
    assert torch.equal(m._rotation.cpu(), o.p["rotation"]), what
    assert torch.equal(m._opacity.cpu(), o.p["opacity"]), what
    assert torch.equal(m.get_emission_mlps.cpu(), o.mlp_flat()), what
    geo, mst = TO.state_flat(o)
    assert np.array_equal(m.array("state").cpu().numpy(), geo), what
    assert np.array_equal(m.array("state_mlp").cpu().numpy(), mst), what
    assert np.array_equal(m.centre_gradient_accum.cpu().numpy(), o.accum.numpy()), what
    assert np.array_equal(m.denom.cpu().numpy(), o.denom.numpy()), what
    assert np.array_equal(m.max_radii2D.cpu().numpy(), o.max_radii2D.numpy()), what
    _close(m.get_scaling.cpu(), torch.exp(o.p["scaling"]), ulps=4, atol=1e-7, what=(what, "scale act"))
    _close(m.get_opacity.cpu(), torch.sigmoid(o.p["opacity"]), ulps=4, atol=1e-9, what=(what, "opacity act"))


def test_densify_and_prune_matches_the_restatement():
    P = 6000
    raw, accum, denom, maxr, geo, mst = _densify_inputs(P, 21)
    m = _gpu_model(raw, headroom=1.0)          # the second bank starts too small: refusal, then regrowth
    _load(m, accum, denom, maxr, geo, mst)
    o = TO.from_arrays(*raw, accum, denom, maxr)
    TO.set_state_flat(o, geo, mst)
    rng = np.random.default_rng(5)
    for rnd, screen in enumerate((None, 20)):
        noise = rng.normal(size=(2 * m.P, 3)).astype(np.float32)
        print("densify round", rnd, "P", m.P, flush=True)
        P_new, nc, ns = m.densify_and_prune(0.0002, 0.01, EXTENT, screen, scene_min=LO, scene_max=HI,
                                            noise=torch.from_numpy(noise).cuda())
        torch.cuda.synchronize()
        onc, ons = TO.densify_and_prune(o, 0.0002, 0.01, EXTENT, screen, LO, HI, noise)
        assert (nc, ns) == (onc, ons) and nc > 50 and ns > 50, (nc, ns, onc, ons)
        _compare_to_oracle(m, o, ("round", rnd))
        # give the next round statistics and moments again
        P2 = m.P
        accum2 = np.random.default_rng(rnd).uniform(0, 0.004, size=P2).astype(np.float32)
        denom2 = np.random.default_rng(rnd + 9).integers(1, 20, size=P2).astype(np.float32)
        maxr2 = np.zeros(P2, np.float32)
        geo2 = np.random.default_rng(rnd + 3).normal(size=22 * P2).astype(np.float32)
        mst2 = np.random.default_rng(rnd + 4).normal(size=388 * P2).astype(np.float32)
        _load(m, accum2, denom2, maxr2, geo2, mst2)
        o.accum = torch.from_numpy(accum2).reshape(P2, 1)
        o.denom = torch.from_numpy(denom2).reshape(P2, 1)
        o.max_radii2D = torch.from_numpy(maxr2)
        TO.set_state_flat(o, geo2, mst2)


def test_prune_only_carries_statistics():
    P = 4000
    raw, accum, denom, maxr, geo, mst = _densify_inputs(P, 33)
    m = _gpu_model(raw)
    _load(m, accum, denom, maxr, geo, mst)
    o = TO.from_arrays(*raw, accum, denom, maxr)
    TO.set_state_flat(o, geo, mst)
    n = m.prune(0.001, EXTENT, LO, HI)
    torch.cuda.synchronize()
    assert n == TO.prune_post(o, 0.001, EXTENT, LO, HI) and n > 0
    _compare_to_oracle(m, o, "prune")
    # nothing to prune: the set stays as it is
    P1 = m.P
    assert m.prune(0.001, EXTENT, LO, HI) == 0 and m.P == P1
    _compare_to_oracle(m, o, "prune again")


def test_densify_stats_match_the_restatement():
    P, R = 5000, 6
    rng = np.random.default_rng(8)
    m = _gpu_model(_raw(P, 8))
    g2d = rng.normal(size=(R, P, 3)).astype(np.float32)
    radii = rng.integers(-2, 30, size=(R, P)).clip(0).astype(np.int32)
    o = TO.Model(P)
    for rep in range(2):
        m.add_densification_stats(torch.from_numpy(g2d).cuda(), torch.from_numpy(radii).cuda(), grad_scale=R)
        TO.add_stats(o, g2d, radii, grad_scale=R)
    torch.cuda.synchronize()
    _close(m.centre_gradient_accum.cpu(), o.accum, ulps=8, atol=1e-6, what="accum")
    assert np.array_equal(m.denom.cpu().numpy(), o.denom.numpy())
    assert np.array_equal(m.max_radii2D.cpu().numpy(), o.max_radii2D.numpy())


# ---------------------------------------------------------------- trainer
def _trainer(P=1500, R=4, seed=71, graphs=True, k_capacity=1000, **opt):
    import cuda_rasterizer as cr  # noqa: F401
    import gsparc_scenes as S
    import gsparc_train as T
    import gsparc_trainer as TR
    import wireless_rasterizer as wr
    from helpers import to_dev
    sc = S.small_scene(seed, P, R=R, walls=True)
    d = to_dev(sc)
    tg = to_dev(S.small_scene(seed + 1, P, R=R, walls=True))
    with torch.no_grad():
        target = wr.forward_batched(tg["means"], tg["mlp"], tg["opac"], tg["scales"], tg["rots"], sc.H, sc.W, 1.0,
                                    d["views"], d["tx"], d["rx"])[1].clone()
    m = G.GaussianModel()
    m.load_raw(d["means"], torch.log(d["scales"]), d["rots"], torch.logit(d["opac"].double()).float(), d["mlp"])
    torch.manual_seed(seed)
    net = T.ConfidenceMLP(32, device="cuda")
    conf = net.packed().detach().clone().contiguous()
    o = G.OptimizationParams(**opt)
    tr = TR.Trainer(m, o, d["views"], d["tx"], d["rx"], target, sc.H, sc.W, xyz_min=[0.0, 0.0, 0.0],
                    xyz_max=S.ROOM, confidence=conf, use_graphs=graphs, k_capacity=k_capacity)
    return tr, sc, d, target, net


def test_trainer_follows_the_schedule():
    tr, sc, d, target, net = _trainer(densify_from_iter=2, densification_interval=5, opacity_reset_interval=10,
                                      densify_until_iter=18, densify_grad_threshold=2e-5)
    losses = []
    for it in range(1, 25):
        print("iteration", it, "P", tr.m.P, flush=True)
        tr.step()
        losses.append(float(tr.loss))
    torch.cuda.synchronize()
    ev = [(i, e) for i, e, _, _ in tr.events]
    assert ev == [(5, "densify"), (10, "densify"), (10, "reset_opacity"), (15, "densify")], ev
    grew = [(a, b) for i, e, a, b in tr.events if e == "densify"]
    assert grew[0][1] > grew[0][0], grew
    c = tr.m.counters()
    assert c["iteration"] == 25
    assert c["steps"] == {"centre": 21, "scaling": 21, "rotation": 21, "opacity": 21, "emission": 21,
                          "confidence": 24}, c
    assert not c["nan_centre"] and not tr.overflowed()
    assert all(math.isfinite(x) for x in losses)
    # the statistics stop at densify_until_iter; the arrays stay finite
    for name in ("centre", "log_scale", "rotation", "logit", "emission_mlps"):
        assert bool(torch.isfinite(tr.m.array(name)).all()), name
    tr.close()


def test_graphed_iterations_match_eager_ones():
    runs = []
    for graphs in (True, False):
        tr, sc, d, target, net = _trainer(graphs=graphs, densify_from_iter=10 ** 6, k_capacity=1 << 22)
        ls = []
        for it in range(6):
            tr.step()
            ls.append(float(tr.loss))
        torch.cuda.synchronize()
        runs.append((ls, {n: tr.m.array(n).detach().clone() for n in ("centre", "log_scale", "rotation", "logit",
                                                                    "emission_mlps")}, tr.m.conf.clone()))
        tr.close()
    (la, pa, ca), (lb, pb, cb) = runs
    for x, y in zip(la, lb):
        assert abs(x - y) <= 1e-4 * abs(y), (la, lb)
    lr_max = 0.0055 * 6 * 2.1
    for n in pa:
        a, b = pa[n], pb[n]
        diff = (a - b).abs()
        tight = diff <= 1e-5 * b.abs() + 1e-6
        assert float(tight.float().mean()) > 0.9995, (n, float(tight.float().mean()))
        assert float(diff.max()) <= lr_max, (n, float(diff.max()))
    assert float((ca - cb).abs().max()) <= 1e-4


def test_first_iteration_matches_autograd_and_torch_adam():
    import cuda_rasterizer as cr
    import gsparc_train as T
    tr, sc, d, target, net = _trainer(densify_from_iter=10 ** 6)
    m = tr.m
    raw0 = {n: m.array(n).detach().clone() for n in ("centre", "log_scale", "rotation", "logit", "emission_mlps")}
    act = [m.get_centre.detach().clone().requires_grad_(True), m.get_scaling.detach().clone().requires_grad_(True),
           m.get_rotation.detach().clone().requires_grad_(True), m.get_opacity.detach().clone().requires_grad_(True),
           m.get_emission_mlps.detach().clone().requires_grad_(True)]
    conf0 = m.conf.detach().clone()
    mu, s_, q, o, w = act
    st = cr.BatchedSetting(sc.H, sc.W, 1.0, d["views"], d["tx"], d["rx"], "reference")
    ore, oim, radii = cr.rasterize_gaussians_batched(mu, w, o, s_, q, st)
    loss = T.confidence_loss(ore, target, d["rx"], net)
    loss.backward()
    tr.step()
    torch.cuda.synchronize()
    b = tr.buf
    P = m.P
    rel = lambda a, r: float((a.double() - r.double()).abs().max() / max(float(r.abs().max()), 1e-30))
    assert abs(float(tr.loss) - float(loss)) <= 1e-5 * abs(float(loss))
    assert rel(b.dcentre[:3 * P].view(P, 3), mu.grad) < 1e-4
    assert rel(b.dscale[:3 * P].view(P, 3), s_.grad) < 1e-4
    assert rel(b.drot[:4 * P].view(P, 4), q.grad) < 1e-4
    assert rel(b.dopac[:P].view(P, 1), o.grad) < 1e-4
    assert rel(b.dmlp[:194 * P], w.grad) < 1e-4
    packed_grad = torch.cat([net.fc1.weight.grad.reshape(-1), net.fc1.bias.grad, net.fc2.weight.grad.reshape(-1),
                             net.fc2.bias.grad])
    assert rel(m.conf_grad, packed_grad) < 1e-5
    # the update: torch.optim.Adam (GPU) on the trainer's own gradients
    opt_p = tr.o
    tp = {n: torch.nn.Parameter(raw0[n].clone()) for n in raw0}
    tc = torch.nn.Parameter(conf0.clone())
    groups = [{"params": [tp["centre"]], "lr": TO.lr_schedule(1, opt_p.position_lr_init, opt_p.position_lr_final,
                                                               opt_p.position_lr_max_steps)},
              {"params": [tp["log_scale"]], "lr": opt_p.scaling_lr},
              {"params": [tp["rotation"]], "lr": opt_p.rotation_lr},
              {"params": [tp["logit"]], "lr": opt_p.opacity_lr},
              {"params": [tp["emission_mlps"]], "lr": opt_p.emission_mlp_lr},
              {"params": [tc], "lr": opt_p.confidence_lr}]
    opt = torch.optim.Adam(groups, lr=0.0, eps=1e-15)
    sc_ = torch.exp(raw0["log_scale"])
    qq = raw0["rotation"].view(P, 4)
    nq = qq.norm(dim=-1, keepdim=True).clamp_min(1e-12)
    y = qq / nq
    dy = b.drot[:4 * P].view(P, 4)
    op = torch.sigmoid(raw0["logit"])
    gc = b.dcentre[:3 * P].clone()
    gc[torch.isnan(gc)] = 0.0
    tp["centre"].grad = gc
    tp["log_scale"].grad = b.dscale[:3 * P] * sc_
    tp["rotation"].grad = ((dy - y * (y * dy).sum(-1, keepdim=True)) / nq).reshape(-1)
    tp["logit"].grad = b.dopac[:P] * (1 - op) * op
    tp["emission_mlps"].grad = b.dmlp[:194 * P].clone()
    tc.grad = m.conf_grad.clone()
    opt.step()
    with torch.no_grad():
        lo = torch.tensor([-4.0, -4.0, -4.0], device="cuda")
        hi = torch.tensor([10.0, 10.0, 7.0], device="cuda")
        tp["centre"].view(P, 3).copy_(torch.max(torch.min(tp["centre"].view(P, 3), hi), lo))
    for n in raw0:
        _close(m.array(n).cpu(), tp[n].detach().cpu(), ulps=8, atol=1e-8, what=("first step", n))
    _close(m.conf.cpu(), tc.detach().cpu(), ulps=8, atol=1e-8, what="first step conf")
    tr.close()
