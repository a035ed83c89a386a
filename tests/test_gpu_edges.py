"""GPU edge cases of the blend decisions and the instance count.

Near-threshold pairs: every visible Gaussian of a scene gets the opacity that
puts its pair with one chosen pixel exactly on the alpha = 1/255 boundary
(o * exp(power) reaches 1/255 at this opacity and not one float below), or
one float below it.  The forward must keep or skip each such pair as the
oracle does (contributor counts bit-exact), and the backward must walk the
same contributors: gradients within 1e-4 of the largest magnitude, in both
gradient modes.

Saturation: a batch whose tile lists would hold more than 2^32 instances
reads back as K = 2^32 - 1 (never a wrapped count); the exact forward refuses
it and a capacity forward reports it as an overflow with NaN spectra."""
import numpy as np
import pytest
import torch

import gsparc_scenes as S
import oracle as O
import wireless_rasterizer as wr
from helpers import gpu_forward, read_state

pytestmark = pytest.mark.gpu

AMIN = np.float32(1.0 / 255.0)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _boundary_opacity(e):
    """Smallest float32 o in (0, 1] with fl(o * e) >= 1/255, per entry of e
    (NaN where even o = 1 stays below)."""
    e = e.astype(np.float32)
    lo = np.zeros(e.shape, np.uint32)                            # o = 0 fails
    hi = np.full(e.shape, np.float32(1.0).view(np.uint32))        # o = 1
    ok = (np.float32(1.0) * e) >= AMIN
    for _ in range(32):
        mid = lo + (hi - lo) // 2
        passes = mid.view(np.float32) * e >= AMIN
        hi = np.where(passes, mid, hi)
        lo = np.where(passes, lo, mid)
    o = hi.view(np.float32).copy()
    o[~ok] = np.nan
    return o


def _planted_scene(seed):
    sc = S.small_scene(seed, 600, R=1)
    pre = O.preprocess(sc, 0)
    px = np.clip(np.floor(pre.xy[:, 0]).astype(np.int64) + 2, 0, sc.W - 1)
    py = np.clip(np.floor(pre.xy[:, 1]).astype(np.int64) + 1, 0, sc.H - 1)
    e = O.math_vec("exp", O.pair_power(pre, px, py))
    o = _boundary_opacity(e)
    rng = np.random.default_rng(seed)
    below = rng.uniform(size=sc.P) < 0.5
    o_plant = np.where(below, np.nextafter(o, np.float32(0)), o)
    plant = (pre.radii > 0) & np.isfinite(o)
    sc.opacities = np.where(plant[:, None], o_plant[:, None], sc.opacities).astype(np.float32)
    return sc, plant, px, py


@pytest.mark.parametrize("mode", ["exact", "reference"])
def test_near_threshold_pairs_match_oracle(mode):
    sc, plant, px, py = _planted_scene(131)
    assert plant.sum() > 300
    print("planted pairs", int(plant.sum()), mode, flush=True)
    # the planted pairs sit on the boundary in the oracle: kept at o, skipped one float below
    pre = O.preprocess(sc, 0)
    e = O.math_vec("exp", O.pair_power(pre, px, py))
    keep = np.minimum(np.float32(0.99), pre.conic_o[:, 3] * e) >= AMIN
    assert 0 < keep[plant].sum() < plant.sum()
    import gsparc_hip as gh
    gh.profile_enable(True, timing=False, counting=True)
    try:
        d, out = gpu_forward(sc)
        K, ore, oim, radii, geom, binbuf, img = out
        g1 = torch.ones_like(ore)
        wr.backward_batched(d["means"], d["mlp"], radii, d["opac"], d["scales"], d["rots"], 1.0, d["views"], g1, g1,
                            d["tx"], d["rx"], geom, K, binbuf, img, grad_mode=mode)
        torch.cuda.synchronize()
        cnt = gh.profile_counts()
    finally:
        gh.profile_enable(False)
    # the backward walked exactly the forward's contributors
    assert cnt["fwd_contrib"] == cnt["bwd_contrib"], cnt
    st = read_state(sc, out)
    pre, keys, vals, ranges, fwd = O.render(sc, 0)
    assert cnt["fwd_contrib"] == int(fwd["stats"][3])
    np.testing.assert_array_equal(st["n_contrib"][0], fwd["n_contrib"])
    np.testing.assert_array_equal(st["final_T"][0], fwd["final_T"])
    peak = max(np.abs(fwd["out_re"]).max(), 1e-30)
    assert np.abs(st["out_re"][0] - fwd["out_re"]).max() <= 1e-5 * peak
    rng = np.random.default_rng(7)
    dre = rng.normal(size=(1, sc.H, sc.W)).astype(np.float32)
    dim = rng.normal(size=(1, sc.H, sc.W)).astype(np.float32)
    K, ore, oim, radii, geom, binbuf, img = out
    g = wr.backward_batched(d["means"], d["mlp"], radii, d["opac"], d["scales"], d["rots"], 1.0, d["views"],
                            torch.from_numpy(dre).cuda(), torch.from_numpy(dim).cuda(), d["tx"], d["rx"], geom, K,
                            binbuf, img, grad_mode=mode)
    torch.cuda.synchronize()
    m = O.GRAD_EXACT if mode == "exact" else O.GRAD_REFERENCE
    acc, g8 = O.backward(sc, pre, vals, ranges, fwd, dre[0], dim[0], m)
    d2, d3, dmlp, dop, dcov, dsc, drot, _ = [None if x is None else x.cpu().numpy() for x in g]
    assert _rel(d3, acc.means) < 1e-4
    assert _rel(dmlp, acc.mlp) < 1e-4
    assert _rel(dop, acc.opacities) < 1e-4
    assert _rel(dsc, acc.scales) < 1e-4
    assert _rel(drot, acc.rotations) < 1e-4
    # the opacity gradient of each planted Gaussian on its own scale: a pair
    # walked by one side only would move it by G dL/dalpha at its pixel
    ref = acc.opacities[:, 0]
    err = np.abs(dop.reshape(-1) - ref)
    assert np.all(err[plant] <= 1e-4 * np.abs(ref).max() + 2e-3 * np.abs(ref[plant]))


def test_instance_count_saturates_past_2_32():
    """1M Gaussians larger than the room seen from 32 receivers: every one of
    the 32M records covers all 138 tiles, K = 4.4e9 > 2^32."""
    P, R, H, W = 1_000_000, 32, 90, 360
    dev = "cuda"
    g = torch.Generator(device=dev).manual_seed(3)
    means = torch.rand((P, 3), device=dev, generator=g) * torch.tensor([6.0, 6.0, 3.0], device=dev)
    scales = torch.full((P, 3), 20.0, device=dev)
    rots = torch.zeros((P, 4), device=dev)
    rots[:, 0] = 1.0
    opac = torch.full((P,), 0.5, device=dev)
    rxn = (0.5 + np.random.default_rng(3).uniform(size=(R, 3)) * (S.ROOM - 1.0)).astype(np.float32)
    rx = torch.from_numpy(rxn).to(dev)
    views = torch.from_numpy(np.stack([S.view_matrix(r) for r in rxn])).to(dev)
    tx = torch.from_numpy(S.TX_POS.copy()).to(dev)
    em = torch.zeros((R, P, 2), device=dev)
    print("saturation: exact forward", flush=True)
    with pytest.raises(RuntimeError, match="2\\^32 - 1 or more"):
        wr.forward_batched(means, None, opac, scales, rots, H, W, 1.0, views, tx, rx, emission=em)
    torch.cuda.synchronize()
    print("saturation: capacity forward", flush=True)
    Kc, ore, oim, radii, geom, binbuf, img = wr.forward_batched(means, None, opac, scales, rots, H, W, 1.0, views,
                                                                tx, rx, emission=em, capacity=1 << 20)
    torch.cuda.synchronize()
    assert int((radii > 0).sum()) > 0.99 * P * R
    assert wr.instance_count(P, R, H, W, img) == (1 << 32) - 1
    assert bool(torch.isnan(ore).all()) and bool(torch.isnan(oim).all())
    dre = torch.ones_like(ore)
    gr = wr.backward_batched(means, None, radii, opac, scales, rots, 1.0, views, dre, dre, tx, rx, geom, Kc, binbuf,
                             img, emission=em)
    torch.cuda.synchronize()
    assert float(gr[1].abs().max()) == 0.0 and float(gr[3].abs().max()) == 0.0


@pytest.mark.parametrize("depth", [32, 64, 128])
@pytest.mark.parametrize("ppl", [1, 2, 4])
def test_blend_counters_count_the_reference_loops(ppl, depth):
    """The counting variants count, in the reference's terms, the pairs its
    forward visits, the pairs its backward walks and the contributing pairs;
    the oracle's loops give the same numbers exactly, and the forward and the
    backward find the same contributing pairs, under every pixel mapping and
    batch depth."""
    import gsparc_hip as gh
    sc = S.small_scene(141, 2000, R=3, walls=True)
    gh.set_option("render_fwd_ppl", ppl)
    gh.set_option("render_bwd_ppl", ppl)
    gh.set_option("render_fwd_depth", depth)
    gh.set_option("render_bwd_depth", depth)
    gh.profile_enable(True, timing=False, counting=True)
    try:
        d, out = gpu_forward(sc)
        K, ore, oim, radii, geom, binbuf, img = out
        g = torch.ones_like(ore)
        wr.backward_batched(d["means"], d["mlp"], radii, d["opac"], d["scales"], d["rots"], 1.0, d["views"], g, g,
                            d["tx"], d["rx"], geom, K, binbuf, img)
        torch.cuda.synchronize()
        cnt = gh.profile_counts()
    finally:
        gh.profile_enable(False)
        for name in ("render_fwd_ppl", "render_bwd_ppl", "render_fwd_depth", "render_bwd_depth"):
            gh.set_option(name, 0)
    visited = blended = walked = 0
    for r in range(sc.R):
        pre, keys, vals, ranges, fwd = O.render(sc, r)
        visited += int(fwd["stats"][0])
        blended += int(fwd["stats"][3])
        walked += int(fwd["n_contrib"].sum())
    assert cnt == {"fwd_pairs": visited, "bwd_pairs": walked, "fwd_contrib": blended, "bwd_contrib": blended}, cnt
