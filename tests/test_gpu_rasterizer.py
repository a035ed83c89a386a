"""GPU parity of the rasterizer (through the C ABI via wireless_rasterizer)
against the CPU oracle on the same seeded scenes.

Bar: per-Gaussian discrete outputs (radii), depth keys, point lists and tile
ranges bit-exact; spectra within 1e-5 of the peak (they are bit-exact in
practice: both sides use the same float32 operation order and gs_math.h);
gradients within 1e-4 relative to the largest magnitude of each gradient.
"""
import numpy as np
import pytest
import torch

import gsparc_scenes as S
import oracle as O
from helpers import gpu_forward, read_state, to_dev
import wireless_rasterizer as wr

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[1, 2, 4], ids=["ppl1", "ppl2", "ppl4"])
def ppl(request):
    """Run under every blend pixel mapping (quarter, half or whole tile per wave)."""
    import gsparc_hip as gh
    gh.set_option("render_fwd_ppl", request.param)
    gh.set_option("render_bwd_ppl", request.param)
    yield request.param
    gh.set_option("render_fwd_ppl", 0)
    gh.set_option("render_bwd_ppl", 0)

SPEC_TOL = 1e-5     # relative to the peak of the spectrum
GRAD_TOL = 1e-4     # relative to the largest |gradient| of the tensor


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return float(np.abs(a - b).max() / scale)


def _check_forward(scene, st, r):
    pre, keys, vals, ranges, fwd = O.render(scene, r)
    vis = pre.radii > 0
    np.testing.assert_array_equal(st["radii"][r], pre.radii)
    sp = st["splat"][r]
    np.testing.assert_array_equal(sp[vis, 0], pre.xy[vis, 0])
    np.testing.assert_array_equal(sp[vis, 1], pre.xy[vis, 1])
    np.testing.assert_array_equal(-2.0 * sp[vis, 2].astype(np.float64), pre.conic_o[vis, 0])
    np.testing.assert_array_equal(-sp[vis, 3], pre.conic_o[vis, 1])
    np.testing.assert_array_equal(-2.0 * sp[vis, 4].astype(np.float64), pre.conic_o[vis, 2])
    np.testing.assert_array_equal(sp[vis, 5], pre.conic_o[vis, 3])
    np.testing.assert_array_equal(st["depth"][r][vis], pre.depth[vis])
    np.testing.assert_array_equal(st["sig"][r][vis, 0], pre.sig_re[vis])
    np.testing.assert_array_equal(st["sig"][r][vis, 1], pre.sig_im[vis])
    # tile ranges of a batch are global offsets over all receivers; empty tiles
    # are (0, 0) as in the reference
    rg = st["ranges"][r].astype(np.int64)
    ref = ranges.astype(np.int64)
    nonempty = ref[:, 1] > ref[:, 0]
    np.testing.assert_array_equal(rg[~nonempty], 0)
    base = rg[nonempty, 0].min() if nonempty.any() else 0
    np.testing.assert_array_equal(rg[nonempty] - base, ref[nonempty])
    if r == 0:
        assert base == 0
    Kr = len(vals)
    np.testing.assert_array_equal(st["plist"][base:base + Kr], vals)
    np.testing.assert_array_equal(st["keys"][base:base + Kr], (keys & 0xffffffff).astype(np.uint32))
    peak = max(np.abs(fwd["out_re"]).max(), 1e-30)
    assert np.abs(st["out_re"][r] - fwd["out_re"]).max() <= SPEC_TOL * peak
    assert np.abs(st["out_im"][r] - fwd["out_im"]).max() <= SPEC_TOL * max(np.abs(fwd["out_im"]).max(), 1e-30)
    np.testing.assert_array_equal(st["n_contrib"][r], fwd["n_contrib"])
    assert np.abs(st["final_T"][r] - fwd["final_T"]).max() <= 1e-6
    return pre, vals, ranges, fwd


@pytest.mark.parametrize("seed,P", [(10, 64), (11, 500), (12, 3000)])
def test_forward_single_receiver(seed, P, ppl):
    sc = S.small_scene(seed, P)
    _, out = gpu_forward(sc)
    st = read_state(sc, out)
    _check_forward(sc, st, 0)


def test_forward_config0():
    sc = S.config0()
    _, out = gpu_forward(sc)
    st = read_state(sc, out)
    _check_forward(sc, st, 0)


def test_forward_batched_equals_single(ppl):
    sc = S.small_scene(21, 2000, R=5, walls=True)
    _, out = gpu_forward(sc)
    st = read_state(sc, out)
    for r in range(sc.R):
        one = S.Scene(sc.means, sc.scales, sc.rotations, sc.opacities, sc.mlp, sc.tx_pos, sc.rx_pos[r:r + 1],
                      sc.views[r:r + 1])
        _, o1 = gpu_forward(one)
        s1 = read_state(one, o1)
        np.testing.assert_array_equal(st["out_re"][r], s1["out_re"][0])
        np.testing.assert_array_equal(st["out_im"][r], s1["out_im"][0])
        np.testing.assert_array_equal(st["radii"][r], s1["radii"][0])
        _check_forward(sc, st, r)


def test_forward_long_segments():
    """Tile lists of more than 8192 instances (the 50k-Gaussian wall scene)."""
    sc = S.config2(R=3)
    _, out = gpu_forward(sc)
    st = read_state(sc, out)
    lens = st["ranges"][..., 1].astype(np.int64) - st["ranges"][..., 0]
    assert lens.max() > 8192
    for r in range(sc.R):
        _check_forward(sc, st, r)


@pytest.mark.parametrize("H,W,R", [(512, 512, 2), (528, 600, 2), (528, 600, 13), (1024, 1024, 1)],
                         ids=["1024_tiles_row_place", "1254_tiles_chunk_scatter", "16302_tile_lists",
                              "4096_tiles_global_counts"])
def test_forward_placement_paths(H, W, R):
    """Maps of up to 1024 tiles place the instances per (chunk, tile row),
    larger ones by the chunked scatter (its tile tables in LDS up to a few
    thousand tiles, in global memory beyond); up to 15360 (receiver, tile)
    lists are scanned through LDS, more by the two-pass scan.  All give the
    reference's lists (the last receiver's ranges carry the whole prefix)."""
    sc = S.small_scene(31, 1500, R=R, H=H, W=W, walls=True)
    print("placement", H, W, R, flush=True)
    _, out = gpu_forward(sc)
    st = read_state(sc, out)
    for r in sorted({0, sc.R - 1}):
        _check_forward(sc, st, r)


def test_equal_depth_ties_in_index_order():
    """Gaussians sharing a centre have identical depth keys; the reference's
    stable sort keeps them in index order (rasterizer_impl.cu:295-300)."""
    sc = S.small_scene(23, 3000, walls=True)
    sc.means[1::2] = sc.means[0::2]
    sc.means[2::3] = sc.means[0::3]
    _, out = gpu_forward(sc)
    st = read_state(sc, out)
    keys = st["keys"]
    assert (keys[1:] == keys[:-1]).sum() > 1000
    _check_forward(sc, st, 0)


def far_gaussians(sc, seed, n=200):
    """Move n Gaussians 15 km to 1000 km from receiver 0, scaled to stay
    visible."""
    rng = np.random.default_rng(seed)
    far = rng.choice(sc.P, n, replace=False)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    dist = np.exp(rng.uniform(np.log(1.5e4), np.log(1e6), size=n))
    sc.means[far] = (sc.rx_pos[0] + dirs * dist[:, None]).astype(np.float32)
    sc.scales[far] = np.repeat(0.05 * dist[:, None], 3, axis=1).astype(np.float32)
    return far


def test_far_gaussians():
    """Depths from 15 km to 1000 km beside the room's (their depth bits differ
    from the room's in every digit of the sort): the oracle's keys, lists and
    ranges."""
    sc = S.small_scene(24, 1500, R=2, walls=True)
    far = far_gaussians(sc, 24)
    _, out = gpu_forward(sc)
    st = read_state(sc, out)
    seen = (st["radii"][0][far] > 0) & (st["depth"][0][far] > 13200.0)
    assert seen.sum() > 20, seen.sum()
    for r in range(sc.R):
        _check_forward(sc, st, r)


def test_empty_and_culled():
    sc = S.small_scene(5, 50)
    # every Gaussian within 0.2 m of the receiver: all culled (auxiliary.h:143)
    sc.means[:] = sc.rx_pos[0] + 0.05
    d, out = gpu_forward(sc)
    K, ore, oim, radii = out[:4]
    assert K == 0 and int(radii.abs().sum()) == 0
    assert float(ore.abs().max()) == 0.0
    # P == 0
    z = torch.zeros((0, 3), device="cuda")
    res = wr.rasterize_gaussians(z, torch.zeros(0, device="cuda"), torch.zeros((0, 1), device="cuda"), z,
                                 torch.zeros((0, 4), device="cuda"), 90, 360, 1.0, d["views"][0], torch.eye(4),
                                 d["tx"], d["rx"][0])
    assert res[0] == 0 and res[1].shape == (1, 90, 360) and float(res[1].abs().max()) == 0.0


def test_bad_means_shape():
    with pytest.raises(RuntimeError, match="means3D"):
        wr.rasterize_gaussians(torch.zeros((4, 2), device="cuda"), None, None, None, None, 90, 360, 1.0,
                               torch.eye(4), torch.eye(4), torch.zeros(3), torch.zeros(3))


def _gpu_backward(sc, d, out, dre, dim, mode):
    K, ore, oim, radii, geom, binbuf, img = out
    res = wr.backward_batched(d["means"], d["mlp"], radii, d["opac"], d["scales"], d["rots"], 1.0, d["views"],
                              torch.from_numpy(dre).cuda(), torch.from_numpy(dim).cuda(), d["tx"], d["rx"], geom, K,
                              binbuf, img, grad_mode=mode)
    torch.cuda.synchronize()
    return [None if x is None else x.cpu().numpy() for x in res]


@pytest.mark.parametrize("mode", ["exact", "reference"])
@pytest.mark.parametrize("seed,P,R", [(31, 400, 1), (32, 3000, 3)])
def test_backward_parity(mode, seed, P, R, ppl):
    sc = S.small_scene(seed, P, R=R)
    d, out = gpu_forward(sc)
    rng = np.random.default_rng(seed)
    dre = rng.normal(size=(R, sc.H, sc.W)).astype(np.float32)
    dim = rng.normal(size=(R, sc.H, sc.W)).astype(np.float32)
    g = _gpu_backward(sc, d, out, dre, dim, mode)
    m = O.GRAD_EXACT if mode == "exact" else O.GRAD_REFERENCE
    acc = O.Grads(sc.P)
    d2 = []
    for r in range(R):
        pre, keys, vals, ranges, fwd = O.render(sc, r)
        _, g8 = O.backward(sc, pre, vals, ranges, fwd, dre[r], dim[r], m, r=r, grads=acc)
        d2.append(np.stack([g8[:, 0] * (0.5 * sc.W), g8[:, 1] * (0.5 * sc.H), 0 * g8[:, 0]], 1))
    dmeans2D, dmeans3D, dmlp, dop, dcov, dsc, drot, _ = g
    assert _rel(dmeans2D, np.stack(d2)) < GRAD_TOL
    assert _rel(dmeans3D, acc.means) < GRAD_TOL
    assert _rel(dmlp, acc.mlp) < GRAD_TOL
    assert _rel(dop, acc.opacities) < GRAD_TOL
    assert _rel(dcov, acc.cov3) < GRAD_TOL
    assert _rel(dsc, acc.scales) < GRAD_TOL
    assert _rel(drot, acc.rotations) < GRAD_TOL


@pytest.mark.parametrize("order", [1, 2, 3], ids=["global", "xcd_mixed", "xcd_receiver"])
def test_work_orders(order):
    """Every blend work order covers every (receiver, tile) exactly once,
    forward and backward, also when R is not a multiple of the 8 XCDs."""
    import gsparc_hip as gh
    gh.set_option("render_order", order)
    try:
        R = 10
        sc = S.small_scene(41, 1500, R=R, walls=True)
        d, out = gpu_forward(sc)
        st = read_state(sc, out)
        for r in range(R):
            _check_forward(sc, st, r)
        rng = np.random.default_rng(41)
        dre = rng.normal(size=(R, sc.H, sc.W)).astype(np.float32)
        dim = rng.normal(size=(R, sc.H, sc.W)).astype(np.float32)
        g = _gpu_backward(sc, d, out, dre, dim, "reference")
        acc = O.Grads(sc.P)
        for r in range(R):
            pre, keys, vals, ranges, fwd = O.render(sc, r)
            O.backward(sc, pre, vals, ranges, fwd, dre[r], dim[r], O.GRAD_REFERENCE, r=r, grads=acc)
        assert _rel(g[1], acc.means) < GRAD_TOL
        assert _rel(g[2], acc.mlp) < GRAD_TOL
        assert _rel(g[3], acc.opacities) < GRAD_TOL
    finally:
        gh.set_option("render_order", 0)


@pytest.mark.parametrize("depth", [32, 128])
@pytest.mark.parametrize("pplv", [1, 2, 4])
def test_batch_depths(pplv, depth):
    """Every LDS batch depth of the blend kernels gives the oracle's spectra,
    final T and contributor counts, and its gradients, on lists that span
    several batches and terminate inside one."""
    import gsparc_hip as gh
    for name, v in (("render_fwd_ppl", pplv), ("render_bwd_ppl", pplv), ("render_fwd_depth", depth),
                    ("render_bwd_depth", depth)):
        gh.set_option(name, v)
    try:
        R = 3
        sc = S.small_scene(43, 3000, R=R, walls=True)
        print(f"batch depth {depth} ppl {pplv}", flush=True)
        d, out = gpu_forward(sc)
        st = read_state(sc, out)
        for r in range(R):
            _check_forward(sc, st, r)
        rng = np.random.default_rng(43)
        dre = rng.normal(size=(R, sc.H, sc.W)).astype(np.float32)
        dim = rng.normal(size=(R, sc.H, sc.W)).astype(np.float32)
        g = _gpu_backward(sc, d, out, dre, dim, "reference")
        acc = O.Grads(sc.P)
        for r in range(R):
            pre, keys, vals, ranges, fwd = O.render(sc, r)
            O.backward(sc, pre, vals, ranges, fwd, dre[r], dim[r], O.GRAD_REFERENCE, r=r, grads=acc)
        assert _rel(g[1], acc.means) < GRAD_TOL
        assert _rel(g[2], acc.mlp) < GRAD_TOL
        assert _rel(g[3], acc.opacities) < GRAD_TOL
        assert _rel(g[5], acc.scales) < GRAD_TOL
    finally:
        for name in ("render_fwd_ppl", "render_bwd_ppl", "render_fwd_depth", "render_bwd_depth"):
            gh.set_option(name, 0)


def test_autograd_drop_in():
    """cuda_rasterizer API: same outputs and gradient order as the reference."""
    import cuda_rasterizer as cr
    sc = S.small_scene(41, 800)
    d = to_dev(sc)
    means = d["means"].clone().requires_grad_(True)
    means2D = torch.zeros_like(means, requires_grad=True)
    mlp = d["mlp"].clone().requires_grad_(True)
    opac = d["opac"].clone().requires_grad_(True)
    scales = d["scales"].clone().requires_grad_(True)
    rots = d["rots"].clone().requires_grad_(True)
    settings = cr.GaussianRasterizerSetting(sc.H, sc.W, 1.0, d["views"][0], torch.eye(4, device="cuda"), d["tx"],
                                            d["rx"][0])
    ras = cr.GaussianRasterizer(settings)
    re, im, radii = ras(means, means2D, mlp, opac, scales, rots)
    assert re.shape == (1, sc.H, sc.W) and radii.shape == (sc.P,) and radii.dtype == torch.int32
    (re.sum() + 0.5 * im.sum()).backward()
    for t in (means, means2D, mlp, opac, scales, rots):
        assert t.grad is not None and torch.isfinite(t.grad).all()
    pre, keys, vals, ranges, fwd = O.render(sc, 0)
    acc, g8 = O.backward(sc, pre, vals, ranges, fwd, np.ones((sc.H, sc.W), np.float32),
                         0.5 * np.ones((sc.H, sc.W), np.float32), O.GRAD_REFERENCE)
    assert _rel(means.grad.cpu().numpy(), acc.means) < GRAD_TOL
    assert _rel(opac.grad.cpu().numpy(), acc.opacities) < GRAD_TOL
    assert _rel(means2D.grad[:, 0].cpu().numpy(), g8[:, 0] * 0.5 * sc.W) < GRAD_TOL
    vis = ras.markVisible(means.detach())
    assert vis.dtype == torch.bool and bool(vis.all())
