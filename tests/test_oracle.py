"""CPU tests that pin the oracle itself (no GPU): closed forms, an independent
brute-force renderer, an independent sort, textbook libraries (scipy
rotations, torch autograd, conv2d SSIM) and FP64 central finite differences.
A mistake such as a dropped term, a wrong sign, a transposed operand or a
wrong index in the oracle fails at least one of these."""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

import gsparc_scenes as S
import oracle as O

W, H = 360, 90


def one_gaussian_scene(mean, scale, quat=(1, 0, 0, 0), opac=0.8, emis=0.6, rx=(3.0, 3.0, 1.5)):
    """Single Gaussian, kept in float64 so that finite differences see the exact inputs."""
    means = np.asarray([mean], np.float64)
    v = np.eye(4)
    v[:3, 3] = -np.asarray(rx, np.float64)
    sc = S.Scene(means, np.asarray([scale], np.float64), np.asarray([quat], np.float64),
                 np.asarray([[opac]], np.float64), S.constant_emission_mlp(means.astype(np.float32), np.array([emis])),
                 S.TX_POS.astype(np.float64), np.asarray([rx], np.float64), v.T.copy()[None])
    return sc


# ------------------------------------------------------------- projection
def test_projection_closed_form_on_axis():
    """Gaussian straight ahead (+z) at distance r: px, py at the map centre and
    cov2D = diag((W/2pi)^2 s^2/r^2, (H/pi)^2 s^2/r^2) + 0.3."""
    rx = np.array([3.0, 3.0, 1.5])
    r, s = 2.0, 0.07
    sc = one_gaussian_scene(rx + [0, 0, r], [s, s, s], rx=rx)
    pre = O.preprocess(sc, dtype=np.float64)
    assert pre.radii[0] > 0
    np.testing.assert_allclose(pre.xy[0], [(W - 1) / 2, (H - 1) / 2], atol=1e-9)
    a = (W / (2 * np.pi)) ** 2 * s * s / (r * r) + 0.3
    c = (H / np.pi) ** 2 * s * s / (r * r) + 0.3
    np.testing.assert_allclose(pre.conic_o[0, :3], [1 / a, 0.0, 1 / c], rtol=1e-6, atol=1e-12)
    assert pre.depth[0] == pytest.approx(r)
    assert pre.radii[0] == int(np.ceil(3 * np.sqrt(max(a, c))))


def test_cov2d_matches_finite_difference_jacobian():
    """cov2D = J Sigma J^T + 0.3 I with J from FD of the projected centre and
    Sigma from scipy's rotation of the quaternion (independent of the oracle)."""
    rng = np.random.default_rng(3)
    rx = np.array([3.0, 3.0, 1.5])
    for _ in range(20):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if abs(d[1]) > 0.9:
            continue
        mean = rx + d * rng.uniform(0.8, 3.0)
        s = np.exp(rng.normal(-2.5, 0.3, 3))
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        sc = one_gaussian_scene(mean, s, q, rx=rx)
        pre = O.preprocess(sc, dtype=np.float64)
        if pre.radii[0] == 0:
            continue
        J = np.zeros((2, 3))
        h = 1e-6
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            p1 = O.preprocess(one_gaussian_scene(mean + e, s, q, rx=rx), dtype=np.float64).xy[0]
            p0 = O.preprocess(one_gaussian_scene(mean - e, s, q, rx=rx), dtype=np.float64).xy[0]
            J[:, k] = (p1 - p0) / (2 * h)
        Rm = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()
        Sig = Rm @ np.diag(s * s) @ Rm.T
        iu = np.triu_indices(3)
        np.testing.assert_allclose(pre.cov3[0], Sig[iu], rtol=1e-9, atol=1e-14)
        cov = J @ Sig @ J.T + 0.3 * np.eye(2)
        qa, qb, qc = pre.conic_o[0, :3]
        got = np.linalg.inv(np.array([[qa, qb], [qb, qc]]))
        np.testing.assert_allclose(got, cov, rtol=2e-5)


def test_rect_integer_semantics():
    """The tile rect of auxiliary.h:14-24 in integer terms: tiles that hold the
    pixels floor(px - r) .. floor(px + r) - 1 (and the same in y), clamped to
    the grid.  Checked for every visible Gaussian of a scene."""
    sc = S.small_scene(4, 400)
    pre = O.preprocess(sc)
    gx, gy = (W + 15) // 16, (H + 15) // 16
    nvis = 0
    for i in np.nonzero(pre.radii > 0)[0]:
        r = np.float32(pre.radii[i])
        for c, (lo, hi, g) in enumerate(((0, 2, gx), (1, 3, gy))):
            p = np.float32(pre.xy[i, c])
            a = int(np.floor(np.float32(p - r)))
            b = int(np.floor(np.float32(p + r)))
            assert pre.rect[i, lo] == min(g, max(0, a) // 16)
            assert pre.rect[i, hi] == min(g, -(-max(0, b) // 16))
        x0, y0, x1, y1 = pre.rect[i]
        assert pre.tiles[i] == (x1 - x0) * (y1 - y0) > 0
        nvis += 1
    assert nvis > 300


def test_cull_near_receiver():
    rx = np.array([3.0, 3.0, 1.5])
    pre = O.preprocess(one_gaussian_scene(rx + [0.1, 0.1, 0.1], [0.05] * 3, rx=rx))
    assert pre.radii[0] == 0 and pre.tiles[0] == 0


# ------------------------------------------------------------- emission MLP
def test_mlp_constant_emission_closed_form():
    """W1 = b1 = W2 = 0 gives sigmoid(b2 - log(d + 1e-6)) (forward.cu:46-72)."""
    sc = S.config0(P=300)
    pre = O.preprocess(sc, dtype=np.float64)
    rng = np.random.Generator(np.random.PCG64(0))
    S.uniform_means(rng, 300)
    S._common(rng, 300)
    emis = 1 / (1 + np.exp(-rng.normal(size=300)))
    vis = pre.radii > 0
    np.testing.assert_allclose(pre.sig_re[vis], emis[vis], rtol=2e-6)
    np.testing.assert_allclose(pre.sig_im[vis], emis[vis], rtol=2e-6)


def test_mlp_backward_matches_torch_autograd():
    """MLP part of the backward vs torch autograd of the 3->32->2 definition."""
    sc = S.small_scene(8, 40)
    P = sc.P
    pre = O.preprocess(sc, dtype=np.float64)
    rng = np.random.default_rng(8)
    g8 = np.zeros((P, 8))
    g8[:, 6] = rng.normal(size=P)
    g8[:, 7] = rng.normal(size=P)
    g = O.preprocess_bwd(sc, pre, g8, O.GRAD_EXACT, dtype=np.float64)
    m = torch.tensor(sc.mlp, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(sc.means, dtype=torch.float64, requires_grad=True)
    W1 = m[:P * 96].view(P, 32, 3)
    b1 = m[P * 96:P * 128].view(P, 32)
    W2 = m[P * 128:P * 192].view(P, 2, 32)
    b2 = m[P * 192:].view(P, 2)
    x = torch.tensor(sc.rx_pos[0], dtype=torch.float64)
    # the reference's float literals 0.01f and 1e-6f (forward.cu:27,48)
    slope, eps = float(np.float32(0.01)), float(np.float32(1e-6))
    h = torch.nn.functional.leaky_relu(W1 @ x + b1, slope)
    z = (W2 @ h.unsqueeze(-1)).squeeze(-1) + b2
    d = torch.linalg.norm(torch.tensor(sc.tx_pos, dtype=torch.float64) - mu, dim=1)
    sig = torch.sigmoid(z - torch.log(d + eps)[:, None])
    vis = torch.tensor(pre.radii > 0)
    loss = (sig * torch.tensor(g8[:, 6:8]) * vis[:, None]).sum()
    loss.backward()
    np.testing.assert_allclose(g.mlp, m.grad.numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(g.means, mu.grad.numpy(), rtol=1e-9, atol=1e-12)


# ------------------------------------------------------------- binning
def test_binning_matches_independent_sort():
    sc = S.small_scene(9, 1500, walls=True)
    pre = O.preprocess(sc)
    keys, vals, ranges = O.bin_gaussians(pre, W, H)
    gx = (W + 15) // 16
    tiles, depth, idx = [], [], []
    for i in np.nonzero(pre.radii > 0)[0]:
        x0, y0, x1, y1 = pre.rect[i]
        for y in range(y0, y1):
            for x in range(x0, x1):
                tiles.append(y * gx + x)
                depth.append(pre.depth[i].view(np.uint32))
                idx.append(i)
    tiles, depth, idx = map(np.asarray, (tiles, depth, idx))
    order = np.lexsort((idx, depth, tiles))
    np.testing.assert_array_equal(vals, idx[order])
    np.testing.assert_array_equal(keys >> np.uint64(32), tiles[order])
    for t in range(O.n_tiles(W, H)):
        sel = np.nonzero(tiles[order] == t)[0]
        if len(sel):
            assert ranges[t, 0] == sel[0] and ranges[t, 1] == sel[-1] + 1
        else:
            assert ranges[t, 0] == ranges[t, 1] == 0


# ------------------------------------------------------------- blend
def brute_render(pre, W, H, sig_re):
    """Independent per-pixel renderer in numpy (float64): front-to-back over the
    Gaussians whose tile rect contains the pixel's tile, sorted by depth."""
    out = np.zeros((H, W))
    T_out = np.ones((H, W))
    vis = np.nonzero(pre.radii > 0)[0]
    for py in range(H):
        for px in range(W):
            tx, ty = px // 16, py // 16
            cand = [i for i in vis if pre.rect[i, 0] <= tx < pre.rect[i, 2] and pre.rect[i, 1] <= ty < pre.rect[i, 3]]
            cand.sort(key=lambda i: (np.float32(pre.depth[i]).view(np.uint32), i))
            T, acc = 1.0, 0.0
            for i in cand:
                dx, dy = pre.xy[i, 0] - px, pre.xy[i, 1] - py
                qa, qb, qc, o = pre.conic_o[i]
                power = -0.5 * (qa * dx * dx + qc * dy * dy) - qb * dx * dy
                if power > 0:
                    continue
                alpha = min(0.99, o * np.exp(power))
                if alpha < np.float32(1 / 255):
                    continue
                if T * (1 - alpha) < np.float32(1e-4):
                    break
                acc += sig_re[i] * alpha * T
                T *= 1 - alpha
            out[py, px] = acc
            T_out[py, px] = T
    return out, T_out


def test_blend_matches_brute_force():
    sc = S.small_scene(12, 60)
    pre, keys, vals, ranges, fwd = O.render(sc, dtype=np.float64)
    ref, T = brute_render(pre, W, H, pre.sig_re)
    np.testing.assert_allclose(fwd["out_re"], ref, atol=1e-12)
    np.testing.assert_allclose(fwd["final_T"], T, atol=1e-12)


def test_blend_two_gaussians_closed_form():
    rx = np.array([3.0, 3.0, 1.5])
    means = np.array([rx + [0, 0, 1.0], rx + [0.01, 0, 2.0]], np.float32)
    sc = S.Scene(means, np.full((2, 3), 0.08, np.float32), np.tile(np.float32([1, 0, 0, 0]), (2, 1)),
                 np.array([[0.6], [0.9]], np.float32), S.constant_emission_mlp(means, np.array([0.3, 0.7])),
                 S.TX_POS.copy(), rx[None].astype(np.float32), S.view_matrix(rx)[None])
    pre, keys, vals, ranges, fwd = O.render(sc, dtype=np.float64)
    px, py = int(round(pre.xy[0, 0])), int(round(pre.xy[0, 1]))

    def alpha(i):
        dx, dy = pre.xy[i, 0] - px, pre.xy[i, 1] - py
        qa, qb, qc, o = pre.conic_o[i]
        return min(0.99, o * np.exp(-0.5 * (qa * dx * dx + qc * dy * dy) - qb * dx * dy))

    a1, a2 = alpha(0), alpha(1)
    assert a1 > 1 / 255 and a2 > 1 / 255
    expect = pre.sig_re[0] * a1 + pre.sig_re[1] * a2 * (1 - a1)
    assert fwd["out_re"][py, px] == pytest.approx(expect, rel=1e-12)
    assert fwd["final_T"][py, px] == pytest.approx((1 - a1) * (1 - a2), rel=1e-12)
    assert fwd["n_contrib"][py, px] == 2


# ------------------------------------------------------------- finite differences
def _decisions(sc):
    pre, keys, vals, ranges, fwd = O.render(sc, dtype=np.float64)
    f = O.render_fwd(pre, vals, ranges, sc.W, sc.H, np.float64, want_hash=True)
    return (pre.radii.tobytes(), vals.tobytes(), f["hash"].tobytes()), f


def _loss(sc, wre, wim):
    dec, f = _decisions(sc)
    return dec, float((f["out_re"] * wre).sum() + (f["out_im"] * wim).sum())


FD_FIELDS = [("means", 3), ("scales", 3), ("rotations", 4), ("opacities", 1), ("mlp", 1)]


def test_exact_gradients_match_fp64_finite_differences():
    """200 random parameters of a small scene: the exact-mode backward equals
    central differences of the FP64 forward (skipping probes that change a
    discrete decision: culling, rects, depth order, alpha skips, termination)."""
    sc = S.small_scene(13, 25, H=90, W=360)
    sc.mlp *= 3.0
    rng = np.random.default_rng(13)
    wre = rng.normal(size=(H, W))
    wim = rng.normal(size=(H, W))
    pre, keys, vals, ranges, fwd = O.render(sc, dtype=np.float64)
    g, _ = O.backward(sc, pre, vals, ranges, fwd, wre, wim, O.GRAD_EXACT, dtype=np.float64)
    base_dec, _ = _decisions(sc)
    checked, worst = 0, 0.0
    for trial in range(400):
        if checked >= 200:
            break
        field, _ = FD_FIELDS[trial % len(FD_FIELDS)]
        arr = getattr(sc, field)
        flat = arr.reshape(-1)
        if field == "mlp":
            j = int(rng.integers(0, flat.size))
        else:
            i = int(rng.choice(np.nonzero(pre.radii > 0)[0]))
            per = flat.size // sc.P
            j = i * per + int(rng.integers(0, per))
        h = 1e-4 * max(1.0, abs(float(flat[j])))
        sc64 = S.Scene(*(np.asarray(getattr(sc, f), np.float64).copy() for f in
                         ("means", "scales", "rotations", "opacities", "mlp", "tx_pos", "rx_pos", "views")))
        a64 = getattr(sc64, field).reshape(-1)
        x0 = a64[j]
        a64[j] = x0 + h
        dp, lp = _loss(sc64, wre, wim)
        a64[j] = x0 - h
        dm, lm = _loss(sc64, wre, wim)
        if dp != base_dec or dm != base_dec:
            continue
        fd = (lp - lm) / (2 * h)
        gfield = {"means": g.means, "scales": g.scales, "rotations": g.rotations, "opacities": g.opacities,
                  "mlp": g.mlp}[field].reshape(-1)
        an = gfield[j]
        err = abs(fd - an) / max(abs(fd), abs(an), 1e-6)
        worst = max(worst, err)
        assert err < 2e-4, (field, j, fd, an)
        checked += 1
    assert checked >= 150, checked


def test_reference_mode_equals_exact_where_reference_is_exact():
    """Q1/Q2 of the reference backward affect only geometry and the imaginary
    channel.  With dL/dout_imag = 0 and no clamped alpha, the reference mode's
    opacity, MLP and screen-space gradients equal the exact ones."""
    sc = S.small_scene(14, 300)
    sc.opacities[:] = np.minimum(sc.opacities, 0.9)
    pre, keys, vals, ranges, fwd = O.render(sc, dtype=np.float64)
    rng = np.random.default_rng(14)
    wre = rng.normal(size=(H, W))
    zero = np.zeros((H, W))
    ge, g8e = O.backward(sc, pre, vals, ranges, fwd, wre, zero, O.GRAD_EXACT, dtype=np.float64)
    gr, g8r = O.backward(sc, pre, vals, ranges, fwd, wre, zero, O.GRAD_REFERENCE, dtype=np.float64)
    np.testing.assert_allclose(g8r, g8e, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(gr.opacities, ge.opacities, rtol=1e-12)
    np.testing.assert_allclose(gr.mlp, ge.mlp, rtol=1e-12, atol=1e-18)
    # the halved x-row makes the geometry gradients differ (backward.cu:172-173)
    assert np.abs(gr.means - ge.means).max() > 1e-3 * np.abs(ge.means).max()


# ------------------------------------------------------------- SSIM
def textbook_ssim_np(a, b):
    import torch.nn.functional as F
    # window built as in fused-ssim/tests/test.py:14-16 (float32 samples / float32 sum)
    g = torch.Tensor([np.exp(-(x - 5) ** 2 / (2 * 1.5 ** 2)) for x in range(11)])
    g = (g / g.sum()).double()
    ta, tb = torch.from_numpy(a), torch.from_numpy(b)
    ch = ta.size(1)
    w = (g[:, None] @ g[None, :]).expand(ch, 1, 11, 11)
    mu1 = F.conv2d(ta, w, padding=5, groups=ch)
    mu2 = F.conv2d(tb, w, padding=5, groups=ch)
    s1 = F.conv2d(ta * ta, w, padding=5, groups=ch) - mu1 ** 2
    s2 = F.conv2d(tb * tb, w, padding=5, groups=ch) - mu2 ** 2
    s12 = F.conv2d(ta * tb, w, padding=5, groups=ch) - mu1 * mu2
    return (((2 * mu1 * mu2 + O.C1) * (2 * s12 + O.C2)) /
            ((mu1 ** 2 + mu2 ** 2 + O.C1) * (s1 + s2 + O.C2))).numpy()


def test_ssim_taps_equal_reference_constants():
    # G_00..G_05 of fused-ssim/ssim.cu:9-14 (data values)
    ref = [0.001028380123898387, 0.0075987582094967365, 0.036000773310661316, 0.10936068743467331,
           0.21300552785396576, 0.26601171493530273]
    # the oracle's taps, read back through the SSIM of a unit impulse
    img = np.zeros((1, 1, 11, 11), np.float32)
    img[0, 0, 5, 5] = 1.0
    m, d1, d2, d3 = O.ssim_fwd(img, np.zeros_like(img), dtype=np.float32)
    g = np.ones_like(img)
    zero = np.zeros_like(img)
    # backward with dL/dmap = 1 and only dm/dmu1 = impulse returns the window row x column
    imp = np.zeros_like(img)
    imp[0, 0, 5, 5] = 1.0
    win = O.ssim_bwd(img, np.zeros_like(img), g, imp, zero, zero, dtype=np.float32)[0, 0]
    row = win[5] / win[5, 5] * np.float32(ref[5])
    np.testing.assert_allclose(row[:6], np.float32(ref), rtol=1e-6)
    np.testing.assert_allclose(row[5:], row[5::-1], rtol=1e-6)


def test_ssim_forward_matches_textbook():
    rng = np.random.default_rng(5)
    a = rng.uniform(size=(2, 3, 40, 70))
    b = rng.uniform(size=(2, 3, 40, 70))
    m, *_ = O.ssim_fwd(a, b)
    np.testing.assert_allclose(m, textbook_ssim_np(a, b), rtol=1e-6, atol=1e-9)
    same, *_ = O.ssim_fwd(a, a)
    np.testing.assert_allclose(same, 1.0, atol=1e-12)


def test_ssim_backward_matches_fd():
    rng = np.random.default_rng(6)
    a = rng.uniform(size=(1, 1, 20, 30))
    b = rng.uniform(size=(1, 1, 20, 30))
    g = rng.normal(size=a.shape)
    m, d1, d2, d3 = O.ssim_fwd(a, b)
    an = O.ssim_bwd(a, b, g, d1, d2, d3)
    h = 1e-6
    for _ in range(40):
        y, x = int(rng.integers(0, 20)), int(rng.integers(0, 30))
        ap = a.copy(); ap[0, 0, y, x] += h
        am = a.copy(); am[0, 0, y, x] -= h
        fd = ((O.ssim_fwd(ap, b)[0] - O.ssim_fwd(am, b)[0]) * g).sum() / (2 * h)
        assert abs(fd - an[0, 0, y, x]) <= 1e-6 * max(1.0, abs(fd))
