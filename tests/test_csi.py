"""CPU tests of the multi-frequency (CSI) oracle: gs_sincos2pi against double
precision; the reductions that pin the C-channel blend (zero frequency gives
the single-channel render bit for bit; each subcarrier equals a single-channel
render of the phased emission s exp(-j 2 pi f_k L / c)); mirror against twin;
and the exact backward (including the path-length gradient to the means)
against FP64 central finite differences."""
import numpy as np

import gsparc_scenes as S
import oracle as O


def test_sincos2pi_accuracy():
    x = np.linspace(-0.5, 0.5, 1_000_001).astype(np.float32)
    c, s = O.sincos2pi(x)
    ref_c = np.cos(2 * np.pi * x.astype(np.float64))
    ref_s = np.sin(2 * np.pi * x.astype(np.float64))
    assert np.abs(c - ref_c).max() < 1.5e-7
    assert np.abs(s - ref_s).max() < 1.5e-7
    c0, s0 = O.sincos2pi(np.array([0.0, 0.25, -0.25, 0.5, -0.5], np.float32))
    np.testing.assert_allclose(c0, [1, 0, 0, -1, -1], atol=1e-7)
    np.testing.assert_allclose(s0, [0, 1, -1, 0, 0], atol=1e-7)


def _scene(seed=21, P=40, R=2):
    sc = S.small_scene(seed, P, R=R)
    sc.mlp *= 3.0
    return sc


def test_zero_frequency_reduces_to_single_channel():
    sc = _scene()
    pre, keys, vals, ranges, fwd = O.render(sc, 0)
    plen = O.path_length(sc, 0)
    f = O.render_fwd_csi(pre, plen, vals, ranges, sc.W, sc.H, np.float32, f0=0.0, df=0.0)
    for k in (0, 17, 51):
        np.testing.assert_array_equal(f["out"][:, :, k, 0], fwd["out_re"])
        np.testing.assert_array_equal(f["out"][:, :, k, 1], fwd["out_im"])
    np.testing.assert_array_equal(f["final_T"], fwd["final_T"])
    np.testing.assert_array_equal(f["n_contrib"], fwd["n_contrib"])


def test_each_subcarrier_is_a_single_channel_render_of_the_phased_emission():
    sc = _scene(22)
    pre, keys, vals, ranges, _ = O.render(sc, 1, np.float64)
    plen = O.path_length(sc, 1, np.float64)
    f = O.render_fwd_csi(pre, plen, vals, ranges, sc.W, sc.H, np.float64)
    s = pre.sig_re + 1j * pre.sig_im
    for k in (0, 25, 51):
        fk = (O.CSI_F0 + k * O.CSI_DF) / O.LIGHT_SPEED
        e = s * np.exp(-2j * np.pi * plen * fk)
        pk = O.Pre(sc.P, np.float64)
        pk.__dict__.update(pre.__dict__)
        pk.sig_re, pk.sig_im = e.real.copy(), e.imag.copy()
        one = O.render_fwd(pk, vals, ranges, sc.W, sc.H, np.float64)
        peak = np.abs(one["out_re"]).max() + np.abs(one["out_im"]).max()
        assert np.abs(f["out"][:, :, k, 0] - one["out_re"]).max() <= 1e-9 * peak
        assert np.abs(f["out"][:, :, k, 1] - one["out_im"]).max() <= 1e-9 * peak


def test_mirror_matches_twin():
    """FP32 mirror (split constants, gs_sincos2pi) against the FP64 twin on
    the same float32 preprocess outputs and path lengths."""
    sc = _scene(23, P=60)
    pre, keys, vals, ranges, _ = O.render(sc, 0)
    plen = O.path_length(sc, 0)
    f32 = O.render_fwd_csi(pre, plen, vals, ranges, sc.W, sc.H, np.float32)
    p64 = O.Pre(sc.P, np.float64)
    for k, v in pre.__dict__.items():
        setattr(p64, k, v.astype(np.float64) if v.dtype == np.float32 else v)
    f64 = O.render_fwd_csi(p64, plen.astype(np.float64), vals, ranges, sc.W, sc.H, np.float64)
    peak = np.abs(f64["out"]).max()
    assert np.abs(f32["out"] - f64["out"]).max() <= 3e-6 * peak
    rng = np.random.default_rng(3)
    dout = rng.normal(size=(sc.H, sc.W, O.CSI_NC, 2))
    g32 = O.render_bwd_csi(pre, plen, vals, ranges, f32, dout.astype(np.float32), sc.W, sc.H, np.float32)
    g64 = O.render_bwd_csi(p64, plen.astype(np.float64), vals, ranges, f64, dout, sc.W, sc.H, np.float64)
    for c in range(9):
        scale = np.abs(g64[:, c]).max()
        assert np.abs(g32[:, c] - g64[:, c]).max() <= 2e-4 * scale, c


def _as64(sc):
    return S.Scene(*(np.asarray(getattr(sc, f), np.float64).copy() for f in
                     ("means", "scales", "rotations", "opacities", "mlp", "tx_pos", "rx_pos", "views")))


def _decisions(sc, r):
    pre, plen, vals, ranges, fwd = O.render_csi(sc, r, np.float64)
    h = O.render_fwd(pre, vals, ranges, sc.W, sc.H, np.float64, want_hash=True)["hash"]
    return (pre.radii.tobytes(), vals.tobytes(), h.tobytes()), fwd


def test_csi_exact_gradients_match_fp64_finite_differences():
    sc = _as64(_scene(24, P=20, R=1))
    r = 0
    rng = np.random.default_rng(24)
    dout = rng.normal(size=(sc.H, sc.W, O.CSI_NC, 2))
    pre, plen, vals, ranges, fwd = O.render_csi(sc, r, np.float64)
    g, g9 = O.backward_csi(sc, pre, plen, vals, ranges, fwd, dout, r, np.float64)
    base, _ = _decisions(sc, r)
    fields = [("means", 3), ("opacities", 1), ("mlp", 1), ("scales", 3), ("rotations", 4)]
    vis = np.nonzero(pre.radii > 0)[0]
    checked = 0
    for trial in range(300):
        if checked >= 120:
            break
        field, per = fields[trial % len(fields)]
        flat = getattr(sc, field).reshape(-1)
        if field == "mlp":
            j = int(rng.integers(0, flat.size))
        else:
            j = int(rng.choice(vis)) * per + int(rng.integers(0, per))
        # (the phase turns ~150 rad per metre of path: a 1e-5 step keeps the
        # central difference's truncation error near 1e-7 relative)
        h = 1e-5 * max(1.0, abs(float(flat[j])))
        x0 = flat[j]
        vals_ = []
        for sgn in (1, -1):
            flat[j] = x0 + sgn * h
            dec, f = _decisions(sc, r)
            vals_.append((dec, float((f["out"] * dout).sum())))
        flat[j] = x0
        if vals_[0][0] != base or vals_[1][0] != base:
            continue
        fd = (vals_[0][1] - vals_[1][1]) / (2 * h)
        gf = {"means": g.means, "opacities": g.opacities, "mlp": g.mlp, "scales": g.scales,
              "rotations": g.rotations}[field].reshape(-1)
        an = gf[j]
        # relative bar, with a floor for entries far below the field's scale
        # (the loss sums ~3M terms: its rounding noise / 2h is ~1e-8)
        assert abs(fd - an) <= 2e-4 * max(abs(fd), abs(an)) + 1e-6 * np.abs(gf).max(), (field, j, fd, an)
        checked += 1
    assert checked >= 100, checked
