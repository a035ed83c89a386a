"""Config 1 (single-receiver training step with the shared 68->128->128->1
emission MLP).  CPU, FP64: torch float64 for the MLP, the oracle's FP64 twin
for the rasterizer and SSIM.

The full step's gradient (oracle backward in exact mode for the rasterizer,
chained through autograd of the MLP and its input features) is checked by
central finite differences of the FP64 loss on 200 random parameters: MLP
weights and biases, Gaussian means (which enter both the projection and the
MLP features) and opacities.  Probes that flip a discrete decision (culling,
rects, list order, alpha skips, termination) are skipped."""
import numpy as np
import torch

import mlp_reference as GE
import gsparc_scenes as S
import oracle as O


def _mlp64(seed=0):
    g = torch.Generator().manual_seed(seed)
    m = GE.TorchSharedMLP(generator=g, dtype=torch.float64)
    with torch.no_grad():
        m.b3 += 0.5   # positive amplitudes on average
    return m


def test_positional_encoding_closed_form():
    mu = torch.tensor([[0.0, 0.25, -0.5]], dtype=torch.float64)
    pe = GE.positional_encoding(mu)[0]
    assert pe.shape == (63,)
    np.testing.assert_allclose(pe[:3], [0.0, 0.25, -0.5])
    # k = 0: sin(pi mu), cos(pi mu); k = 1: sin(2 pi mu), cos(2 pi mu)
    np.testing.assert_allclose(pe[3:6], [0.0, np.sin(np.pi / 4), -1.0], atol=1e-15)
    np.testing.assert_allclose(pe[6:9], [1.0, np.cos(np.pi / 4), 0.0], atol=1e-15)
    np.testing.assert_allclose(pe[9:12], [0.0, 1.0, 0.0], atol=1e-15)
    np.testing.assert_allclose(pe[12:15], [1.0, 0.0, -1.0], atol=1e-15)


def test_receiver_features_and_split_layer():
    rng = np.random.default_rng(0)
    means = torch.from_numpy(rng.uniform(0, 6, (7, 3)))
    rx = torch.tensor([[3.0, 3.0, 1.5], [1.0, 2.0, 0.5]], dtype=torch.float64)
    tx = torch.tensor([0.0, 0.5, 3.0], dtype=torch.float64)
    f = GE.receiver_features(means, rx, tx)
    d = np.linalg.norm(rx.numpy()[:, None] - means.numpy()[None], axis=-1)
    np.testing.assert_allclose(f[..., 0], d)
    np.testing.assert_allclose(f[..., 1], 1.0 / np.maximum(d * d, 0.04))
    t = tx.numpy() - means.numpy()
    np.testing.assert_allclose(f[1, :, 2:], t / np.linalg.norm(t, axis=-1, keepdims=True))
    # the split first layer equals the plain 68-input layer
    m = _mlp64()
    x = torch.cat([GE.positional_encoding(means)[None].expand(2, -1, -1), f], dim=-1)
    h2 = torch.relu(torch.relu(x @ m.W1.t() + m.b1) @ m.W2.t() + m.b2)
    ref = (h2 @ m.w3.t()).squeeze(-1) + m.b3
    np.testing.assert_allclose(m.amplitude(means, rx, tx).detach(), ref.detach(), rtol=1e-12, atol=1e-12)


def _scene():
    sc = S.small_scene(61, 40, H=90, W=360)
    return S.Scene(*(np.asarray(getattr(sc, f), np.float64).copy() for f in
                     ("means", "scales", "rotations", "opacities", "mlp", "tx_pos", "rx_pos", "views")))


def _forward(sc, mlp, target, dtype=np.float64):
    """Step in the oracle's FP64 twin (or FP32 mirror), MLP in torch float64:
    (decisions, loss, state for the backward)."""
    means = torch.from_numpy(np.asarray(sc.means, np.float64)).requires_grad_(True)
    E = mlp(means, torch.from_numpy(np.asarray(sc.rx_pos, np.float64)), torch.from_numpy(np.asarray(sc.tx_pos, np.float64)))
    pre = O.preprocess(sc, 0, dtype)
    pre.sig_re[:] = E[0, :, 0].detach().numpy()
    pre.sig_im[:] = 0.0
    keys, vals, ranges = O.bin_gaussians(pre, sc.W, sc.H)
    f = O.render_fwd(pre, vals, ranges, sc.W, sc.H, dtype, want_hash=True)
    img = f["out_re"][None, None].astype(np.float64)
    m, d1, d2, d3 = O.ssim_fwd(img, target)
    N = img.size
    loss = 0.8 * np.abs(img - target).sum() / N + 0.2 * (1.0 - m.sum() / N)
    dec = (pre.radii.tobytes(), vals.tobytes(), f["hash"].tobytes(), np.sign(img - target).tobytes())
    return dec, float(loss), (pre, vals, ranges, f, E, means, d1, d2, d3, img)


def _gradients(sc, mlp, target, dtype=np.float64):
    """Gradient of the step: oracle backward (exact mode) chained through
    autograd of the MLP; dL/dmeans sums the projection and feature paths."""
    _, _, (pre, vals, ranges, f, E, means, d1, d2, d3, img) = _forward(sc, mlp, target, dtype)
    N = img.size
    dimg = 0.8 * np.sign(img - target) / N + O.ssim_bwd(img, target, np.full_like(img, -0.2 / N), d1, d2, d3)
    g8 = O.render_bwd(pre, vals, ranges, f, dimg[0, 0], np.zeros_like(dimg[0, 0]), sc.W, sc.H, O.GRAD_EXACT,
                      dtype)
    # geometry: the emission enters only through g8[:, 6:8], handled by autograd below
    g8_geo = g8.copy()
    g8_geo[:, 6:] = 0.0
    g = O.preprocess_bwd(sc, pre, g8_geo, O.GRAD_EXACT, 0, dtype)
    mlp.zero_grad()
    E[0, :, 0].backward(torch.from_numpy(g8[:, 6].copy()))
    grads = {n: p.grad.numpy().copy() for n, p in mlp.named_parameters()}
    grads["means"] = g.means + means.grad.numpy()
    grads["opacities"] = g.opacities
    return grads


def test_config1_step_gradients_match_fp64_finite_differences():
    sc = _scene()
    mlp = _mlp64(1)
    rng = np.random.default_rng(62)
    target = rng.uniform(0.0, 0.3, (1, 1, sc.H, sc.W))
    an = _gradients(sc, mlp, target)
    base_dec, _, (pre, *_rest) = _forward(sc, mlp, target)
    visible = np.nonzero(pre.radii > 0)[0]
    fields = ["W1", "b1", "W2", "b2", "w3", "b3", "means", "opacities"]
    checked = 0
    for trial in range(600):
        if checked >= 200:
            break
        field = fields[trial % len(fields)]
        if field in ("means", "opacities"):
            arr = getattr(sc, field).reshape(-1)
            per = arr.size // sc.P
            j = int(rng.choice(visible)) * per + int(rng.integers(0, per))
        else:
            arr = getattr(mlp, field).data.view(-1)
            j = int(rng.integers(0, arr.numel()))
        x0 = float(arr[j])
        # means also feed sin/cos(2^9 pi mu) of the encoding: a small step
        # keeps the central difference's truncation error below 1e-6
        h = 2e-7 if field == "means" else 1e-5 * max(1.0, abs(x0))
        vals = []
        for sgn in (1, -1):
            arr[j] = x0 + sgn * h
            dec, loss, _ = _forward(sc, mlp, target)
            vals.append((dec, loss))
        arr[j] = x0
        if vals[0][0] != base_dec or vals[1][0] != base_dec:
            continue
        fd = (vals[0][1] - vals[1][1]) / (2 * h)
        a = float(an[field].reshape(-1)[j])
        assert abs(fd - a) <= 1e-4 * max(abs(fd), abs(a)) + 1e-9, (field, j, fd, a)
        checked += 1
    assert checked >= 150, checked
