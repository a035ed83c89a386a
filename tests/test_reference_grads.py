"""CPU pins of the reference-mode geometry gradients.

The reference backward is not the adjoint of its forward: it recomputes the
2D covariance with the projection Jacobian's x row halved
(cuda_wireless_rasterizer/backward.cu:172-173), scales the conic chain by
det^2 / (det^2 + 1e-7) (:214), pairs dL/dJ of that halved Jacobian with the
derivatives of the un-halved one (:241-292), feeds the halved x row into the
direct mean gradient (:179-181, 416-430) and drops one scale-modifier factor
from dL/dscale (:353-355).  Each of those is characterised here by FP64
central differences of forward quantities, computed in this file from the
forward's definitions (projection auxiliary.h:101-113 with ndc2Pix
auxiliary.h:10-13, Jacobian forward.cu:80-97, cov2D forward.cu:99-121, cov3D
forward.cu:124-149), never from the backward's formulas:

  (i)   conic gradients 0: dL/dmean = Rv^T (dL/dpx grad(px)/2 + dL/dpy grad(py))
  (ii)  dL/dcov3D = dL(conic(cov2D(J_h, Sigma)))/dSigma * det_h^2/(det_h^2 + 1e-7),
        J_h = the Jacobian with its x row halved, held fixed
  (iii) screen-space gradients 0: dL/dmean = Rv^T (2 FD_x + FD_y) det_h^2/(det_h^2 + 1e-7),
        FD_x / FD_y = d/dt of the same loss with only J_h's x / y row following t
  (iv)  dL/dscale = d<dL/dcov3D, Sigma(s, q, mod)>/ds / mod, dL/drot = d<...>/dq

with L = dqa qa + dqb qb + dqc qc linear in the conic.  The oracle's FP64
twin (reference mode) must match each prediction, and must NOT match the
same prediction with one slip planted in it (the 1/2 on the other row, no
1/2, FD_x counted once, the det factor dropped, the modifier kept): each
check is shown to have the power to catch that slip.  The GPU's reference
mode is compared with the oracle in the GPU tests, so these pin it too."""
import numpy as np
import pytest

import gsparc_scenes as S
import oracle as O

W, H = 360, 90
EPS = float(np.float32(1e-7))
ONE_OVER_PI = float(np.float32(1.0 / np.pi))
TWO_OVER_PI = float(np.float32(2.0 / np.pi))
LOWPASS = float(np.float32(0.3))
IJ = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


# ---- forward definitions, FP64 ------------------------------------------------------------
def jacobian(t):
    """d(px, py)/dt as forward.cu:80-97 writes it (epsilons included)."""
    tx, ty, tz = t
    rxz2 = tx * tx + tz * tz
    rho = np.sqrt(rxz2)
    r2 = rxz2 + ty * ty
    wd = W * 0.5 * ONE_OVER_PI
    hd = H * ONE_OVER_PI
    return np.array([[wd * tz / (rxz2 + EPS), 0.0, -wd * tx / (rxz2 + EPS)],
                     [-hd * tx * ty / ((rho + EPS) * (r2 + EPS)), hd * rho / (r2 + EPS),
                      -hd * tz * ty / ((rho + EPS) * (r2 + EPS))]])


def pixel(t):
    """(px, py) of view-space t: longitude atan2(tx, tz), latitude
    asin(ty / (r + 1e-7)), then ndc2Pix (auxiliary.h:10-13, 101-113)."""
    r = np.linalg.norm(t)
    lon = np.arctan2(t[0], t[2])
    lat = np.arcsin(t[1] / (r + EPS))
    return np.array([((lon * ONE_OVER_PI + 1.0) * W - 1.0) * 0.5, ((lat * TWO_OVER_PI + 1.0) * H - 1.0) * 0.5])


def rotation(q):
    r, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y)],
                     [2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x)],
                     [2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y)]])


def sigma(s, q, mod):
    Rq = rotation(q)
    return Rq @ np.diag((mod * np.asarray(s)) ** 2) @ Rq.T


def cov2d(J, Rv, Sig):
    M = J @ Rv
    return M @ Sig @ M.T + LOWPASS * np.eye(2)


def conic_loss(C2, g):
    a, b, c = C2[0, 0], C2[0, 1], C2[1, 1]
    det = a * c - b * b
    return g[0] * (c / det) + g[1] * (-b / det) + g[2] * (a / det)


def det_factor(C2):
    det = C2[0, 0] * C2[1, 1] - C2[0, 1] ** 2
    return det * det / (det * det + EPS)


def central(f, x0, h):
    x0 = np.asarray(x0, np.float64)
    out = []
    for k in range(x0.size):
        xp, xm = x0.copy(), x0.copy()
        xp.flat[k] += h
        xm.flat[k] -= h
        out.append((f(xp) - f(xm)) / (2 * h))
    return np.array(out)


# ---- the scene and the oracle's reference-mode backward --------------------------------
def _view(rng):
    """Random rotation and translation, column-major as the reference stores it."""
    from scipy.spatial.transform import Rotation
    Rv = Rotation.random(random_state=int(rng.integers(1 << 30))).as_matrix()
    trans = rng.uniform(-0.5, 0.5, size=3)
    V = np.eye(4)
    V[:3, :3] = Rv
    V[:3, 3] = trans
    return V.T.reshape(16).copy(), Rv, trans


def _scene(seed, P=24, mod_small=True):
    rng = np.random.default_rng(seed)
    view16, Rv, trans = _view(rng)
    # view-space centres 1.5-4 m away, away from the poles (rho > 0.5 r)
    ts = []
    while len(ts) < P:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if np.hypot(d[0], d[2]) < 0.5:
            continue
        ts.append(d * rng.uniform(1.5, 4.0))
    ts = np.array(ts)
    means = (ts - trans) @ Rv          # t = Rv mu + trans
    half = P // 2
    scales = np.concatenate([np.exp(rng.uniform(np.log(0.004), np.log(0.02), size=(half, 3))),
                             np.exp(rng.uniform(np.log(0.05), np.log(0.25), size=(P - half, 3)))])
    q = rng.normal(size=(P, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    mlp = np.zeros(P * S.MLP_FLOATS)
    sc = S.Scene(means, scales, q, np.full((P, 1), 0.5), mlp, S.TX_POS.astype(np.float64),
                 np.zeros((1, 3)), view16.reshape(1, 4, 4))
    return sc, Rv, trans, rng


class _Pre:
    def __init__(self, P):
        self.radii = np.ones(P, np.int32)


def _oracle_grads(sc, g8, mode, mod):
    return O.preprocess_bwd(sc, _Pre(sc.P), g8, mode, r=0, dtype=np.float64, mod=mod)


def _close(a, b, tol):
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


# ---- (i) the direct term ---------------------------------------------------------------
def test_direct_mean_gradient_halves_the_x_row():
    sc, Rv, trans, rng = _scene(41)
    g8 = np.zeros((sc.P, 8))
    g8[:, 0] = rng.normal(size=sc.P) * 3.0      # dL/dpx (pixels)
    g8[:, 1] = rng.normal(size=sc.P) * 3.0      # dL/dpy
    got = _oracle_grads(sc, g8, O.GRAD_REFERENCE, 1.0).means
    exact = _oracle_grads(sc, g8 * np.array([0.5, 1, 1, 1, 1, 1, 1, 1]), O.GRAD_EXACT, 1.0).means
    for i in range(sc.P):
        t0 = Rv @ sc.means[i] + trans
        Jp = central(pixel, t0, 1e-6)          # [3, 2]: d(px, py)/dt_k
        predict = lambda fx, fy: Rv.T @ (fx * g8[i, 0] * Jp[:, 0] + fy * g8[i, 1] * Jp[:, 1])
        assert _close(got[i], predict(0.5, 1.0), 1e-6), i
        assert not _close(got[i], predict(1.0, 1.0), 1e-3), i       # no 1/2
        assert not _close(got[i], predict(1.0, 0.5), 1e-3), i       # 1/2 on the y row
    # the same statement through the oracle's exact mode: reference(dpx, dpy) = exact(dpx / 2, dpy)
    assert _close(got, exact, 1e-6)


# ---- (ii) dL/dcov3D ----------------------------------------------------------------------
def _conic_g8(sc, rng):
    g8 = np.zeros((sc.P, 8))
    g8[:, 2:5] = rng.normal(size=(sc.P, 3))     # dL/dqa, dL/dqb (counted once), dL/dqc
    return g8


def _halved(J, row=0):
    Jh = J.copy()
    Jh[row] *= 0.5
    return Jh


def test_cov3d_gradient_is_fd_of_the_halved_jacobian_cov2d_with_det_factor():
    sc, Rv, trans, rng = _scene(42)
    g8 = _conic_g8(sc, rng)
    got = _oracle_grads(sc, g8, O.GRAD_REFERENCE, 1.0).cov3
    det_effect = 0
    for i in range(sc.P):
        t0 = Rv @ sc.means[i] + trans
        Sig0 = sigma(sc.scales[i], sc.rotations[i], 1.0)
        J = jacobian(t0)

        def predict(Jfix, use_det=True):
            def f(u):
                Sg = np.zeros((3, 3))
                for e, (a, b) in enumerate(IJ):
                    Sg[a, b] = Sg[b, a] = u[e]
                return conic_loss(cov2d(Jfix, Rv, Sg), g8[i, 2:5])
            u0 = np.array([Sig0[a, b] for a, b in IJ])
            fd = central(f, u0, 1e-7 * max(1.0, np.abs(u0).max()))
            return fd * (det_factor(cov2d(Jfix, Rv, Sig0)) if use_det else 1.0)

        good = predict(_halved(J))
        assert _close(got[i], good, 1e-7), i
        assert not _close(got[i], predict(J), 1e-3), i                   # no 1/2
        assert not _close(got[i], predict(_halved(J, 1)), 1e-3), i       # 1/2 on the y row
        if not _close(good, predict(_halved(J), use_det=False), 2e-6):
            det_effect += 1
            assert not _close(got[i], predict(_halved(J), use_det=False), 1e-7), i   # det factor dropped
    assert det_effect >= sc.P // 3        # the small Gaussians make the 1e-7 visible


# ---- (iii) the mean gradient through the covariance -------------------------------------
def test_mean_gradient_through_cov_pairs_halved_dJ_with_unhalved_derivative():
    sc, Rv, trans, rng = _scene(43)
    g8 = _conic_g8(sc, rng)
    got = _oracle_grads(sc, g8, O.GRAD_REFERENCE, 1.0).means
    det_effect = 0
    for i in range(sc.P):
        t0 = Rv @ sc.means[i] + trans
        Sig = sigma(sc.scales[i], sc.rotations[i], 1.0)
        J0h = _halved(jacobian(t0))

        def fd_row(row, hrow=0):
            # only row `row` of the halved-on-`hrow` Jacobian follows t
            def f(t):
                Jt = _halved(jacobian(t), hrow)
                Jm = _halved(jacobian(t0), hrow)
                Jm[row] = Jt[row]
                return conic_loss(cov2d(Jm, Rv, Sig), g8[i, 2:5])
            return central(f, t0, 1e-6)

        fx, fy = fd_row(0), fd_row(1)
        dfac = det_factor(cov2d(J0h, Rv, Sig))

        def predict(cx, cy, use_det=True, fx=fx, fy=fy):
            return Rv.T @ ((cx * fx + cy * fy) * (dfac if use_det else 1.0))

        good = predict(2.0, 1.0)
        assert _close(got[i], good, 2e-6), i
        assert not _close(got[i], predict(1.0, 1.0), 1e-3), i       # the true derivative of L(J_h)
        assert not _close(got[i], predict(2.0, 2.0), 1e-3), i       # the pairing applied to the y row as well
        gx, gy = fd_row(0, 1), fd_row(1, 1)
        assert not _close(got[i], predict(1.0, 2.0, fx=gx, fy=gy), 1e-3), i   # 1/2 on the y row
        if not _close(good, predict(2.0, 1.0, use_det=False), 1e-5):
            det_effect += 1
            assert not _close(got[i], predict(2.0, 1.0, use_det=False), 2e-6), i
    assert det_effect >= 1


# ---- (iv) scale and rotation -------------------------------------------------------------
@pytest.mark.parametrize("mod", [1.0, 1.7])
def test_scale_and_rotation_gradients_from_the_cov3d_gradient(mod):
    sc, Rv, trans, rng = _scene(44)
    g8 = _conic_g8(sc, rng)
    gr = _oracle_grads(sc, g8, O.GRAD_REFERENCE, mod)
    ge = _oracle_grads(sc, g8, O.GRAD_EXACT, mod)
    for i in range(sc.P):
        dc = gr.cov3[i]
        lin = lambda s, q: sum(dc[e] * sigma(s, q, mod)[a, b] for e, (a, b) in enumerate(IJ))
        fd_s = central(lambda s: lin(s, sc.rotations[i]), sc.scales[i], 1e-7)
        fd_q = central(lambda q: lin(sc.scales[i], q), sc.rotations[i], 1e-7)
        assert _close(gr.scales[i], fd_s / mod, 1e-6), i
        assert _close(gr.rotations[i], fd_q, 1e-6), i
        if mod != 1.0:
            assert not _close(gr.scales[i], fd_s, 1e-3), i          # the modifier kept
        # the exact mode keeps it, on its own cov3D gradient
        dce = ge.cov3[i]
        line = lambda s: sum(dce[e] * sigma(s, sc.rotations[i], mod)[a, b] for e, (a, b) in enumerate(IJ))
        assert _close(ge.scales[i], central(line, sc.scales[i], 1e-7), 1e-6), i
