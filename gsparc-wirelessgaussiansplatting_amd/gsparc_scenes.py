"""Seeded synthetic scenes for the equirectangular RF rasterizer.

All tensors are generated on the CPU with numpy's PCG64 so that the GPU path,
the CPU checker and the CPU baseline see identical inputs.  Conventions follow
the reference:

* Room box [0,6] x [0,6] x [0,3] m; fixed transmitter at (0, 0.5, 3.0)
  (preprocess_data.py:154).
* Activated parameters as the renderer passes them (renderer.py:44-48,
  gaussian.py:33-38): scales = exp(N(-2.5, 0.3)), rotations = normalised
  N(0, I4) quaternions (w, x, y, z), opacities = sigmoid(N(0, 1)).
* Per-Gaussian emission MLP 3->32->2 packed as
  [W1 P*32*3 | b1 P*32 | W2 P*2*32 | b2 P*2] (gaussian.py:554-572).
* View matrix: receiver orientation = identity, translation = -rx, stored
  column-major (the transpose of the world->receiver matrix), as built by
  preprocess_data.py:177-202.
"""
from __future__ import annotations

import dataclasses

import numpy as np

ROOM = np.array([6.0, 6.0, 3.0])
TX_POS = np.array([0.0, 0.5, 3.0], dtype=np.float32)
MLP_FLOATS = 194  # 32*3 + 32 + 2*32 + 2


@dataclasses.dataclass
class Scene:
    means: np.ndarray       # [P,3] f32
    scales: np.ndarray      # [P,3] f32 (activated)
    rotations: np.ndarray   # [P,4] f32 (normalised, w x y z)
    opacities: np.ndarray   # [P,1] f32 (activated)
    mlp: np.ndarray         # [P*194] f32 packed emission MLPs
    tx_pos: np.ndarray      # [3] f32
    rx_pos: np.ndarray      # [R,3] f32
    views: np.ndarray       # [R,4,4] f32 column-major view matrices
    H: int = 90
    W: int = 360

    @property
    def P(self) -> int:
        return int(self.means.shape[0])

    @property
    def R(self) -> int:
        return int(self.rx_pos.shape[0])


def view_matrix(rx: np.ndarray) -> np.ndarray:
    """World->receiver transform with identity orientation, column-major."""
    v = np.eye(4, dtype=np.float32)
    v[:3, 3] = -np.asarray(rx, dtype=np.float32)
    return np.ascontiguousarray(v.T)


def _sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def _common(rng, P):
    scales = np.exp(rng.normal(-2.5, 0.3, size=(P, 3))).astype(np.float32)
    q = rng.normal(size=(P, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    opac = _sigmoid(rng.normal(size=(P, 1))).astype(np.float32)
    return scales, q.astype(np.float32), opac


def constant_emission_mlp(means: np.ndarray, emission: np.ndarray, tx=TX_POS) -> np.ndarray:
    """MLP weights whose output equals `emission` for every receiver.

    With W1 = b1 = W2 = 0 the reference MLP returns sigmoid(b2 - log(|tx-mu| + 1e-6))
    (forward.cu:46-72), so b2 = logit(e) + log(d + 1e-6) reproduces e up to FP32
    rounding.  emission: [P] (same value on both channels) or [P,2].
    """
    P = means.shape[0]
    e = np.asarray(emission, dtype=np.float64)
    if e.ndim == 1:
        e = np.stack([e, e], axis=1)
    d = np.linalg.norm(np.asarray(tx, np.float32)[None, :].astype(np.float32) - means, axis=1).astype(np.float64)
    b2 = np.log(e / (1.0 - e)) + np.log(d + 1e-6)[:, None]
    mlp = np.zeros(P * MLP_FLOATS, dtype=np.float32)
    mlp[P * 192:] = b2.astype(np.float32).reshape(-1)
    return mlp


def random_mlp(rng, P: int, scale: float = 0.3) -> np.ndarray:
    """Random per-Gaussian MLP weights (the reference initialises with randn,
    gaussian.py:559-563; a smaller scale keeps the sigmoid out of saturation)."""
    return (rng.normal(size=P * MLP_FLOATS) * scale).astype(np.float32)


def uniform_means(rng, P):
    return (rng.uniform(size=(P, 3)) * ROOM).astype(np.float32)


def wall_means(rng, P, frac_wall=0.7, band=0.3):
    """70% within `band` m of the six walls (wall chosen by area), 30% uniform."""
    n_wall = int(round(P * frac_wall))
    pts = rng.uniform(size=(P, 3)) * ROOM
    areas = np.array([ROOM[1] * ROOM[2], ROOM[1] * ROOM[2], ROOM[0] * ROOM[2], ROOM[0] * ROOM[2],
                      ROOM[0] * ROOM[1], ROOM[0] * ROOM[1]])
    wall = rng.choice(6, size=n_wall, p=areas / areas.sum())
    off = rng.uniform(0.0, band, size=n_wall)
    axis = wall // 2
    low = (wall % 2) == 0
    w = pts[:n_wall]
    w[np.arange(n_wall), axis] = np.where(low, off, ROOM[axis] - off)
    return pts.astype(np.float32)


def config0(seed: int = 0, P: int = 2000) -> Scene:
    """Plain forward render: P uniform Gaussians, one receiver at the room centre,
    scalar emission sigmoid(N(0,1)) through the closed-form MLP weights."""
    rng = np.random.Generator(np.random.PCG64(seed))
    means = uniform_means(rng, P)
    scales, rots, opac = _common(rng, P)
    emis = _sigmoid(rng.normal(size=P))
    mlp = constant_emission_mlp(means, emis)
    rx = np.array([[3.0, 3.0, 1.5]], dtype=np.float32)
    return Scene(means, scales, rots, opac, mlp, TX_POS.copy(), rx, view_matrix(rx[0])[None])


def config1(seed: int = 1, P: int = 20000) -> Scene:
    """Single receiver, P Gaussians drawn as in config 0, random MLP weights."""
    rng = np.random.Generator(np.random.PCG64(seed))
    means = uniform_means(rng, P)
    scales, rots, opac = _common(rng, P)
    mlp = random_mlp(rng, P)
    rx = np.array([[3.0, 3.0, 1.5]], dtype=np.float32)
    return Scene(means, scales, rots, opac, mlp, TX_POS.copy(), rx, view_matrix(rx[0])[None])


def config2(seed: int = 2, P: int = 50000, R: int = 64) -> Scene:
    """Batched receivers: 70% of the Gaussians near the walls, R receivers
    uniform in the room (kept 0.25 m inside the walls)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    means = wall_means(rng, P)
    scales, rots, opac = _common(rng, P)
    mlp = random_mlp(rng, P)
    rx = (0.25 + rng.uniform(size=(R, 3)) * (ROOM - 0.5)).astype(np.float32)
    views = np.stack([view_matrix(r) for r in rx])
    return Scene(means, scales, rots, opac, mlp, TX_POS.copy(), rx, views)


def small_scene(seed: int, P: int, R: int = 1, H: int = 90, W: int = 360, walls: bool = False) -> Scene:
    """Small seeded scene for parity tests."""
    rng = np.random.Generator(np.random.PCG64(seed))
    means = wall_means(rng, P) if walls else uniform_means(rng, P)
    scales, rots, opac = _common(rng, P)
    mlp = random_mlp(rng, P)
    rx = (0.5 + rng.uniform(size=(R, 3)) * (ROOM - 1.0)).astype(np.float32)
    views = np.stack([view_matrix(r) for r in rx])
    return Scene(means, scales, rots, opac, mlp, TX_POS.copy(), rx, views, H=H, W=W)


def target_spectra(seed: int, shape) -> np.ndarray:
    """A smooth positive synthetic target map (stand-in for a measured spectrum)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(0.0, 0.5, size=shape).astype(np.float32)


def config4(seed: int = 4, n_scenes: int = 16, p_min: int = 5000, p_max: int = 200000, r_min: int = 4,
            r_max: int = 64, variant: int = 0) -> list:
    """Uneven scenes: n_scenes independent rooms with Gaussian counts
    log-uniform in [p_min, p_max] and r_min..r_max receivers (uniform integer),
    each drawn as in config2.  `variant` changes only the Gaussians (sizes,
    receivers and transmitter stay), e.g. for the target spectra."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Ps = np.rint(np.exp(rng.uniform(np.log(p_min), np.log(p_max), size=n_scenes))).astype(int)
    Rs = rng.integers(r_min, r_max + 1, size=n_scenes)
    out = []
    for s in range(n_scenes):
        rr = np.random.Generator(np.random.PCG64([seed, s]))
        rx = (0.25 + rr.uniform(size=(int(Rs[s]), 3)) * (ROOM - 0.5)).astype(np.float32)
        g = np.random.Generator(np.random.PCG64([seed, s, variant + 1]))
        P = int(Ps[s])
        means = wall_means(g, P)
        scales, rots, opac = _common(g, P)
        mlp = random_mlp(g, P)
        views = np.stack([view_matrix(r) for r in rx])
        out.append(Scene(means, scales, rots, opac, mlp, TX_POS.copy(), rx, views))
    return out
