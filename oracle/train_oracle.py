"""CPU restatement of the reference's training surround, for the tests.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline may import this module, as the checker; nothing in the product
path calls it.

Arrays follow the reference's tensors (gaussian.py:31-555): centre [P,3],
scaling [P,3] (log), rotation [P,4] (raw), opacity [P,1] (logit), the four
emission MLP tensors, the Adam moments of each per-Gaussian group, and the
statistics centre_gradient_accum [P,1], denom [P,1], max_radii2D [P].  Torch
on the CPU in float32, as the reference computes on the GPU.

Functions:
  densify_and_prune(m, ...)   gaussian.py:510-535 with 450-508, 395-429,
                              308-337 and the optimizer-state surgery of
                              266-393; split children from injected normals
  prune_post(m, ...)          train.py:539-555
  reset_opacity(m)            gaussian.py:223-253
  add_stats(m, g2d, radii)    gaussian.py:538-551, train.py:510-516
The Adam step itself is torch.optim.Adam (the reference's optimizer) in the
tests; nothing is restated for it here.

Pins (tests/test_train_oracle.py): hand-built sets whose result is fixed by
the reference's rules (which Gaussians clone, split, survive, and in what
order), closed forms of a split child (offset R(q)(s z), |offset| = |s z|,
scale s / 1.2) and a brute-force loop restatement of the order on random
sets.
"""
from __future__ import annotations

import math

import numpy as np
import torch

GROUPS = ("centre", "scaling", "rotation", "opacity", "fc1_weights", "fc1_bias", "fc2_weights", "fc2_bias")
WIDTH = {"centre": (3,), "scaling": (3,), "rotation": (4,), "opacity": (1,), "fc1_weights": (32, 3),
         "fc1_bias": (32,), "fc2_weights": (2, 32), "fc2_bias": (2,)}
SECTIONS = (("fc1_weights", 96), ("fc1_bias", 32), ("fc2_weights", 64), ("fc2_bias", 2))


class Model:
    """The per-Gaussian tensors of a GaussianModel and their Adam moments."""

    def __init__(self, P):
        self.p = {g: torch.zeros((P,) + WIDTH[g]) for g in GROUPS}
        self.m = {g: torch.zeros((P,) + WIDTH[g]) for g in GROUPS}
        self.v = {g: torch.zeros((P,) + WIDTH[g]) for g in GROUPS}
        self.accum = torch.zeros((P, 1))
        self.denom = torch.zeros((P, 1))
        self.max_radii2D = torch.zeros(P)

    @property
    def P(self):
        return self.p["centre"].shape[0]

    def scaling(self):
        return torch.exp(self.p["scaling"])

    def opacity(self):
        return torch.sigmoid(self.p["opacity"])

    def mlp_flat(self):
        """EmissionMLPs.get_all_params (gaussian.py:565-572)."""
        return torch.cat([self.p[g].reshape(-1) for g, _ in SECTIONS])

    def set_mlp_flat(self, flat, which="p"):
        d = getattr(self, which)
        P = self.P
        o = 0
        for g, w in SECTIONS:
            d[g] = flat[o: o + w * P].reshape((P,) + WIDTH[g]).clone()
            o += w * P

    # ---- the optimizer-state surgery
    def _cat(self, new: dict):
        """cat_tensors_to_optimizer (gaussian.py:339-393): append, moments zero."""
        for g in GROUPS:
            self.p[g] = torch.cat([self.p[g], new[g]])
            self.m[g] = torch.cat([self.m[g], torch.zeros_like(new[g])])
            self.v[g] = torch.cat([self.v[g], torch.zeros_like(new[g])])

    def _postfix(self, new: dict):
        """densification_postfix (gaussian.py:395-429)."""
        self._cat(new)
        P = self.P
        self.accum = torch.zeros((P, 1))
        self.denom = torch.zeros((P, 1))
        self.max_radii2D = torch.zeros(P)

    def prune_points(self, mask):
        """prune_points (gaussian.py:308-337): drop where mask is True."""
        keep = ~mask
        for g in GROUPS:
            self.p[g] = self.p[g][keep]
            self.m[g] = self.m[g][keep]
            self.v[g] = self.v[g][keep]
        self.accum = self.accum[keep]
        self.denom = self.denom[keep]
        self.max_radii2D = self.max_radii2D[keep]


def rotation_matrices(q):
    """build_rotation (utils.py:127-148): rotation of the normalised quaternion."""
    n = torch.sqrt(q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2] + q[:, 3] * q[:, 3])
    u = q / n[:, None]
    r, x, y, z = u[:, 0], u[:, 1], u[:, 2], u[:, 3]
    R = torch.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - r * z)
    R[:, 0, 2] = 2 * (x * z + r * y)
    R[:, 1, 0] = 2 * (x * y + r * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - r * x)
    R[:, 2, 0] = 2 * (x * z - r * y)
    R[:, 2, 1] = 2 * (y * z + r * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def densify_and_prune(m: Model, max_grad, min_opacity, extent, max_screen_size, scene_min=None, scene_max=None,
                      noise=None, percent_dense=0.01):
    """densify_and_prune in place; noise: standard normals [2 n_split, 3] (or
    more rows), row k the first child of the k-th split Gaussian, row
    n_split + k the second.  Returns (n_clone, n_split)."""
    grads = m.accum / m.denom
    grads[grads.isnan()] = 0.0
    thr = np.float32(max_grad)
    # clone (gaussian.py:493-508): |grad| >= thr, max scale <= percent_dense * extent
    gsel = torch.norm(grads, dim=-1) >= float(thr)
    small = torch.max(m.scaling(), dim=1).values <= float(np.float32(percent_dense * extent))
    clone = gsel & small
    n_clone = int(clone.sum())
    m._postfix({g: m.p[g][clone] for g in GROUPS})
    # split (gaussian.py:450-491) over the set with the clones appended
    n0 = m.P
    padded = torch.zeros(n0)
    padded[: grads.shape[0]] = grads.squeeze(-1)
    sel = (padded >= float(thr)) & (torch.max(m.scaling(), dim=1).values > float(np.float32(percent_dense * extent)))
    n_split = int(sel.sum())
    N = 2
    s = m.scaling()[sel].repeat(N, 1)
    z = torch.as_tensor(np.asarray(noise, np.float32)[: N * n_split].reshape(-1, 3)) if n_split else torch.zeros((0, 3))
    samples = s * z
    R = rotation_matrices(m.p["rotation"][sel]).repeat(N, 1, 1)
    new = {
        "centre": torch.bmm(R, samples.unsqueeze(-1)).squeeze(-1) + m.p["centre"][sel].repeat(N, 1),
        "scaling": torch.log(m.scaling()[sel].repeat(N, 1) / (0.6 * N)),
        "rotation": m.p["rotation"][sel].repeat(N, 1),
        "opacity": m.p["opacity"][sel].repeat(N, 1),
    }
    for g, _ in SECTIONS:
        new[g] = m.p[g][sel].repeat((N,) + (1,) * len(WIDTH[g]))
    m._postfix(new)
    m.prune_points(torch.cat([sel, torch.zeros(N * n_split, dtype=torch.bool)]))
    # prune (gaussian.py:519-530)
    mask = (m.opacity() < float(np.float32(min_opacity))).squeeze(-1)
    if max_screen_size:
        big_vs = m.max_radii2D > max_screen_size
        big_ws = m.scaling().max(dim=1).values > float(np.float32(0.1 * extent))
        mask = mask | big_vs | big_ws
    if scene_min is not None:
        lo = torch.as_tensor(np.asarray(scene_min, np.float32))
        hi = torch.as_tensor(np.asarray(scene_max, np.float32))
        c = m.p["centre"]
        mask = mask | (c < lo).any(dim=1) | (c > hi).any(dim=1)
    m.prune_points(mask)
    return n_clone, n_split


def prune_post(m: Model, min_opacity, extent, scene_min, scene_max):
    """train.py:539-555; returns the number removed."""
    mask = (m.opacity() < float(np.float32(min_opacity))).squeeze(-1)
    mask = mask | (m.scaling().max(dim=1).values > float(np.float32(0.1 * extent)))
    lo = torch.as_tensor(np.asarray(scene_min, np.float32))
    hi = torch.as_tensor(np.asarray(scene_max, np.float32))
    c = m.p["centre"]
    mask = mask | (c < lo).any(dim=1) | (c > hi).any(dim=1)
    n = int(mask.sum())
    if n:
        m.prune_points(mask)
    return n


def reset_opacity(m: Model):
    """gaussian.py:223-253."""
    o = torch.min(m.opacity(), torch.ones_like(m.opacity()) * 0.01)
    z = torch.log(o / (1 - o))
    m.p["opacity"] = torch.nan_to_num(z, nan=0.0, posinf=1.0, neginf=0.0)
    m.m["opacity"] = torch.zeros_like(z)
    m.v["opacity"] = torch.zeros_like(z)


def add_stats(m: Model, grad2d, radii, grad_scale=1.0):
    """Per receiver r: visible = radii[r] > 0; accum[visible] += |g[r, visible, :2]|
    * grad_scale; denom[visible] += 1; max_radii2D = max(., radii)."""
    g = torch.as_tensor(np.asarray(grad2d, np.float32))
    rad = torch.as_tensor(np.asarray(radii, np.int32))
    for r in range(rad.shape[0]):
        vis = rad[r] > 0
        m.max_radii2D[vis] = torch.max(m.max_radii2D[vis], rad[r][vis].float())
        m.accum[vis] += torch.norm(g[r][vis, :2], dim=-1, keepdim=True) * grad_scale
        m.denom[vis] += 1


def expected_order(flags, keep_mask_fn):
    """Brute-force loop restatement of the candidate order: originals not
    split, clones, first children, second children (each in index order);
    then the kept ones.  flags: 0 / 1 clone / 2 split per original.
    keep_mask_fn(kind, src, child) -> bool.  Returns [(kind, src, child)]."""
    cand = []
    for i, f in enumerate(flags):
        if f != 2:
            cand.append((0, i, 0))
    for i, f in enumerate(flags):
        if f == 1:
            cand.append((1, i, 0))
    split = [i for i, f in enumerate(flags) if f == 2]
    for copy in range(2):
        for k, i in enumerate(split):
            cand.append((2, i, copy * len(split) + k))
    return [c for c in cand if keep_mask_fn(*c)]


def from_arrays(centre, log_scale, rotation, logit, mlp_flat, accum=None, denom=None, max_radius=None):
    P = int(np.asarray(centre).shape[0])
    m = Model(P)
    m.p["centre"] = torch.as_tensor(np.asarray(centre, np.float32)).reshape(P, 3).clone()
    m.p["scaling"] = torch.as_tensor(np.asarray(log_scale, np.float32)).reshape(P, 3).clone()
    m.p["rotation"] = torch.as_tensor(np.asarray(rotation, np.float32)).reshape(P, 4).clone()
    m.p["opacity"] = torch.as_tensor(np.asarray(logit, np.float32)).reshape(P, 1).clone()
    m.set_mlp_flat(torch.as_tensor(np.asarray(mlp_flat, np.float32)))
    if accum is not None:
        m.accum = torch.as_tensor(np.asarray(accum, np.float32)).reshape(P, 1).clone()
        m.denom = torch.as_tensor(np.asarray(denom, np.float32)).reshape(P, 1).clone()
        m.max_radii2D = torch.as_tensor(np.asarray(max_radius, np.float32)).reshape(P).clone()
    return m


def state_flat(m: Model):
    """Adam moments in the library's layout: [2][centre | scaling | rotation |
    logit] and [2][W1 | b1 | W2 | b2]."""
    geo = [m.m[g].reshape(-1) for g in ("centre", "scaling", "rotation", "opacity")] + \
          [m.v[g].reshape(-1) for g in ("centre", "scaling", "rotation", "opacity")]
    mlp = [m.m[g].reshape(-1) for g, _ in SECTIONS] + [m.v[g].reshape(-1) for g, _ in SECTIONS]
    return torch.cat(geo).numpy(), torch.cat(mlp).numpy()


def set_state_flat(m: Model, geo, mlp):
    P = m.P
    geo = torch.as_tensor(np.asarray(geo, np.float32))
    mlp = torch.as_tensor(np.asarray(mlp, np.float32))
    o = 0
    for which in ("m", "v"):
        d = getattr(m, which)
        for g, w in (("centre", 3), ("scaling", 3), ("rotation", 4), ("opacity", 1)):
            d[g] = geo[o: o + w * P].reshape((P,) + WIDTH[g]).clone()
            o += w * P
    o = 0
    for which in ("m", "v"):
        d = getattr(m, which)
        for g, w in SECTIONS:
            d[g] = mlp[o: o + w * P].reshape((P,) + WIDTH[g]).clone()
            o += w * P


def unit_quaternion_rotate(q, v):
    """|R(q) v| = |v| and R(q) = the quaternion sandwich q v q*, an independent
    closed form to pin rotation_matrices (Hamilton product, w first)."""
    q = np.asarray(q, np.float64)
    q = q / np.linalg.norm(q)
    w, x, y, z = q
    vx, vy, vz = v
    # t = 2 q_vec x v ; v' = v + w t + q_vec x t
    tx = 2 * (y * vz - z * vy)
    ty = 2 * (z * vx - x * vz)
    tz = 2 * (x * vy - y * vx)
    return np.array([vx + w * tx + (y * tz - z * ty), vy + w * ty + (z * tx - x * tz), vz + w * tz + (x * ty - y * tx)])


def lr_schedule(step, lr_init, lr_final, max_steps, delay_steps=0, delay_mult=1.0):
    """get_expon_lr_func (utils.py:46-79)."""
    if step < 0 or (lr_init == 0.0 and lr_final == 0.0):
        return 0.0
    delay = 1.0
    if delay_steps > 0:
        delay = delay_mult + (1 - delay_mult) * np.sin(0.5 * np.pi * np.clip(step / delay_steps, 0, 1))
    t = np.clip(step / max_steps, 0, 1)
    return float(delay * np.exp(np.log(lr_init) * (1 - t) + np.log(lr_final) * t))


def lr_closed_form(step, lr_init, lr_final, max_steps):
    """lr_init (lr_final / lr_init)^(step / max_steps): the geometric
    interpolation the log-linear form equals (no delay)."""
    t = min(max(step / max_steps, 0.0), 1.0)
    return lr_init * math.pow(lr_final / lr_init, t)
