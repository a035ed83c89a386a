"""The Gaussian set of a training run and the optimizer surround of the
rasterizer, on the library's kernels (k_train.hip through include/gsparc.h).

GaussianModel mirrors what gaussian.py:GaussianModel offers a training loop:
get_centre / get_scaling / get_rotation / get_opacity / get_emission_mlps,
initialize_gaussians, training_setup, update_learning_rate, reset_opacity,
add_densification_stats, densify_and_prune and optimizer_step (the
reference's optimizer.step() with the NaN guard and clamp around it,
train.py:441-444, 558-566).  The parameters, their activations, the Adam
moments and the statistics live in device arrays allocated for a capacity;
densify and prune gather one array set into a second one (two banks used in
turn), so P changes without a per-tensor reallocation.

The optimizer state follows torch.optim.Adam(eps=1e-15) of training_setup
(gaussian.py:171-212) exactly in what it keeps: exp_avg and exp_avg_sq per
parameter, a step count per group (the reference's per-parameter
state["step"]; its four emission MLP groups and its four confidence groups
step together), and the groups that the reference has just replaced (after
densify / prune, after reset_opacity for the opacity) do not step in that
iteration because torch skips a parameter without gradient.  The
'clip_grad_norm' entry training_setup adds to every param group
(gaussian.py:199-201) is read by nothing in torch.optim.Adam, so no clipping
happens there and none here.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

import gsparc_hip as gh

MLP_FLOATS = 194                     # 3 -> 32 -> 2 per Gaussian: W1 96 | b1 32 | W2 64 | b2 2
PERCENT_DENSE = 0.01                 # gaussian.py:54
SPLIT_N = 2                          # densify_and_split's N (gaussian.py:450)

# (name, floats per Gaussian) of one bank, in gsparc_gaussians order
_FIELDS = (("centre", 3), ("log_scale", 3), ("rotation", 4), ("logit", 1), ("emission_mlps", MLP_FLOATS),
           ("scale", 3), ("rotation_n", 4), ("opacity", 1), ("state", 22), ("state_mlp", 2 * MLP_FLOATS),
           ("accum", 1), ("denom", 1), ("max_radius", 1))


class OptimizationParams:
    """The reference's OptimizationParams defaults (optimizer.py:4-45); keyword
    arguments override them, unknown keys are ignored as there."""

    def __init__(self, **kwargs):
        self.iterations = 200_000
        self.lambda1 = 0.2
        self.position_lr_init = 0.00016
        self.position_lr_final = 0.00000016
        self.position_lr_delay_mult = 0.01
        self.position_lr_max_steps = 200000
        self.feature_lr = 0.0025
        self.opacity_lr = 0.0055
        self.scaling_lr = 0.005
        self.rotation_lr = 0.001
        self.emission_mlp_lr = 0.002
        self.percent_dense = 0.01
        self.lambda_dssim = 0.2
        self.densification_interval = 300
        self.opacity_reset_interval = 4000
        self.densify_from_iter = 1000
        self.densify_until_iter = 120000
        self.densify_grad_threshold = 0.0002
        self.cube_size = 0.3
        self.confidence_lr = 0.001
        self.confidence_alpha = 0.3
        self.confidence_hidden_size = 32
        for k, v in kwargs.items():
            if hasattr(self, k):
                setattr(self, k, v)


def inverse_sigmoid(x: torch.Tensor) -> torch.Tensor:
    """utils.py inverse_sigmoid: log(x / (1 - x))."""
    return torch.log(x / (1 - x))


def expon_lr(step, lr_init, lr_final, lr_delay_steps=0, lr_delay_mult=1.0, max_steps=1000000):
    """get_expon_lr_func(...)(step) (utils.py:46-79), for the host side."""
    if step < 0 or (lr_init == 0.0 and lr_final == 0.0):
        return 0.0
    if lr_delay_steps > 0:
        delay = lr_delay_mult + (1 - lr_delay_mult) * math.sin(0.5 * math.pi * min(max(step / lr_delay_steps, 0), 1))
    else:
        delay = 1.0
    t = min(max(step / max_steps, 0), 1)
    return delay * math.exp(math.log(lr_init) * (1 - t) + math.log(lr_final) * t)


class _Bank:
    """One set of device arrays for `capacity` Gaussians."""

    def __init__(self, capacity: int, device):
        self.capacity = int(capacity)
        self.t = {n: torch.zeros(max(1, w * self.capacity), dtype=torch.float32, device=device) for n, w in _FIELDS}

    def struct(self) -> gh.GaussiansC:
        return gh.GaussiansC(*[self.t[n].data_ptr() for n, _ in _FIELDS])


class GaussianModel:
    """Device-resident Gaussians with their optimizer state (gaussian.py:31-555).

    Views (valid until the next densify / prune): get_centre [P,3], the raw
    _scaling [P,3] / _rotation [P,4] / _opacity [P,1], the activations
    get_scaling, get_rotation, get_opacity and get_emission_mlps [P*194]
    (sections W1 [P,32,3] | b1 [P,32] | W2 [P,2,32] | b2 [P,2], the layout of
    EmissionMLPs.get_all_params, gaussian.py:565-572).  `headroom` sizes a
    bank for that many times the Gaussians it must hold when it grows."""

    def __init__(self, device="cuda", headroom: float = 1.5):
        self.device = torch.device(device)
        self.headroom = float(headroom)
        self.P = 0
        self.banks = [None, None]
        self.cur = 0
        self.percent_dense = PERCENT_DENSE
        self.spatial_lr_scale = 1.0
        self.adam = None
        self.train_state = None
        self.conf = None           # packed confidence network [5 hidden + 1] (gsparc_train.ConfidenceMLP.packed)
        self.conf_hidden = 0
        self.conf_grad = None
        self.conf_state = None

    # ---- the arrays
    @property
    def bank(self) -> _Bank:
        return self.banks[self.cur]

    def _view(self, name, width):
        t = self.bank.t[name][: width * self.P]
        return t.view(self.P, width) if width in (3, 4) else t

    @property
    def get_num_gaussians(self) -> int:
        return self.P

    @property
    def get_centre(self):
        return self._view("centre", 3)

    @property
    def _scaling(self):
        return self._view("log_scale", 3)

    @property
    def _rotation(self):
        return self._view("rotation", 4)

    @property
    def _opacity(self):
        return self.bank.t["logit"][: self.P].view(self.P, 1)

    @property
    def get_scaling(self):
        return self._view("scale", 3)

    @property
    def get_rotation(self):
        return self._view("rotation_n", 4)

    @property
    def get_opacity(self):
        return self.bank.t["opacity"][: self.P].view(self.P, 1)

    @property
    def get_emission_mlps(self):
        return self.bank.t["emission_mlps"][: MLP_FLOATS * self.P]

    def array(self, name: str) -> torch.Tensor:
        """Any array of the current bank in its P-entry layout (flat)."""
        w = dict(_FIELDS)[name]
        return self.bank.t[name][: w * self.P]

    @property
    def centre_gradient_accum(self):
        return self.bank.t["accum"][: self.P].view(self.P, 1)

    @property
    def denom(self):
        return self.bank.t["denom"][: self.P].view(self.P, 1)

    @property
    def max_radii2D(self):
        return self.bank.t["max_radius"][: self.P]

    def struct(self) -> gh.GaussiansC:
        return self.bank.struct()

    def _ensure(self, which: int, n: int):
        b = self.banks[which]
        if b is None or b.capacity < n:
            cap = max(n, int(math.ceil(n * self.headroom)), 1)
            self.banks[which] = _Bank(cap, self.device)

    # ---- set-up
    def load_raw(self, centre, log_scale, rotation, logit, emission_mlps):
        """Set the raw parameters (centre [P,3], log-scale [P,3], quaternion
        [P,4], logit [P] or [P,1], emission MLPs [P*194] sectioned); moments
        and statistics start at zero, the activations are computed."""
        P = int(centre.shape[0])
        self._ensure(self.cur, P)
        b = self.bank
        for t in b.t.values():
            t.zero_()
        self.P = P
        for name, src, w in (("centre", centre, 3), ("log_scale", log_scale, 3), ("rotation", rotation, 4),
                             ("logit", logit, 1), ("emission_mlps", emission_mlps, MLP_FLOATS)):
            src = torch.as_tensor(src, dtype=torch.float32, device=self.device).reshape(-1)
            if src.numel() != w * P:
                raise RuntimeError(f"{name} must hold {w} floats per Gaussian")
            b.t[name][: w * P].copy_(src)
        self.activate()

    def initialize_gaussians(self, points, generator: torch.Generator | None = None):
        """initialize_gaussians (gaussian.py:122-169): centres at the points,
        log-scale log(sqrt(max(d2, 1e-7))) with d2 the squared distance to the
        nearest other point, identity rotations, opacity 0.1, MLP weights
        standard normal (four draws in the order fc1_weights [P,32,3],
        fc1_bias [P,32], fc2_weights [P,2,32], fc2_bias [P,2])."""
        c = torch.as_tensor(points, dtype=torch.float32, device=self.device).reshape(-1, 3)
        P = c.shape[0]
        d2 = torch.full((P,), float("inf"), device=self.device)
        for s in range(0, P, 1024):
            e = min(s + 1024, P)
            dd = ((c[s:e, None, :] - c[None, :, :]) ** 2).sum(dim=2)
            dd[torch.arange(e - s, device=self.device), torch.arange(s, e, device=self.device)] = float("inf")
            d2[s:e] = dd.min(dim=1).values
        d2 = torch.clamp_min(d2, 1e-7)
        ls = torch.log(torch.sqrt(d2))[:, None].repeat(1, 3)
        rot = torch.zeros((P, 4), device=self.device)
        rot[:, 0] = 1
        logit = inverse_sigmoid(0.1 * torch.ones((P, 1), dtype=torch.float, device=self.device))
        g = generator
        mlp = torch.cat([torch.randn((P, 32, 3), device=self.device, generator=g).reshape(-1),
                         torch.randn((P, 32), device=self.device, generator=g).reshape(-1),
                         torch.randn((P, 2, 32), device=self.device, generator=g).reshape(-1),
                         torch.randn((P, 2), device=self.device, generator=g).reshape(-1)])
        self.load_raw(c, ls, rot, logit, mlp)

    def activate(self):
        gh.check(gh.lib().gsparc_gaussians_activate(self.P, C.byref(self.struct()), gh.stream_handle(self.device)),
                 "gsparc_gaussians_activate")

    def training_setup(self, optim_params: OptimizationParams, spatial_lr_scale=None, confidence=None,
                       scene_min=None, scene_max=None, first_iteration: int | None = 1):
        """Adam settings and groups of training_setup (gaussian.py:171-212).
        confidence: the packed confidence network (gsparc_train.ConfidenceMLP.
        packed()) or None.  scene_min / scene_max: the clamp bounds applied
        after every step (train.py:502-505, 561-562: xyz_min / xyz_max with a
        margin of 4), None for no clamp.  first_iteration: the iteration the
        first step completes (the reference's loop starts at 1); None keeps
        the step counts and moments a restored checkpoint brought, and a
        restored confidence network when `confidence` is None."""
        o = optim_params
        if spatial_lr_scale is not None:
            self.spatial_lr_scale = float(spatial_lr_scale)
        self.percent_dense = float(getattr(o, "percent_dense", PERCENT_DENSE))
        a = gh.AdamConfig()
        a.beta1, a.beta2, a.eps = 0.9, 0.999, 1e-15
        a.lr_centre_init = o.position_lr_init * self.spatial_lr_scale
        a.lr_centre_final = o.position_lr_final * self.spatial_lr_scale
        a.lr_centre_delay_mult = o.position_lr_delay_mult
        a.lr_centre_delay_steps = 0          # training_setup does not pass lr_delay_steps
        a.lr_centre_max_steps = int(o.position_lr_max_steps)
        a.lr_scaling = o.scaling_lr
        a.lr_rotation = o.rotation_lr
        a.lr_opacity = o.opacity_lr
        a.lr_emission = o.emission_mlp_lr
        a.lr_confidence = getattr(o, "confidence_lr", 0.001)
        if scene_min is not None:
            a.clamp_centres = 1
            for k in range(3):
                a.scene_min[k] = float(np.float32(scene_min[k]))
                a.scene_max[k] = float(np.float32(scene_max[k]))
        self.adam = a
        if first_iteration is not None or self.train_state is None:
            self.train_state = torch.zeros(gh.TRAIN_STATE_BYTES // 4, dtype=torch.int32, device=self.device)
            self.train_state[0] = int(first_iteration if first_iteration is not None else 1)
        if confidence is not None:
            self.conf = confidence
            self.conf_hidden = (confidence.numel() - 1) // 5
            self.conf_grad = torch.zeros_like(confidence)
            self.conf_state = torch.zeros(2 * confidence.numel(), dtype=torch.float32, device=self.device)

    # ---- per step
    def update_learning_rate(self, iteration: int):
        """Set the iteration the next step completes; the centre rate is
        evaluated on the device from it (gaussian.py:214-221).  Returns the
        centre rate, as the reference does."""
        self.train_state[0] = int(iteration)
        a = self.adam
        return expon_lr(iteration, a.lr_centre_init, a.lr_centre_final, 0, a.lr_centre_delay_mult,
                        a.lr_centre_max_steps)

    def counters(self) -> dict:
        """The device training state (synchronises): iteration, per-group step
        counts and the NaN-centre flag."""
        v = self.train_state[:8].cpu().tolist()
        names = ("centre", "scaling", "rotation", "opacity", "emission", "confidence")
        return {"iteration": v[0], "steps": dict(zip(names, v[1:7])), "nan_centre": bool(v[7])}

    def optimizer_step(self, dcentre, dscale, drotation, dopacity, demission_mlps, groups: int = gh.GROUPS_ALL,
                       stream=None):
        """One step of the groups in `groups` from gradients with respect to
        the activations (as the rasterizer's backward returns them).  No
        synchronisation; can be captured."""
        conf = self.conf is not None
        L = gh.lib()
        s = stream if stream is not None else gh.stream_handle(self.device)
        gh.check(L.gsparc_adam_step(self.P, int(groups), C.byref(self.adam), C.byref(self.struct()), gh.ptr(dcentre),
                                    gh.ptr(dscale), gh.ptr(drotation), gh.ptr(dopacity), gh.ptr(demission_mlps),
                                    self.conf.numel() if conf else 0, gh.ptr(self.conf) if conf else None,
                                    gh.ptr(self.conf_grad) if conf else None, gh.ptr(self.conf_state) if conf else None,
                                    gh.ptr(self.train_state), s), "gsparc_adam_step")

    def add_densification_stats(self, dmeans2D, radii, grad_scale: float = 1.0, stream=None):
        """Statistics of one batch (dmeans2D [R,P,3], radii [R,P])."""
        R = int(radii.numel() // max(self.P, 1)) if self.P else 0
        s = stream if stream is not None else gh.stream_handle(self.device)
        gh.check(gh.lib().gsparc_densify_stats(self.P, R, float(grad_scale), gh.ptr(dmeans2D), gh.ptr(radii),
                                               C.byref(self.struct()), s), "gsparc_densify_stats")

    def reset_opacity(self):
        gh.check(gh.lib().gsparc_opacity_reset(self.P, C.byref(self.struct()), gh.stream_handle(self.device)),
                 "gsparc_opacity_reset")

    # ---- densify and prune
    def _compact(self, cfg: gh.DensifyConfig, noise):
        if self.P == 0:
            return 0, 0, 0
        L = gh.lib()
        other = 1 - self.cur
        self._ensure(other, self.P)
        if noise is None:
            noise = torch.randn((SPLIT_N * self.P, 3), device=self.device)
        noise = gh.f32(noise)
        if noise.numel() < 3 * SPLIT_N * self.P:
            raise RuntimeError("noise must hold [2P, 3] standard normals")
        work = torch.empty(L.gsparc_densify_work_size(self.P), dtype=torch.uint8, device=self.device)
        pn, nc, ns = C.c_int(0), C.c_int(0), C.c_int(0)
        for _ in range(2):
            st = L.gsparc_densify_and_prune(self.P, C.byref(cfg), C.byref(self.struct()), gh.ptr(noise),
                                            C.byref(self.banks[other].struct()), self.banks[other].capacity,
                                            gh.ptr(work), C.byref(pn), C.byref(nc), C.byref(ns),
                                            gh.stream_handle(self.device))
            if st == gh.ERR_CAPACITY:
                self.banks[other] = None
                self._ensure(other, pn.value)
                continue
            gh.check(st, "gsparc_densify_and_prune")
            break
        self.cur = other
        self.P = pn.value
        return pn.value, nc.value, ns.value

    def densify_and_prune(self, max_grad, min_opacity, extent, max_screen_size, radii=None, scene_min=None,
                          scene_max=None, noise=None):
        """densify_and_prune (gaussian.py:510-535).  noise: standard normals
        [2P, 3] for the split children (row k: first child of the k-th split
        Gaussian, row n_split + k its second), drawn here when None.  radii
        only feeds the reference's tmp_radii, which nothing reads; it is
        accepted and ignored.  Returns (P_new, n_clone, n_split)."""
        cfg = gh.DensifyConfig()
        cfg.densify = 1
        cfg.grad_threshold = float(np.float32(max_grad))
        cfg.clone_max_scale = float(np.float32(self.percent_dense * extent))
        cfg.min_opacity = float(np.float32(min_opacity))
        cfg.prune_max_scale = float(np.float32(0.1 * extent)) if max_screen_size else float("inf")
        self._bounds(cfg, scene_min, scene_max)
        return self._compact(cfg, noise)

    def prune(self, min_opacity, extent, scene_min=None, scene_max=None):
        """The post-densification pruning (train.py:539-555): opacity <
        min_opacity, max scale > 0.1 extent, centre out of bounds.  Returns
        the number removed; when it is 0 the set is left as it was (the
        reference then does not replace its parameters and they step)."""
        cfg = gh.DensifyConfig()
        cfg.densify = 0
        cfg.grad_threshold = float("inf")
        cfg.clone_max_scale = 0.0
        cfg.min_opacity = float(np.float32(min_opacity))
        cfg.prune_max_scale = float(np.float32(0.1 * extent))
        self._bounds(cfg, scene_min, scene_max)
        P0 = self.P
        cur0 = self.cur
        pn, _, _ = self._compact(cfg, None)
        if pn == P0:
            self.cur, self.P = cur0, P0
        return P0 - pn

    @staticmethod
    def _bounds(cfg, scene_min, scene_max):
        cfg.use_bounds = 0 if scene_min is None else 1
        if scene_min is not None:
            for k in range(3):
                cfg.scene_min[k] = float(np.float32(scene_min[k]))
                cfg.scene_max[k] = float(np.float32(scene_max[k]))

    # ---- checkpoints
    def capture(self) -> dict:
        """The model and optimizer state as CPU tensors (P-entry layouts)."""
        out = {n: self.array(n).detach().cpu().clone() for n, _ in _FIELDS}
        out["P"] = self.P
        out["train_state"] = self.train_state.detach().cpu().clone() if self.train_state is not None else None
        if self.conf is not None:
            out["confidence"] = self.conf.detach().cpu().clone()
            out["confidence_state"] = self.conf_state.detach().cpu().clone()
        return out

    def restore(self, st: dict):
        """Load a capture() snapshot (or a checkpoint's, which has no
        activations: they are recomputed)."""
        P = int(st["P"])
        self._ensure(self.cur, P)
        self.P = P
        for t in self.bank.t.values():
            t.zero_()
        for n, w in _FIELDS:
            if st.get(n) is not None:
                self.bank.t[n][: w * P].copy_(st[n].reshape(-1))
        if any(st.get(n) is None for n in ("scale", "rotation_n", "opacity")):
            self.activate()
        if st.get("train_state") is not None:
            if self.train_state is None:
                self.train_state = torch.zeros(gh.TRAIN_STATE_BYTES // 4, dtype=torch.int32, device=self.device)
            self.train_state.copy_(st["train_state"].to(torch.int32).reshape(-1))
        if st.get("confidence") is not None:
            n = int(st["confidence"].numel())
            if self.conf is None or self.conf.numel() != n:
                self.conf = torch.zeros(n, dtype=torch.float32, device=self.device)
                self.conf_grad = torch.zeros_like(self.conf)
                self.conf_state = torch.zeros(2 * n, dtype=torch.float32, device=self.device)
                self.conf_hidden = (n - 1) // 5
            self.conf.copy_(st["confidence"].reshape(-1))
            self.conf_state.copy_(st["confidence_state"].reshape(-1))
