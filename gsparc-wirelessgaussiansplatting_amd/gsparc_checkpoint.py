"""Checkpoints in the reference's format (train.py:225-311).

A checkpoint is the dictionary Trainer.save_checkpoint writes with
torch.save:
  'iteration'                  the last completed iteration
  'model_state_dict'           '_centre' [P,3], '_scaling' [P,3] (log),
                               '_rotation' [P,4] (raw), '_opacity' [P,1]
                               (logit), 'emission_fc1_weights' [P,32,3],
                               'emission_fc1_bias' [P,32],
                               'emission_fc2_weights' [P,2,32],
                               'emission_fc2_bias' [P,2]
  'optimizer_state_dict'       torch.optim.Adam.state_dict() over the groups
                               of training_setup (gaussian.py:175-197):
                               centre, scaling, rotation, opacity, the four
                               emission groups, then confidence_fc1.weight,
                               confidence_fc1.bias, confidence_fc2.weight,
                               confidence_fc2.bias; per parameter 'step',
                               'exp_avg', 'exp_avg_sq'
  'confidence_mlp_state_dict'  the ConfidenceMLP state dict (fc1.weight
                               [h,3], fc1.bias [h], fc2.weight [1,h],
                               fc2.bias [1]) or None
Files written here load in the reference (its load_checkpoint and the
optimizer's load_state_dict), and the reference's files load here.  Reading
uses torch.load(weights_only=True): tensors, numbers and strings only.
"""
from __future__ import annotations

import collections
import os

import torch

from gsparc_gaussian import MLP_FLOATS, OptimizationParams, expon_lr

MODEL_KEYS = ("_centre", "_scaling", "_rotation", "_opacity", "emission_fc1_weights", "emission_fc1_bias",
              "emission_fc2_weights", "emission_fc2_bias")
GROUPS = ("centre", "scaling", "rotation", "opacity", "emission_fc1_weights", "emission_fc1_bias",
          "emission_fc2_weights", "emission_fc2_bias")
CONF_PARAMS = ("fc1.weight", "fc1.bias", "fc2.weight", "fc2.bias")
# the library's moment layout: [2][centre 3 | scaling 3 | rotation 4 | logit 1] P
# and [2][W1 96 | b1 32 | W2 64 | b2 2] P
_GEO = (("centre", 3, (3,)), ("scaling", 3, (3,)), ("rotation", 4, (4,)), ("opacity", 1, (1,)))
_MLP = (("emission_fc1_weights", 96, (32, 3)), ("emission_fc1_bias", 32, (32,)),
        ("emission_fc2_weights", 64, (2, 32)), ("emission_fc2_bias", 2, (2,)))
# train-state step counter of each reference group
_STEP_SLOT = {"centre": 1, "scaling": 2, "rotation": 3, "opacity": 4, "emission_fc1_weights": 5,
              "emission_fc1_bias": 5, "emission_fc2_weights": 5, "emission_fc2_bias": 5}
_CONF_SLOT = 6


def _conf_shapes(hidden):
    return ((hidden, 3), (hidden,), (1, hidden), (1,))


def _split(flat, P, table, half=None):
    """Sections of a [P*w] (or [2][P*w], half 0/1) array as [P, ...] tensors."""
    tot = sum(w for _, w, _ in table)
    base = 0 if half is None else half * tot * P
    out, o = {}, base
    for name, w, shape in table:
        out[name] = flat[o: o + w * P].reshape((P,) + shape).clone()
        o += w * P
    return out


def _join(parts, table):
    return torch.cat([parts[name].reshape(-1) for name, _, _ in table])


def _adam_layout(optim_params, spatial_lr_scale, iteration, hidden):
    """An empty torch.optim.Adam.state_dict() with the reference's groups,
    built by torch itself so that its keys are this torch's."""
    o = optim_params
    lrs = {"centre": expon_lr(iteration, o.position_lr_init * spatial_lr_scale, o.position_lr_final * spatial_lr_scale,
                              0, o.position_lr_delay_mult, o.position_lr_max_steps),
           "scaling": o.scaling_lr, "rotation": o.rotation_lr, "opacity": o.opacity_lr}
    groups = [{"params": [torch.nn.Parameter(torch.zeros(1))], "lr": lrs.get(n, o.emission_mlp_lr), "name": n}
              for n in GROUPS]
    if hidden:
        groups += [{"params": [torch.nn.Parameter(torch.zeros(1))], "lr": getattr(o, "confidence_lr", 0.001),
                    "name": f"confidence_{n}"} for n in CONF_PARAMS]
    opt = torch.optim.Adam(groups, lr=0.0, eps=1e-15)
    for g in opt.param_groups:
        g["clip_grad_norm"] = 1.0        # training_setup adds it (gaussian.py:199-201)
    return opt.state_dict()


def to_checkpoint(snap: dict, iteration: int, optim_params: OptimizationParams | None = None,
                  spatial_lr_scale: float = 1.0, hidden: int = 32) -> dict:
    """The reference's checkpoint dictionary from GaussianModel.capture()."""
    o = optim_params or OptimizationParams()
    P = int(snap["P"])
    model = collections.OrderedDict()
    model["_centre"] = snap["centre"].reshape(P, 3).clone()
    model["_scaling"] = snap["log_scale"].reshape(P, 3).clone()
    model["_rotation"] = snap["rotation"].reshape(P, 4).clone()
    model["_opacity"] = snap["logit"].reshape(P, 1).clone()
    mlp = _split(snap["emission_mlps"].reshape(-1), P, _MLP)
    for name, _, _ in _MLP:
        model[name] = mlp[name]
    conf = snap.get("confidence")
    has_conf = conf is not None
    sd = _adam_layout(o, spatial_lr_scale, iteration, hidden if has_conf else 0)
    counters = snap["train_state"].reshape(-1) if snap.get("train_state") is not None else torch.zeros(8)
    m_geo = _split(snap["state"].reshape(-1), P, _GEO, 0)
    v_geo = _split(snap["state"].reshape(-1), P, _GEO, 1)
    m_mlp = _split(snap["state_mlp"].reshape(-1), P, _MLP, 0)
    v_mlp = _split(snap["state_mlp"].reshape(-1), P, _MLP, 1)
    m_all = {**m_geo, **m_mlp}
    v_all = {**v_geo, **v_mlp}
    for i, name in enumerate(GROUPS):
        t = int(counters[_STEP_SLOT[name]])
        if t > 0:
            sd["state"][i] = {"step": torch.tensor(float(t)), "exp_avg": m_all[name], "exp_avg_sq": v_all[name]}
    conf_sd = None
    if has_conf:
        conf = conf.reshape(-1)
        n = conf.numel()
        cst = snap["confidence_state"].reshape(-1)
        conf_sd = collections.OrderedDict()
        off = 0
        t = int(counters[_CONF_SLOT])
        for k, (pname, shape) in enumerate(zip(CONF_PARAMS, _conf_shapes(hidden))):
            w = 1
            for d in shape:
                w *= d
            conf_sd[pname] = conf[off: off + w].reshape(shape).clone()
            if t > 0:
                sd["state"][len(GROUPS) + k] = {"step": torch.tensor(float(t)),
                                                "exp_avg": cst[off: off + w].reshape(shape).clone(),
                                                "exp_avg_sq": cst[n + off: n + off + w].reshape(shape).clone()}
            off += w
    return {"iteration": int(iteration), "model_state_dict": model, "optimizer_state_dict": sd,
            "confidence_mlp_state_dict": conf_sd}


def from_checkpoint(ck: dict) -> dict:
    """A GaussianModel.restore() snapshot (CPU tensors, the library's
    layouts) from the reference's checkpoint dictionary.  Optimizer groups
    are matched by their 'name'; a group without state restarts at zero."""
    ms = ck["model_state_dict"]
    P = int(ms["_centre"].shape[0])
    f = lambda t: torch.as_tensor(t).detach().float().cpu()
    snap = {"P": P, "centre": f(ms["_centre"]).reshape(-1), "log_scale": f(ms["_scaling"]).reshape(-1),
            "rotation": f(ms["_rotation"]).reshape(-1), "logit": f(ms["_opacity"]).reshape(-1),
            "emission_mlps": torch.cat([f(ms[k]).reshape(-1) for k in MODEL_KEYS[4:]])}
    if snap["emission_mlps"].numel() != MLP_FLOATS * P:
        raise RuntimeError("the checkpoint's emission MLPs are not 3 -> 32 -> 2 per Gaussian")
    zeros = lambda shape: torch.zeros((P,) + shape)
    m = {n: zeros(s) for n, _, s in _GEO + _MLP}
    v = {n: zeros(s) for n, _, s in _GEO + _MLP}
    counters = torch.zeros(64, dtype=torch.int32)
    counters[0] = int(ck.get("iteration", 0)) + 1
    osd = ck.get("optimizer_state_dict") or {"state": {}, "param_groups": []}
    by_name = {}
    for g in osd.get("param_groups", []):
        for pid in g["params"]:
            by_name[g.get("name")] = osd["state"].get(pid)
    for name in GROUPS:
        st = by_name.get(name)
        if st:
            m[name] = f(st["exp_avg"]).reshape(m[name].shape)
            v[name] = f(st["exp_avg_sq"]).reshape(v[name].shape)
            counters[_STEP_SLOT[name]] = int(float(st["step"]))
    snap["state"] = torch.cat([_join(m, _GEO), _join(v, _GEO)])
    snap["state_mlp"] = torch.cat([_join(m, _MLP), _join(v, _MLP)])
    for k in ("accum", "denom", "max_radius"):
        snap[k] = torch.zeros(P)
    conf_sd = ck.get("confidence_mlp_state_dict")
    snap["confidence"] = None
    if conf_sd is not None:
        parts = [f(conf_sd[p_]).reshape(-1) for p_ in CONF_PARAMS]
        snap["confidence"] = torch.cat(parts)
        mm, vv = [], []
        for p_, t_ in zip(CONF_PARAMS, parts):
            st = by_name.get(f"confidence_{p_}")
            if st:
                mm.append(f(st["exp_avg"]).reshape(-1))
                vv.append(f(st["exp_avg_sq"]).reshape(-1))
                counters[_CONF_SLOT] = int(float(st["step"]))
            else:
                mm.append(torch.zeros_like(t_))
                vv.append(torch.zeros_like(t_))
        snap["confidence_state"] = torch.cat(mm + vv)
        snap["confidence_hidden"] = int(conf_sd["fc1.weight"].shape[0])
    snap["train_state"] = counters
    return snap


def save_checkpoint(model, iteration: int, path: str, optim_params: OptimizationParams | None = None):
    """Trainer.save_checkpoint (train.py:225-254) for a GaussianModel."""
    snap = model.capture()
    ck = to_checkpoint(snap, iteration, optim_params, model.spatial_lr_scale, model.conf_hidden or 32)
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    torch.save(ck, path)
    return path


def load_checkpoint(model, path: str) -> int:
    """Trainer.load_checkpoint (train.py:256-311) into a GaussianModel: the
    parameters, the Adam moments and step counts, the confidence network and
    its moments (when the model was set up with one).  Returns the
    checkpoint's iteration; the model's next step is iteration + 1."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"Checkpoint file not found: {path}")
    ck = torch.load(path, map_location="cpu", weights_only=True)
    snap = from_checkpoint(ck)
    model.restore(snap)
    return int(ck["iteration"])
