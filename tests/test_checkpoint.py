"""Checkpoint compatibility with the reference (train.py:225-311), on the CPU.

A checkpoint written the reference's way (its parameter tensors, the state of
a torch.optim.Adam with training_setup's groups after a few steps, the
ConfidenceMLP state dict) goes through torch.save / torch.load
(weights_only=True), into the library's layouts and back: every tensor and
step count survives unchanged, and the result loads into a fresh optimizer of
the reference's structure (optimizer.load_state_dict) and into a
ConfidenceMLP."""
import numpy as np
import torch

import gsparc_checkpoint as CK
import gsparc_gaussian as G
import gsparc_train as T


def _reference_style(P=50, hidden=32, steps=3, seed=0):
    g = torch.Generator().manual_seed(seed)
    p = {"_centre": torch.randn((P, 3), generator=g), "_scaling": torch.randn((P, 3), generator=g),
         "_rotation": torch.randn((P, 4), generator=g), "_opacity": torch.randn((P, 1), generator=g),
         "emission_fc1_weights": torch.randn((P, 32, 3), generator=g),
         "emission_fc1_bias": torch.randn((P, 32), generator=g),
         "emission_fc2_weights": torch.randn((P, 2, 32), generator=g),
         "emission_fc2_bias": torch.randn((P, 2), generator=g)}
    params = {k: torch.nn.Parameter(v) for k, v in p.items()}
    torch.manual_seed(seed)
    net = T.ConfidenceMLP(hidden)
    o = G.OptimizationParams()
    names = ["centre", "scaling", "rotation", "opacity", "emission_fc1_weights", "emission_fc1_bias",
             "emission_fc2_weights", "emission_fc2_bias"]
    lrs = [o.position_lr_init, o.scaling_lr, o.rotation_lr, o.opacity_lr] + [o.emission_mlp_lr] * 4
    groups = [{"params": [params[k]], "lr": lr, "name": n} for k, lr, n in zip(p, lrs, names)]
    groups += [{"params": [q], "lr": o.confidence_lr, "name": f"confidence_{n}"} for n, q in net.named_parameters()]
    opt = torch.optim.Adam(groups, lr=0.0, eps=1e-15)
    for gr in opt.param_groups:
        gr["clip_grad_norm"] = 1.0
    for s in range(steps):
        for q in list(params.values()) + list(net.parameters()):
            q.grad = torch.randn(q.shape, generator=g)
        if s == steps - 1:
            params["_opacity"].grad = None       # one group a step behind, as after a reset
        opt.step()
    ck = {"iteration": 1234, "model_state_dict": params, "optimizer_state_dict": opt.state_dict(),
          "confidence_mlp_state_dict": net.state_dict()}
    return ck, params, net, opt


def test_reference_checkpoint_round_trip(tmp_path):
    ck, params, net, opt = _reference_style()
    path = str(tmp_path / "checkpoint_iter_1234.pth")
    torch.save(ck, path)
    loaded = torch.load(path, map_location="cpu", weights_only=True)
    snap = CK.from_checkpoint(loaded)
    assert snap["P"] == 50 and int(snap["train_state"][0]) == 1235
    steps = snap["train_state"][1:7].tolist()
    assert steps == [3, 3, 3, 2, 3, 3], steps
    back = CK.to_checkpoint(snap, 1234, G.OptimizationParams(), 1.0, 32)
    for k, v in params.items():
        assert torch.equal(back["model_state_dict"][k], v.detach()), k
    for k, v in net.state_dict().items():
        assert torch.equal(back["confidence_mlp_state_dict"][k], v), k
    so, sb = ck["optimizer_state_dict"]["state"], back["optimizer_state_dict"]["state"]
    assert sorted(so) == sorted(sb)
    for i in so:
        assert float(so[i]["step"]) == float(sb[i]["step"])
        assert torch.equal(so[i]["exp_avg"], sb[i]["exp_avg"].reshape(so[i]["exp_avg"].shape)), i
        assert torch.equal(so[i]["exp_avg_sq"], sb[i]["exp_avg_sq"].reshape(so[i]["exp_avg_sq"].shape)), i
    go, gb = ck["optimizer_state_dict"]["param_groups"], back["optimizer_state_dict"]["param_groups"]
    assert [g["name"] for g in go] == [g["name"] for g in gb]
    assert [g["params"] for g in go] == [g["params"] for g in gb]
    # the written checkpoint loads into the reference's structures
    path2 = str(tmp_path / "ours.pth")
    torch.save(back, path2)
    again = torch.load(path2, map_location="cpu", weights_only=True)
    _, params2, net2, opt2 = _reference_style(steps=0, seed=9)
    opt2.load_state_dict(again["optimizer_state_dict"])
    net2.load_state_dict(again["confidence_mlp_state_dict"])
    for i, gr in enumerate(opt2.param_groups):
        q = gr["params"][0]
        st = opt2.state.get(q)
        if i in so:
            assert torch.equal(st["exp_avg"], so[i]["exp_avg"])
    assert torch.equal(net2.fc1.weight, net.fc1.weight)


def test_library_layout_survives_the_format():
    """A snapshot in the library's layouts (as GaussianModel.capture gives)
    goes to the reference's dictionary and back unchanged."""
    P, h = 40, 16
    rng = np.random.default_rng(3)
    t = lambda n: torch.from_numpy(rng.normal(size=n).astype(np.float32))
    st = torch.zeros(64, dtype=torch.int32)
    st[:8] = torch.tensor([77, 5, 5, 5, 4, 5, 9, 0])
    snap = {"P": P, "centre": t(3 * P), "log_scale": t(3 * P), "rotation": t(4 * P), "logit": t(P),
            "emission_mlps": t(194 * P), "state": t(22 * P), "state_mlp": t(388 * P), "train_state": st,
            "confidence": t(5 * h + 1), "confidence_state": t(2 * (5 * h + 1))}
    ck = CK.to_checkpoint(snap, 76, G.OptimizationParams(), 1.0, h)
    back = CK.from_checkpoint(ck)
    for k in ("centre", "log_scale", "rotation", "logit", "emission_mlps", "state", "state_mlp", "confidence",
              "confidence_state"):
        assert torch.equal(back[k].reshape(-1), snap[k].reshape(-1)), k
    assert back["train_state"][:7].tolist() == [77, 5, 5, 5, 4, 5, 9]
    # the centre's rate in the checkpoint is the schedule's at that iteration
    lr = ck["optimizer_state_dict"]["param_groups"][0]["lr"]
    assert abs(lr - G.expon_lr(76, 0.00016, 1.6e-7, 0, 0.01, 200000)) < 1e-18
