"""The evaluation path (train.py:613-946) on the library's kernels.

evaluate_split renders a split's receivers in batches, computes the
reference's per-sample metrics on the GPU in one SSIM pass
(gsparc_eval_metrics: loss, PSNR, SSIM, MAE, MSE, RSSI difference,
confidence) and returns what Trainer.evaluate's evaluate_split returns:
averages, the per-sample value lists and rows, and the rendered spectra.
With an output folder it also writes the reference's files
gs_output_real_{split}.npy and per_sample_metrics_{split}.csv.  The plots the
reference draws next to them are not produced.
"""
from __future__ import annotations

import csv
import os

import numpy as np
import torch

import gsparc_hip as gh
import wireless_rasterizer as wr
from gsparc_train import LAMBDA_SSIM, SSIM_C1, SSIM_C2

METRICS = ("loss", "psnr", "ssim", "mae", "mse", "rssi", "rssi_diff", "confidence")


def eval_metrics(img, target, rx_pos=None, conf=None, lambda_ssim: float = LAMBDA_SSIM):
    """Per-spectrum metrics [R, 8] (columns METRICS) of img against target
    [R,H,W]; conf: the packed confidence network (C = 1 without it)."""
    x = gh.f32(img)
    y = gh.f32(target)
    if x.dim() == 2:
        x, y = x.unsqueeze(0), y.unsqueeze(0)
    if x.shape != y.shape or x.dim() != 3:
        raise RuntimeError("img and target must both be [H,W] or [R,H,W]")
    R, H, W = x.shape
    mlp = None if conf is None else gh.f32(conf).reshape(-1)
    rx = None if rx_pos is None else gh.f32(rx_pos).reshape(-1, 3)
    if mlp is not None and (rx is None or rx.shape[0] != R):
        raise RuntimeError("the confidence needs one receiver position per spectrum")
    gh.require_cuda(x, y, rx, mlp)
    L = gh.lib()
    out = torch.empty((R, len(METRICS)), dtype=torch.float32, device=x.device)
    work = torch.empty(L.gsparc_eval_work_size(R, H, W), dtype=torch.uint8, device=x.device)
    hidden = 0 if mlp is None else (mlp.numel() - 1) // 5
    gh.check(L.gsparc_eval_metrics(R, H, W, float(lambda_ssim), SSIM_C1, SSIM_C2, hidden, gh.ptr(mlp), gh.ptr(rx),
                                   gh.ptr(x), gh.ptr(y), gh.ptr(work), gh.ptr(out), gh.stream_handle(x.device)),
             "evaluation metrics")
    return out


def evaluate_split(model, views, tx_pos, rx_pos, targets, H, W, split_name="test", batch: int = 64,
                   lambda_ssim: float = LAMBDA_SSIM, out_dir: str | None = None, sample_ids=None):
    """Render and score every receiver of a split.  model: a GaussianModel
    (its activations and confidence network are used); views [N,4,4],
    rx_pos [N,3], tx_pos [3], targets [N,H,W] on the device."""
    N = int(rx_pos.shape[0])
    views = gh.f32(views).reshape(N, 16)
    rx = gh.f32(rx_pos).reshape(N, 3)
    tx = gh.f32(tx_pos).reshape(3)
    conf = model.conf
    dev = model.device
    rendered, metrics, times = [], [], []
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    with torch.no_grad():
        for b0 in range(0, N, batch):
            b1 = min(N, b0 + batch)
            start.record()
            out = wr.forward_batched(model.get_centre, model.get_emission_mlps, model.get_opacity,
                                     model.get_scaling, model.get_rotation, H, W, 1.0, views[b0:b1], tx, rx[b0:b1])
            end.record()
            img = out[1]
            m = eval_metrics(img, targets[b0:b1], rx[b0:b1], conf, lambda_ssim)
            torch.cuda.synchronize(dev)
            ms = start.elapsed_time(end)
            times += [ms / (b1 - b0)] * (b1 - b0)
            rendered.append(img.cpu().numpy())
            metrics.append(m.cpu().numpy())
    rendered = np.concatenate(rendered, axis=0) if rendered else np.zeros((0, H, W), np.float32)
    mt = np.concatenate(metrics, axis=0) if metrics else np.zeros((0, len(METRICS)), np.float32)
    ids = list(range(N)) if sample_ids is None else list(sample_ids)
    rows = []
    for i in range(N):
        rows.append({"sample_idx": i, "sample_id": ids[i], "split": split_name, "psnr": float(mt[i, 1]),
                     "ssim": float(mt[i, 2]), "mae": float(mt[i, 3]), "rssi_diff": float(mt[i, 6]),
                     "render_time_ms": float(times[i]), "loss": float(mt[i, 0]), "confidence": float(mt[i, 7])})
    mean = lambda k: float(np.mean([r[k] for r in rows])) if rows else 0.0
    res = {"avg_loss": mean("loss"), "avg_psnr": mean("psnr"), "avg_ssim": mean("ssim"),
           "avg_rssi_diff": mean("rssi_diff"), "avg_mae": mean("mae"), "avg_confidence": mean("confidence"),
           "avg_render_time": mean("render_time_ms"),
           "psnr_values": [r["psnr"] for r in rows], "ssim_values": [r["ssim"] for r in rows],
           "rssi_diff_values": [r["rssi_diff"] for r in rows], "mae_values": [r["mae"] for r in rows],
           "confidence_values": [r["confidence"] for r in rows], "per_sample_rows": rows, "rendered": rendered,
           "metrics": mt}
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        np.save(os.path.join(out_dir, f"gs_output_real_{split_name}.npy"), rendered)
        with open(os.path.join(out_dir, f"per_sample_metrics_{split_name}.csv"), "w", newline="") as f:
            w = csv.writer(f)
            cols = ["sample_idx", "sample_id", "split", "psnr", "ssim", "mae", "rssi_diff", "render_time_ms",
                    "loss", "confidence"]
            w.writerow(cols)
            for r in rows:
                w.writerow([r[c] for c in cols])
    return res
