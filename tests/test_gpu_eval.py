"""GPU tests of the evaluation path and of checkpoints through GaussianModel.

Metrics: gsparc_eval_metrics against the reference's formulas
(train.py:76-102) evaluated in FP64 on the CPU, with the SSIM map from the
oracle's FP64 twin (pinned against textbook SSIM in the oracle tests), and
the confidence against ConfidenceMLP.forward.  Checkpoints: a trained model
saved in the reference's format and loaded into a new model gives identical
arrays; training resumed from it continues as the original does; the split
evaluation writes the reference's files."""
import csv
import os

import numpy as np
import pytest
import torch

import gsparc_checkpoint as CK
import gsparc_eval as E
import gsparc_gaussian as G
import gsparc_scenes as S
import gsparc_train as T
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("R,H,W", [(5, 90, 360), (3, 37, 50)])
def test_eval_metrics_match_the_reference_formulas(R, H, W):
    g = torch.Generator().manual_seed(R + H)
    img = torch.rand((R, H, W), generator=g) * 0.3
    tgt = torch.rand((R, H, W), generator=g) * 0.3
    rx = torch.rand((R, 3), generator=g) * 5
    torch.manual_seed(3)
    net = T.ConfidenceMLP(32)
    conf = net.packed().detach().clone()
    m = E.eval_metrics(img.cuda(), tgt.cuda(), rx.cuda(), conf.cuda()).cpu().numpy().astype(np.float64)
    x = img.numpy().astype(np.float64)
    y = tgt.numpy().astype(np.float64)
    ssim_map = O.ssim_fwd(x[:, None], y[:, None], dtype=np.float64, train=False)[0][:, 0]
    for r in range(R):
        d = x[r] - y[r]
        mse = np.mean(d * d)
        mae = np.mean(np.abs(d))
        ssim = ssim_map[r].mean()
        rp = 10 * np.log10(np.mean(x[r] ** 2)) - 30
        rt = 10 * np.log10(np.mean(y[r] ** 2)) - 30
        want = [0.8 * mae + 0.2 * (1 - ssim), 20 * np.log10(1 / np.sqrt(mse)), ssim, mae, mse, rp, rp - rt,
                float(net(rx[r]).detach())]
        np.testing.assert_allclose(m[r], want, rtol=2e-5, atol=1e-6, err_msg=str(r))
    # without a network the confidence is 1; identical inputs give PSNR inf and SSIM 1
    m2 = E.eval_metrics(img.cuda(), img.cuda()).cpu().numpy()
    assert np.all(m2[:, 7] == 1.0) and np.all(np.isinf(m2[:, 1])) and np.allclose(m2[:, 2], 1.0, atol=1e-6)


def _trained(iters=4, seed=91):
    from test_gpu_train_surround import _trainer
    tr, sc, d, target, net = _trainer(seed=seed, densify_from_iter=2, densification_interval=3,
                                      opacity_reset_interval=100, densify_until_iter=50,
                                      densify_grad_threshold=2e-5)
    for _ in range(iters):
        tr.step()
    torch.cuda.synchronize()
    return tr, sc, d, target


def test_checkpoint_save_load_and_resume(tmp_path):
    tr, sc, d, target = _trained()
    it = tr.iteration - 1
    path = CK.save_checkpoint(tr.m, it, str(tmp_path / f"checkpoint_iter_{it}.pth"), tr.o)
    m2 = G.GaussianModel()
    assert CK.load_checkpoint(m2, path) == it
    for n in ("centre", "log_scale", "rotation", "logit", "emission_mlps", "state", "state_mlp", "scale",
              "rotation_n", "opacity"):
        assert torch.equal(m2.array(n), tr.m.array(n)), n
    assert torch.equal(m2.conf, tr.m.conf) and torch.equal(m2.conf_state, tr.m.conf_state)
    assert m2.counters()["steps"] == tr.m.counters()["steps"]
    assert m2.counters()["iteration"] == it + 1
    # resume: the restored model continues as the original does
    import gsparc_trainer as TR
    import gsparc_scenes as S_
    tr2 = TR.Trainer(m2, tr.o, d["views"], d["tx"], d["rx"], target, sc.H, sc.W, xyz_min=[0.0, 0.0, 0.0],
                     xyz_max=S_.ROOM, first_iteration=None, k_capacity=1 << 22)
    assert tr2.iteration == tr.iteration
    tr.step()
    tr2.step()
    torch.cuda.synchronize()
    assert abs(float(tr.loss) - float(tr2.loss)) <= 1e-4 * abs(float(tr.loss))
    for n in ("centre", "log_scale", "logit", "emission_mlps"):
        a, b = tr.m.array(n), tr2.m.array(n)
        tight = (a - b).abs() <= 1e-5 * b.abs() + 1e-6
        assert float(tight.float().mean()) > 0.9995, n
    assert tr.m.counters()["steps"] == tr2.m.counters()["steps"]
    tr.close()
    tr2.close()


def test_evaluate_split_writes_the_reference_files(tmp_path):
    tr, sc, d, target = _trained(iters=2, seed=93)
    res = E.evaluate_split(tr.m, d["views"], d["tx"], d["rx"], target, sc.H, sc.W, split_name="test", batch=3,
                           out_dir=str(tmp_path))
    R = sc.R
    assert len(res["per_sample_rows"]) == R and res["rendered"].shape == (R, sc.H, sc.W)
    # the rendering is the model's forward; the metrics those of eval_metrics on it
    m = E.eval_metrics(torch.from_numpy(res["rendered"]).cuda(), target, d["rx"], tr.m.conf).cpu().numpy()
    np.testing.assert_array_equal(res["metrics"], m)
    assert abs(res["avg_psnr"] - float(np.mean(m[:, 1]))) < 1e-4
    arr = np.load(os.path.join(tmp_path, "gs_output_real_test.npy"), allow_pickle=False)
    assert np.array_equal(arr, res["rendered"])
    with open(os.path.join(tmp_path, "per_sample_metrics_test.csv")) as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["sample_idx", "sample_id", "split", "psnr", "ssim", "mae", "rssi_diff", "render_time_ms",
                       "loss", "confidence"]
    assert len(rows) == R + 1
    tr.close()
