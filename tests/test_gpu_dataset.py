"""The dataset on the device: view matrices against build_view_matrix's
computation (scipy's Rotation in double, rounded to float, and the float
product for the translation; preprocess_data.py:177-202), the resident split
against the files, the batch gather, and a trainer fed by the dataset whose
graph replays draw the reference's sample order."""
import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

import gsparc_dataset as D
import gsparc_gaussian as G
import gsparc_hip as gh
import gsparc_scenes as S
from dataset_fixture import QUAT, make_dataset

pytestmark = pytest.mark.gpu


def _reference_view(rx, quat):
    R_c2w = torch.from_numpy(Rotation.from_quat(np.asarray(quat, np.float64)).as_matrix()).float()
    R_w2c = R_c2w.transpose(0, 1)
    C = torch.as_tensor(rx, dtype=torch.float32)
    t = -(R_w2c @ C)
    v = torch.eye(4)
    v[:3, :3] = R_w2c
    v[:3, 3] = t
    return v.transpose(0, 1).contiguous()


@pytest.mark.parametrize("quat", [QUAT, (0.0, 0.0, 0.0, 1.0), (0.5, 0.5, -0.5, 0.5), (0.0, 0.7, 0.0, 0.7)])
def test_view_matrices_match_build_view_matrix(quat):
    rng = np.random.default_rng(1)
    rx = rng.uniform(-8, 8, size=(777, 3)).astype(np.float32)
    got = D.view_matrices(torch.from_numpy(rx).cuda(), quat).cpu()
    for i in range(rx.shape[0]):
        ref = _reference_view(rx[i], quat)
        # the rotation part exactly; the translation within 2 float ulps (product order)
        assert torch.equal(got[i][:3, :3], ref[:3, :3]), i
        assert torch.equal(got[i][:, 3], ref[:, 3]), i
        t, tr = got[i][3, :3].numpy(), ref[3, :3].numpy()
        assert np.all(np.abs(t - tr) <= 2 * np.spacing(np.abs(tr)) + 1e-12), (i, t, tr)
    if quat == (0.0, 0.0, 0.0, 1.0):
        # identity orientation: the synthetic scenes' view matrices
        for i in range(20):
            assert np.array_equal(got[i].numpy().reshape(-1), S.view_matrix(rx[i]).reshape(-1))


def test_resident_split_and_batch_gather(tmp_path):
    root = str(tmp_path)
    rx, spectra = make_dataset(root, n=30, H=12, W=40)
    idx = [f"{i:05d}" for i in (5, 1, 17, 30, 9, 22, 2)]
    np.savetxt(str(tmp_path / "train_index.txt"), np.array(idx), fmt="%s")
    ds = D.SpectrumDataset(root, str(tmp_path / "train_index.txt"))
    assert len(ds) == 7 and (ds.n_elevation, ds.n_azimuth) == (12, 40)
    rows = [int(s) - 1 for s in idx]
    assert np.array_equal(ds.spectra.cpu().numpy(), spectra[rows])
    assert np.allclose(ds.rx_pos.cpu().numpy(), rx[rows], atol=1e-6)
    assert torch.equal(ds.viewmatrices.cpu(), D.view_matrices(ds.rx_pos, QUAT).cpu())
    w = ds[3]
    assert torch.equal(w.spectrum_real, ds.spectra[3]) and w.spectrum_elevation == 12
    # the batch of an iteration: order[(it - 1) R + r], wrapping
    order = torch.tensor([6, 0, 3, 3, 5, 1, 2, 4, 0, 6], dtype=torch.int32, device="cuda")
    state = torch.zeros(64, dtype=torch.int32, device="cuda")
    R = 4
    out_s = torch.empty((R, 12, 40), device="cuda")
    out_r = torch.empty((R, 3), device="cuda")
    out_v = torch.empty((R, 16), device="cuda")
    for it in (1, 2, 3, 4):
        state[0] = it
        D.gather_batch(R, ds, order, state, out_s, out_r, out_v)
        torch.cuda.synchronize()
        sel = [int(order[((it - 1) * R + r) % 10]) for r in range(R)]
        assert torch.equal(out_s, ds.spectra[sel]), it
        assert torch.equal(out_r, ds.rx_pos[sel]) and torch.equal(out_v, ds.viewmatrices.reshape(-1, 16)[sel])


def test_trainer_draws_the_sample_order_inside_the_graph(tmp_path):
    import gsparc_trainer as TR
    root = str(tmp_path)
    make_dataset(root, n=9, H=90, W=360, seed=3)
    np.savetxt(str(tmp_path / "train_index.txt"), np.array([f"{i:05d}" for i in range(1, 10)]), fmt="%s")
    ds = D.SpectrumDataset(root, str(tmp_path / "train_index.txt"))
    sc = S.small_scene(5, 1200, R=1, walls=True)
    m = G.GaussianModel()
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    m.load_raw(t(sc.means) + torch.tensor([-3.0, -2.0, 0.0], device="cuda"), torch.log(t(sc.scales)),
               t(sc.rotations), torch.logit(t(sc.opacities).double()).float(), t(sc.mlp))
    o = G.OptimizationParams(densify_from_iter=10 ** 6)
    R = 4
    tr = TR.Trainer.from_dataset(m, o, ds, R, xyz_min=[-7.0, -4.0, 0.0], xyz_max=[7.0, 6.0, 4.0], iterations=20,
                                 seed=11)
    order = tr.data[1].cpu().numpy()
    assert np.array_equal(order, D.sample_order(9, 80, np.random.RandomState(11)))
    for it in range(1, 6):
        tr.step()
        torch.cuda.synchronize()
        sel = order[(it - 1) * R: it * R].tolist()
        assert torch.equal(tr.targets, ds.spectra[sel]), it
        assert torch.equal(tr.views, ds.viewmatrices.reshape(-1, 16)[sel]), it
        assert np.isfinite(float(tr.loss))
    assert tr.graph_key is not None            # iterations 2.. replayed the captured step
    tr.close()
