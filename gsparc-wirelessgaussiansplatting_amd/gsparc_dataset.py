"""The RF spectrum dataset (preprocess_data.py:17-243) resident in device
memory, with its view matrices built on the device and the training order
of its samples.

Layout of a dataset folder (the reference's RFID spectrum format):
  rx_pos.csv          receiver positions, one row per spectrum (a header
                      row, then x, y, z); spectrum "00042" uses row 42
  gateway_info.yml    gateway1: position [3], orientation [x, y, z, w]
  spectrum/*.npy      one [elevation, azimuth] float spectrum per receiver
  train_index.txt, test_index.txt   spectrum names of each split
SpectrumDataset loads a split into device arrays (spectra [N,H,W], rx_pos
[N,3], viewmatrices [N,16]) once; gsparc_gather_batch picks each iteration's
batch from them on the device.  Files are read with loaders that execute
nothing from them: numpy.load without pickles, pandas CSV, yaml.safe_load.
"""
from __future__ import annotations

import ctypes as C
import os
import random

import numpy as np
import torch

import gsparc_hip as gh

TX_POS = (0.00, 0.50, 3.00)      # the fixed transmitter of Spectrum_dataset.load_data (preprocess_data.py:151)


def split_dataset(datadir, ratio=0.8, dataset_type="rfid"):
    """split_dataset (preprocess_data.py:17-40): shuffle the sample names with
    Python's `random` and write train_index.txt / test_index.txt.  For 'rfid'
    the names come from spectrum/*.png as in the reference; a folder with
    only .npy spectra (what Spectrum_dataset reads) uses those names."""
    if dataset_type == "rfid":
        d = os.path.join(datadir, "spectrum")
        names = sorted(f for f in os.listdir(d) if f.endswith(".png"))
        if not names:
            names = sorted(f for f in os.listdir(d) if f.endswith(".npy"))
        index = [x.split(".")[0] for x in names]
        random.shuffle(index)
    elif dataset_type == "ble":
        import pandas as pd
        index = list(pd.read_csv(os.path.join(datadir, "gateway_rssi.csv")).index.values)
        random.shuffle(index)
    elif dataset_type == "mimo":
        index = list(range(np.load(os.path.join(datadir, "csidata.npy"), mmap_mode="r").shape[0]))
        random.shuffle(index)
    else:
        raise ValueError(f"unknown dataset type {dataset_type}")
    n = int(len(index) * ratio)
    np.savetxt(os.path.join(datadir, "train_index.txt"), np.array(index[:n]), fmt="%s")
    np.savetxt(os.path.join(datadir, "test_index.txt"), np.array(index[n:]), fmt="%s")


def read_gateway(datadir):
    import yaml
    with open(os.path.join(datadir, "gateway_info.yml")) as f:
        info = yaml.safe_load(f)
    g = info["gateway1"]
    return np.asarray(g["position"], np.float64), np.asarray(g["orientation"], np.float64)


def view_matrices(rx_pos: torch.Tensor, quat) -> torch.Tensor:
    """build_view_matrix for every row of rx_pos [N,3] (device) under the
    gateway quaternion (x, y, z, w): [N,4,4] column-major, on the device."""
    rx = gh.f32(rx_pos).reshape(-1, 3)
    gh.require_cuda(rx)
    N = int(rx.shape[0])
    out = torch.empty((N, 4, 4), dtype=torch.float32, device=rx.device)
    q = (C.c_double * 4)(*[float(v) for v in quat])
    gh.check(gh.lib().gsparc_view_matrices(N, q, gh.ptr(rx), gh.ptr(out), gh.stream_handle(rx.device)),
             "gsparc_view_matrices")
    return out


class WirelessData:
    """One sample (preprocess_data.py:70-87): views into the resident arrays."""

    def __init__(self, spectrum, tx_pos, rx_pos, view_matrix):
        self.spectrum_real = spectrum
        self.spectrum_imag = torch.empty(0, device=spectrum.device)
        self.tx_pos = tx_pos
        self.rx_pos = rx_pos
        self.spectrum_elevation = int(spectrum.shape[0])
        self.spectrum_azimuth = int(spectrum.shape[1])
        self.view_matrix = view_matrix

    def to(self, device):
        for k in ("spectrum_real", "spectrum_imag", "tx_pos", "rx_pos", "view_matrix"):
            setattr(self, k, getattr(self, k).to(device))
        return self


class SpectrumDataset:
    """Spectrum_dataset (preprocess_data.py:114-243) with its split resident
    on the device: spectra [N,H,W], rx_pos [N,3], viewmatrices [N,4,4], tx_pos
    [3]; dataset_index the split's names; n_elevation, n_azimuth;
    ray_directions [H*W,3] and the gateway's rx_pos / view_matrix of
    gen_rays_spectrum (preprocess_data.py:204-243)."""

    def __init__(self, datadir, indexdir, scale_worldsize=1, device="cuda"):
        import pandas as pd
        self.datadir = datadir
        self.device = torch.device(device)
        sdir = os.path.join(datadir, "spectrum")
        names = sorted(f for f in os.listdir(sdir) if f.endswith(".npy"))
        if not names:
            raise FileNotFoundError(f"no .npy spectra in {sdir}")
        ex = np.load(os.path.join(sdir, names[0]), allow_pickle=False)
        self.n_elevation, self.n_azimuth = ex.shape
        self.rays_per_spectrum = self.n_elevation * self.n_azimuth
        self.dataset_index = np.loadtxt(indexdir, dtype=str, ndmin=1)
        rx_all = torch.tensor(pd.read_csv(os.path.join(datadir, "rx_pos.csv")).values, dtype=torch.float32)
        self.gateway_pos, self.gateway_quat = read_gateway(datadir)
        N = len(self.dataset_index)
        H, W = self.n_elevation, self.n_azimuth
        self.spectra = torch.empty((N, H, W), dtype=torch.float32, device=self.device)
        rows = []
        for i, idx in enumerate(self.dataset_index):
            a = np.load(os.path.join(sdir, idx + ".npy"), allow_pickle=False)
            if a.shape != (H, W):
                raise RuntimeError(f"spectrum {idx} is {a.shape}, expected {(H, W)}")
            self.spectra[i].copy_(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)))
            rows.append(int(idx) - 1)
        self.rx_pos = rx_all[torch.tensor(rows, dtype=torch.long)].to(self.device).contiguous()
        self.tx_pos = torch.tensor(TX_POS, dtype=torch.float32, device=self.device)
        self.viewmatrices = view_matrices(self.rx_pos, self.gateway_quat)
        self.ray_directions = self._ray_directions()

    def _ray_directions(self):
        """gen_rays_spectrum's ray directions in world coordinates [H*W, 3]."""
        from scipy.spatial.transform import Rotation
        H, W = self.n_elevation, self.n_azimuth
        az = torch.linspace(1, 360, W) / 180 * np.pi
        az = az % (2 * np.pi)
        az = torch.where(az > np.pi, az - 2 * np.pi, az)
        el = torch.linspace(1, 90, H) / 180 * np.pi
        az = torch.tile(az, (H,))
        el = torch.repeat_interleave(el, W)
        d = torch.stack([torch.cos(el) * torch.sin(az), torch.sin(el), torch.cos(el) * torch.cos(az)], dim=0)
        R = torch.from_numpy(Rotation.from_quat(self.gateway_quat).as_matrix()).float()
        return (R @ d).T.contiguous()

    def __len__(self):
        return len(self.dataset_index)

    def __getitem__(self, i):
        return WirelessData(self.spectra[i], self.tx_pos, self.rx_pos[i], self.viewmatrices[i])

    def get_data(self):
        return [self[i] for i in range(len(self))]


def sample_order(n: int, count: int, rng=None) -> np.ndarray:
    """The order in which train.py:364-370 draws samples: the first pass pops
    the dataset list from its end (n-1, ..., 0); every later pass shuffles a
    fresh copy with np.random.shuffle and pops from its end.  Returns the
    first `count` draws.  rng: a numpy RandomState (default the global one,
    as the reference uses)."""
    rs = np.random if rng is None else rng
    out = []
    cur = list(range(n))
    while len(out) < count:
        if not cur:
            cur = list(range(n))
            rs.shuffle(cur)
        out.append(cur.pop())
    return np.asarray(out, np.int32)


def gather_batch(R, dataset: SpectrumDataset, order: torch.Tensor, train_state: torch.Tensor, out_spectra,
                 out_rx, out_views, stream=None):
    """gsparc_gather_batch on the dataset's resident arrays."""
    N = len(dataset)
    npix = dataset.n_elevation * dataset.n_azimuth
    s = stream if stream is not None else gh.stream_handle(dataset.device)
    gh.check(gh.lib().gsparc_gather_batch(int(R), N, npix, gh.ptr(order), int(order.numel()), gh.ptr(train_state),
                                          gh.ptr(dataset.spectra), gh.ptr(dataset.rx_pos),
                                          gh.ptr(dataset.viewmatrices), gh.ptr(out_spectra), gh.ptr(out_rx),
                                          gh.ptr(out_views), s), "gsparc_gather_batch")
