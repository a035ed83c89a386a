"""The training loop around the rasterizer (train.py:364-566) on the
library's kernels, with receivers batched.

One iteration: forward of R receivers, the (confidence-weighted) L1 + SSIM
loss, backward, the densification statistics, the schedule's densify /
opacity reset / pruning, the Adam step of the groups the reference steps in
that iteration, the centre clamp.  Iterations without a schedule event
replay one HIP graph that holds all of it; an iteration with an event runs
eagerly (it changes P or replaces parameters), and the next plain iteration
captures a new graph for the new P: it runs once eagerly, with the instance
capacity checked and grown, and is recorded for the replays that follow.

Batching: an iteration renders R receivers where the reference renders one.
The loss is the mean over the R spectra, so the screen-space gradient of
each (receiver, Gaussian) record is 1/R of the per-view gradient the
reference accumulates; the statistics take it scaled by R
(gsparc_densify_stats' grad_scale).  With R = 1 an iteration is the
reference's.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

import gsparc_hip as gh
import wireless_rasterizer as wr
from gsparc_gaussian import GaussianModel, OptimizationParams
from gsparc_train import CONFIDENCE_ALPHA, SSIM_C1, SSIM_C2

BOUNDS_MARGIN = 4.0                   # train.py:502
POST_PRUNE_INTERVAL = 10000           # train.py:540
POST_PRUNE_MIN_OPACITY = 0.001        # train.py:542
DENSIFY_MIN_OPACITY = 0.01            # train.py:526
SCREEN_SIZE_THRESHOLD = 20            # train.py:522


def scene_bounds(xyz_min, xyz_max, margin=BOUNDS_MARGIN):
    """(scene_min, scene_max, extent) as train.py:209-211, 502-504 compute them."""
    lo = np.asarray(xyz_min, np.float64)
    hi = np.asarray(xyz_max, np.float64)
    return (lo - margin).astype(np.float32), (hi + margin).astype(np.float32), float(np.linalg.norm(hi - lo))


class _Buffers:
    """Per-step device buffers for up to `pcap` Gaussians and R receivers."""

    def __init__(self, pcap, R, H, W, kcap, dev, conf: bool):
        L = gh.lib()
        u8 = lambda n: torch.empty(int(n), dtype=torch.uint8, device=dev)
        f32 = dict(dtype=torch.float32, device=dev)
        self.pcap, self.kcap = int(pcap), int(kcap)
        self.geom = u8(L.gsparc_geom_buffer_size(self.pcap, R, W, H))
        self.img = u8(L.gsparc_image_buffer_size(self.pcap, R, W, H))
        self.bin = u8(L.gsparc_binning_buffer_size(self.kcap))
        self.gradbuf = u8(L.gsparc_grad_buffer_size(self.pcap, R))
        self.radii = torch.empty(R * self.pcap, dtype=torch.int32, device=dev)
        self.dmeans2d = torch.empty(R * self.pcap * 3, **f32)
        self.out_real = torch.empty((R, H, W), **f32)
        self.out_imag = torch.empty((R, H, W), **f32)
        self.zero_imag = torch.zeros((R, H, W), **f32)
        self.dimg = torch.empty((R, H, W), **f32)
        self.loss = torch.empty(1, **f32)
        self.rec_loss = torch.empty(R, **f32)
        self.confidence = torch.empty(R, **f32)
        ws = L.gsparc_conf_loss_work_size(R, H, W) if conf else L.gsparc_loss_work_size(R, H, W)
        self.work = u8(ws)
        self.dcentre = torch.empty(3 * self.pcap, **f32)
        self.dscale = torch.empty(3 * self.pcap, **f32)
        self.drot = torch.empty(4 * self.pcap, **f32)
        self.dopac = torch.empty(self.pcap, **f32)
        self.dmlp = torch.empty(194 * self.pcap, **f32)

    def grow_bin(self, kcap, dev):
        self.kcap = int(kcap)
        self.bin = torch.empty(int(gh.lib().gsparc_binning_buffer_size(self.kcap)), dtype=torch.uint8, device=dev)


class Trainer:
    """Batched training of a GaussianModel against spectra of R receivers.

    views [R,4,4], rx_pos [R,3], tx_pos [3] and targets [R,H,W] are device
    tensors read in place at every iteration (change their contents between
    iterations to feed other receivers).  confidence: a packed confidence
    network (gsparc_train.ConfidenceMLP.packed()) for the confidence-weighted
    loss of train.py:418-436, or None for the plain loss (or, with
    first_iteration=None, the network a restored checkpoint brought).
    first_iteration: the iteration the first step runs; None continues from
    the model's training state (a checkpoint: its iteration + 1).  k_capacity: the
    first instance capacity; k_headroom: the capacity as a multiple of the K
    measured whenever an iteration runs eagerly (it grows when K outgrows
    it)."""

    def __init__(self, model: GaussianModel, optim: OptimizationParams, views, tx_pos, rx_pos, targets, H, W,
                 xyz_min, xyz_max, confidence=None, grad_mode="reference", k_headroom=1.5, use_graphs=True,
                 first_iteration=1, k_capacity=1 << 20):
        self.m = model
        self.o = optim
        self.dev = model.device
        self.views = gh.f32(views).reshape(-1, 16)
        self.rx = gh.f32(rx_pos).reshape(-1, 3)
        self.tx = gh.f32(tx_pos).reshape(3)
        self.targets = targets
        self.R, self.H, self.W = int(self.rx.shape[0]), int(H), int(W)
        if tuple(targets.shape) != (self.R, self.H, self.W):
            raise RuntimeError("targets must be [R,H,W]")
        self.scene_min, self.scene_max, self.extent = scene_bounds(xyz_min, xyz_max)
        self.mode = gh.GRAD_MODES[grad_mode] if isinstance(grad_mode, str) else int(grad_mode)
        self.k_headroom = float(k_headroom)
        self.use_graphs = bool(use_graphs)
        self.alpha = float(getattr(optim, "confidence_alpha", CONFIDENCE_ALPHA))
        self.lambda_ssim = float(optim.lambda1)
        model.training_setup(optim, confidence=confidence, scene_min=self.scene_min, scene_max=self.scene_max,
                             first_iteration=first_iteration)
        self.iteration = int(model.train_state[0].item())
        self.buf = None
        self.graph = C.c_void_p(None)
        self.graph_key = None
        self.stream = torch.cuda.Stream(device=self.dev)
        self.events = []            # (iteration, event, P before, P after)
        self.data = None            # (SpectrumDataset, sample order on the device) when fed by a dataset
        self.k_capacity0 = int(k_capacity)

    @classmethod
    def from_dataset(cls, model: GaussianModel, optim: OptimizationParams, dataset, R: int, xyz_min, xyz_max,
                     iterations: int | None = None, seed: int | None = None, **kw):
        """A trainer fed by a resident SpectrumDataset: iteration i trains on
        draws (i-1)R .. iR-1 of the reference's sample order
        (gsparc_dataset.sample_order, numpy's global RNG or RandomState(seed)),
        gathered on the device as part of the step."""
        import gsparc_dataset as D
        dev = model.device
        H, W = dataset.n_elevation, dataset.n_azimuth
        n = int(iterations if iterations is not None else optim.iterations) * int(R)
        order = D.sample_order(len(dataset), n, None if seed is None else np.random.RandomState(seed))
        views = torch.zeros((R, 16), dtype=torch.float32, device=dev)
        rx = torch.zeros((R, 3), dtype=torch.float32, device=dev)
        targets = torch.zeros((R, H, W), dtype=torch.float32, device=dev)
        tr = cls(model, optim, views, dataset.tx_pos, rx, targets, H, W, xyz_min, xyz_max, **kw)
        tr.data = (dataset, torch.from_numpy(order).to(dev))
        return tr

    # ---- the launch sequence
    def _gather(self, s):
        if self.data is not None:
            import gsparc_dataset as D
            ds, order = self.data
            D.gather_batch(self.R, ds, order, self.m.train_state, self.targets, self.rx, self.views, stream=s)

    def _ensure_buffers(self):
        P = self.m.P
        b = self.buf
        if b is None or b.pcap < P:
            cap = self.m.bank.capacity
            kcap = b.kcap if b is not None else self.k_capacity0
            self._drop_graph()
            self.buf = _Buffers(cap, self.R, self.H, self.W, kcap, self.dev, self.m.conf is not None)

    def _forward(self, s):
        L, m, b = gh.lib(), self.m, self.buf
        gh.check(L.gsparc_rasterize_forward(m.P, self.R, self.W, self.H, gh.ptr(m.get_centre), gh.ptr(m.get_scaling),
                                            gh.ptr(m.get_rotation), gh.ptr(m.get_opacity),
                                            gh.ptr(m.get_emission_mlps), None, gh.ptr(self.views), gh.ptr(self.tx),
                                            gh.ptr(self.rx), 1.0, gh.ptr(b.geom), gh.ptr(b.img), gh.ptr(b.bin),
                                            b.kcap, gh.ptr(b.radii), gh.ptr(b.out_real), gh.ptr(b.out_imag), 0, s),
                 "training step: forward")

    def _loss_backward(self, s):
        L, m, b = gh.lib(), self.m, self.buf
        R, H, W = self.R, self.H, self.W
        if m.conf is None:
            gh.check(L.gsparc_loss_l1_ssim(R, H, W, self.lambda_ssim, SSIM_C1, SSIM_C2, gh.ptr(b.out_real),
                                           gh.ptr(self.targets), gh.ptr(b.work), gh.ptr(b.loss), gh.ptr(b.dimg), s),
                     "training step: loss")
        else:
            gh.check(L.gsparc_loss_l1_ssim_conf(R, H, W, self.lambda_ssim, SSIM_C1, SSIM_C2, self.alpha,
                                                m.conf_hidden, gh.ptr(m.conf), gh.ptr(self.rx), gh.ptr(b.out_real),
                                                gh.ptr(self.targets), gh.ptr(b.work), gh.ptr(b.loss),
                                                gh.ptr(b.rec_loss), gh.ptr(b.confidence), gh.ptr(b.dimg),
                                                gh.ptr(m.conf_grad), s), "training step: confidence-weighted loss")
        gh.check(L.gsparc_rasterize_backward(self.mode, m.P, R, W, H, gh.ptr(m.get_centre), gh.ptr(m.get_scaling),
                                             gh.ptr(m.get_rotation), gh.ptr(m.get_emission_mlps), None,
                                             gh.ptr(self.views), gh.ptr(self.tx), gh.ptr(self.rx), 1.0,
                                             gh.ptr(b.radii), gh.ptr(b.geom), gh.ptr(b.img), gh.ptr(b.bin), b.kcap,
                                             gh.ptr(b.gradbuf), gh.ptr(b.dimg), gh.ptr(b.zero_imag),
                                             gh.ptr(b.dmeans2d), gh.ptr(b.dcentre), gh.ptr(b.dmlp), gh.ptr(b.dopac),
                                             None, gh.ptr(b.dscale), gh.ptr(b.drot), None, s),
                 "training step: backward")

    def _stats(self, s):
        b = self.buf
        self.m.add_densification_stats(b.dmeans2d[: self.R * self.m.P * 3], b.radii[: self.R * self.m.P],
                                       grad_scale=float(self.R), stream=s)

    def _adam(self, groups, s):
        b = self.buf
        self.m.optimizer_step(b.dcentre, b.dscale, b.drot, b.dopac, b.dmlp, groups=groups, stream=s)

    def _record_plain(self, s, stats):
        self._gather(s)
        self._forward(s)
        self._loss_backward(s)
        if stats:
            self._stats(s)
        self._adam(gh.GROUPS_ALL, s)

    def _check_capacity(self, s):
        """Forward, then K read back; when K times the headroom exceeds the
        instance capacity, a larger bin buffer replaces the old one and the
        forward runs again into it."""
        self._forward(s)
        K = wr.instance_count(self.m.P, self.R, self.H, self.W, self.buf.img)
        if K >= (1 << 32) - 1:
            raise RuntimeError("the tile lists would hold 2^32 - 1 or more instances")
        want = min(int(K * self.k_headroom) + 1024, (1 << 32) - 2)
        if want > self.buf.kcap:
            # a new bin buffer holds nothing yet and its layout depends on the
            # capacity: the forward runs again into it before any backward
            self._drop_graph()
            self.buf.grow_bin(want, self.dev)
            self._forward(s)
        return K

    def _drop_graph(self):
        if self.graph:
            gh.lib().gsparc_graph_destroy(self.graph)
        self.graph = C.c_void_p(None)
        self.graph_key = None

    # ---- the schedule (train.py:506-566)
    def _event(self, it):
        o = self.o
        if it < o.densify_until_iter:
            dens = it > o.densify_from_iter and it % o.densification_interval == 0
            reset = it % o.opacity_reset_interval == 0 and it > 0
            return dens, reset, False
        return False, False, it % POST_PRUNE_INTERVAL == 0

    def step(self):
        """Run one iteration (the number in self.iteration) and advance."""
        it = self.iteration
        dens, reset, post = self._event(it)
        stats = it < self.o.densify_until_iter
        self._ensure_buffers()
        cur = torch.cuda.current_stream(self.dev)
        if not (dens or reset or post):
            key = (self.m.cur, self.m.P, stats, self.buf.kcap, self.buf.pcap)
            if self.use_graphs and key == self.graph_key:
                gh.check(gh.lib().gsparc_graph_launch(self.graph, gh.stream_handle(self.dev)), "graph launch")
            else:
                self._capture_and_run(stats, cur)
        else:
            h = gh.stream_handle(self.dev)
            if dens or post:
                self._drop_graph()          # the arrays the graph reads are about to change
            self._gather(h)
            self._check_capacity(h)
            self._loss_backward(h)
            if stats:
                self._stats(h)
            groups = gh.GROUPS_ALL
            if dens:
                P0 = self.m.P
                self.m.densify_and_prune(self.o.densify_grad_threshold, DENSIFY_MIN_OPACITY, self.extent,
                                         SCREEN_SIZE_THRESHOLD if it > self.o.opacity_reset_interval else None,
                                         scene_min=self.scene_min, scene_max=self.scene_max)
                groups = gh.GROUP_CONFIDENCE
                self.events.append((it, "densify", P0, self.m.P))
            if reset:
                self.m.reset_opacity()
                groups &= ~gh.GROUP_OPACITY
                self.events.append((it, "reset_opacity", self.m.P, self.m.P))
            if post:
                P0 = self.m.P
                if self.m.prune(POST_PRUNE_MIN_OPACITY, self.extent, self.scene_min, self.scene_max) > 0:
                    groups = gh.GROUP_CONFIDENCE
                self.events.append((it, "prune", P0, self.m.P))
            self._ensure_buffers()
            self._adam(groups, h)
        self.iteration = it + 1

    def _capture_and_run(self, stats, cur):
        self._drop_graph()
        s = self.stream
        s.wait_stream(cur)
        with torch.cuda.stream(s):
            h = gh.stream_handle(self.dev)
            # this iteration, eagerly, with the instance capacity checked
            self._gather(h)
            self._check_capacity(h)
            self._loss_backward(h)
            if stats:
                self._stats(h)
            self._adam(gh.GROUPS_ALL, h)
            if self.use_graphs:
                L = gh.lib()
                gh.check(L.gsparc_graph_begin(h), "graph begin")
                try:
                    self._record_plain(h, stats)
                finally:
                    gh.check(L.gsparc_graph_end(h, C.byref(self.graph)), "graph end")
                self.graph_key = (self.m.cur, self.m.P, stats, self.buf.kcap, self.buf.pcap)
        cur.wait_stream(s)

    def train(self, iterations: int):
        """Run until self.iteration passes `iterations` (inclusive, as
        range(first_iter, iterations + 1) in train.py:364)."""
        while self.iteration <= iterations:
            self.step()

    # ---- results
    @property
    def loss(self) -> torch.Tensor:
        return self.buf.loss

    def overflowed(self) -> bool:
        """Whether the last forward needed more instances than the capacity
        (synchronises)."""
        return wr.instance_count(self.m.P, self.R, self.H, self.W, self.buf.img) > self.buf.kcap

    def nan_centre(self) -> bool:
        """The NaN check after the step (train.py:565-566), sticky
        (synchronises)."""
        return bool(int(self.m.train_state[7].item()))

    def close(self):
        self._drop_graph()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
