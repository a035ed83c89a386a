"""Training-step helpers over the C ABI: the reference's loss as one fused
call, and a whole forward + loss + backward step captured in a HIP graph.

The reference trains one receiver at a time (train.py:382-440): render, then
loss = (1 - lambda1) * L1 + lambda1 * (1 - SSIM) with lambda1 = 0.2
(optimizer.py:8, train.py:418-427), then loss.backward().  Here:

* `l1_ssim_loss(img, target)` computes that loss and its gradient in three
  kernels (SSIM forward with per-workgroup loss partials, a fixed-order
  reduction, SSIM backward with the L1 term fused), for R spectra at once.
* `GraphedTrainingStep` records the sync-free forward (tile lists sized by a
  capacity, K kept on the device), the loss and the backward once into a HIP
  graph owned by the library; `replay()` then re-runs the step on whatever
  the parameter tensors hold, with no host synchronisation and no allocation.
* `confidence_l1_ssim_loss` / `confidence_loss` add the confidence weighting
  of train.py:429-436 (C_r from the receiver-position ConfidenceMLP) inside
  the same fused loss.
"""
from __future__ import annotations

import ctypes as C

import torch

import gsparc_hip as gh
import wireless_rasterizer as wr

SSIM_C1 = 0.01 ** 2
SSIM_C2 = 0.03 ** 2
LAMBDA_SSIM = 0.2          # lambda1 of optimizer.py:8


def _work(planes, H, W, dev):
    n = gh.lib().gsparc_loss_work_size(int(planes), int(H), int(W))
    return torch.empty(n, dtype=torch.uint8, device=dev)


def l1_ssim_loss(img: torch.Tensor, target: torch.Tensor, lambda_ssim: float = LAMBDA_SSIM, want_grad: bool = True):
    """img, target [R,H,W] (or [H,W]) float32 on the GPU.  Returns (loss [1],
    dL/dimg like img or None): loss = (1 - lambda_ssim) * mean|img - target| +
    lambda_ssim * (1 - mean SSIM), the mean over all R*H*W pixels."""
    x = gh.f32(img)
    y = gh.f32(target)
    if x.shape != y.shape or x.dim() not in (2, 3):
        raise RuntimeError("img and target must both be [H,W] or [R,H,W]")
    gh.require_cuda(x, y)
    R = 1 if x.dim() == 2 else x.size(0)
    H, W = x.shape[-2], x.shape[-1]
    loss = torch.empty(1, dtype=torch.float32, device=x.device)
    grad = torch.empty_like(x) if want_grad else None
    work = _work(R, H, W, x.device)
    gh.check(gh.lib().gsparc_loss_l1_ssim(R, H, W, float(lambda_ssim), SSIM_C1, SSIM_C2, gh.ptr(x), gh.ptr(y),
                                          gh.ptr(work), gh.ptr(loss), gh.ptr(grad), gh.stream_handle(x.device)),
             "l1 + ssim loss")
    return loss, grad


class L1SSIMLoss(torch.autograd.Function):
    """Autograd form of l1_ssim_loss (gradient to img only, as fused_ssim)."""

    @staticmethod
    def forward(ctx, img, target, lambda_ssim=LAMBDA_SSIM):
        loss, grad = l1_ssim_loss(img, target, lambda_ssim, want_grad=img.requires_grad)
        ctx.save_for_backward(grad if grad is not None else torch.empty(0, device=img.device))
        return loss.reshape(())

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return grad * g, None, None


CONFIDENCE_ALPHA = 0.3     # confidence_alpha of optimizer.py


class ConfidenceMLP(torch.nn.Module):
    """The receiver-position confidence network of gaussian.py:575-609:
    Linear(3 -> hidden) -> ReLU -> Linear(hidden -> 1), C = 1 + exp(raw) > 1.
    Same parameter names, shapes and initialisation (Xavier-uniform weights,
    zero biases), so its state dict is the reference's.  forward() is the
    plain torch form; the training loss evaluates the network inside the
    fused loss kernel from packed()."""

    def __init__(self, hidden_size: int = 32, device=None):
        super().__init__()
        self.fc1 = torch.nn.Linear(3, hidden_size)
        self.relu = torch.nn.ReLU()
        self.fc2 = torch.nn.Linear(hidden_size, 1)
        torch.nn.init.zeros_(self.fc2.bias)
        torch.nn.init.xavier_uniform_(self.fc1.weight)
        torch.nn.init.zeros_(self.fc1.bias)
        torch.nn.init.xavier_uniform_(self.fc2.weight)
        if device is not None:
            self.to(device)

    @property
    def hidden(self) -> int:
        return self.fc1.out_features

    def forward(self, rx_pos):
        if rx_pos.dim() == 1:
            rx_pos = rx_pos.unsqueeze(0)
        return (1.0 + torch.exp(self.fc2(self.relu(self.fc1(rx_pos))))).squeeze(-1)

    def packed(self) -> torch.Tensor:
        """[fc1.weight | fc1.bias | fc2.weight | fc2.bias], differentiable."""
        return torch.cat([self.fc1.weight.reshape(-1), self.fc1.bias, self.fc2.weight.reshape(-1), self.fc2.bias])


def confidence_l1_ssim_loss(img: torch.Tensor, target: torch.Tensor, rx_pos: torch.Tensor, conf_mlp=None,
                            hidden: int = 32, alpha: float = CONFIDENCE_ALPHA, lambda_ssim: float = LAMBDA_SSIM,
                            want_grad: bool = True):
    """The reference's confidence-weighted loss (train.py:418-436) of R
    spectra img, target [R,H,W] seen from rx_pos [R,3]:
    mean_r (C_r L_r - alpha log C_r), C_r from the packed confidence network
    conf_mlp ([5 hidden + 1] float32, or None: C_r = 1).  Returns (loss [1],
    dL/dimg or None, dL/dconf_mlp or None, L_r [R], C_r [R])."""
    x = gh.f32(img)
    y = gh.f32(target)
    if x.dim() == 2:
        x, y = x.unsqueeze(0), y.unsqueeze(0)
    if x.shape != y.shape or x.dim() != 3:
        raise RuntimeError("img and target must both be [H,W] or [R,H,W]")
    R, H, W = x.shape
    rx = gh.f32(rx_pos).reshape(-1, 3)
    if rx.size(0) != R:
        raise RuntimeError("one receiver position per spectrum is required")
    mlp = None if conf_mlp is None else gh.f32(conf_mlp).reshape(-1)
    if mlp is not None and mlp.numel() != 5 * hidden + 1:
        raise RuntimeError(f"the packed confidence network holds 5 * hidden + 1 = {5 * hidden + 1} floats")
    gh.require_cuda(x, y, rx, mlp)
    dev = x.device
    L = gh.lib()
    loss = torch.empty(1, dtype=torch.float32, device=dev)
    rec = torch.empty(R, dtype=torch.float32, device=dev)
    conf = torch.empty(R, dtype=torch.float32, device=dev)
    dimg = torch.empty_like(x) if want_grad else None
    dmlp = torch.empty_like(mlp) if (want_grad and mlp is not None) else None
    work = torch.empty(L.gsparc_conf_loss_work_size(R, H, W), dtype=torch.uint8, device=dev)
    gh.check(L.gsparc_loss_l1_ssim_conf(R, H, W, float(lambda_ssim), SSIM_C1, SSIM_C2, float(alpha), int(hidden),
                                        gh.ptr(mlp), gh.ptr(rx), gh.ptr(x), gh.ptr(y), gh.ptr(work), gh.ptr(loss),
                                        gh.ptr(rec), gh.ptr(conf), gh.ptr(dimg), gh.ptr(dmlp),
                                        gh.stream_handle(dev)), "confidence-weighted loss")
    if dimg is not None and img.dim() == 2:
        dimg = dimg[0]
    return loss, dimg, dmlp, rec, conf


class ConfidenceL1SSIMLoss(torch.autograd.Function):
    """Autograd form of confidence_l1_ssim_loss: gradients to img and to the
    packed confidence network."""

    @staticmethod
    def forward(ctx, img, target, rx_pos, conf_packed, hidden=32, alpha=CONFIDENCE_ALPHA,
                lambda_ssim=LAMBDA_SSIM):
        loss, dimg, dmlp, rec, conf = confidence_l1_ssim_loss(img, target, rx_pos, conf_packed, hidden, alpha,
                                                              lambda_ssim, want_grad=True)
        ctx.save_for_backward(dimg, dmlp if dmlp is not None else torch.empty(0, device=img.device))
        ctx.has_mlp = dmlp is not None
        return loss.reshape(())

    @staticmethod
    def backward(ctx, g):
        dimg, dmlp = ctx.saved_tensors
        return dimg * g, None, None, (dmlp * g if ctx.has_mlp else None), None, None, None


def confidence_loss(img, target, rx_pos, conf_mlp: ConfidenceMLP = None, alpha: float = CONFIDENCE_ALPHA,
                    lambda_ssim: float = LAMBDA_SSIM):
    """loss = mean_r (C_r L_r - alpha log C_r) with autograd to img and to
    the ConfidenceMLP's parameters (train.py:418-436)."""
    packed = None if conf_mlp is None else conf_mlp.packed()
    hidden = 32 if conf_mlp is None else conf_mlp.hidden
    return ConfidenceL1SSIMLoss.apply(img, target, rx_pos, packed, hidden, alpha, lambda_ssim)


def csi_loss(out: torch.Tensor, target: torch.Tensor, lambda_ssim: float = LAMBDA_SSIM, want_grad: bool = True):
    """Loss of CSI maps out, target [R,H,W,52,2]: (1 - lambda) * mean |out -
    target| over all real and imaginary parts + lambda * (1 - mean SSIM of the
    52 magnitude maps per receiver).  Returns (loss [1], dL/dout or None)."""
    x = gh.f32(out)
    y = gh.f32(target)
    if x.shape != y.shape or x.dim() != 5 or x.shape[-2:] != (52, 2):
        raise RuntimeError("out and target must both be [R,H,W,52,2]")
    gh.require_cuda(x, y)
    R, H, W = x.shape[:3]
    L = gh.lib()
    loss = torch.empty(1, dtype=torch.float32, device=x.device)
    grad = torch.empty_like(x) if want_grad else None
    work = torch.empty(L.gsparc_csi_loss_work_size(R, H, W), dtype=torch.uint8, device=x.device)
    gh.check(L.gsparc_csi_loss(R, H, W, float(lambda_ssim), SSIM_C1, SSIM_C2, gh.ptr(x), gh.ptr(y), gh.ptr(work),
                               gh.ptr(loss), gh.ptr(grad), gh.stream_handle(x.device)), "csi loss")
    return loss, grad


class CSILoss(torch.autograd.Function):
    """Autograd form of csi_loss (gradient to out only)."""

    @staticmethod
    def forward(ctx, out, target, lambda_ssim=LAMBDA_SSIM):
        loss, grad = csi_loss(out, target, lambda_ssim, want_grad=out.requires_grad)
        ctx.save_for_backward(grad if grad is not None else torch.empty(0, device=out.device))
        return loss.reshape(())

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return grad * g, None, None


def _require_param(t, name):
    if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise RuntimeError(f"{name} must be a contiguous float32 GPU tensor (the graph reads it in place)")
    return t


class _GraphedStep:
    """Capture of a step's launch sequence into a library-owned HIP graph
    (gsparc_graph_begin / _end) and its replay.  Subclasses provide
    _record(stream) and instance_count()."""

    def _capture(self, dev):
        L = gh.lib()
        self.graph = C.c_void_p(None)
        self.stream = torch.cuda.Stream(device=dev)
        cur = torch.cuda.current_stream(dev)
        self.stream.wait_stream(cur)
        with torch.cuda.stream(self.stream):
            h = self.stream.cuda_stream
            self._record(h)                       # one eager run: checks every argument
            gh.check(L.gsparc_graph_begin(h), "graph begin")
            try:
                self._record(h)
            finally:
                gh.check(L.gsparc_graph_end(h, C.byref(self.graph)), "graph end")
        cur.wait_stream(self.stream)

    def replay(self):
        """Launch the captured step on the current stream."""
        gh.check(gh.lib().gsparc_graph_launch(self.graph, gh.stream_handle(self.target.device)), "graph launch")

    def run_eager(self):
        """The same launch sequence without the graph, on the current stream."""
        self._record(gh.stream_handle(self.target.device))

    def overflowed(self) -> bool:
        return self.instance_count() > self.capacity

    def close(self):
        if getattr(self, "graph", None):
            gh.lib().gsparc_graph_destroy(self.graph)
            self.graph = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GraphedTrainingStep(_GraphedStep):
    """Forward, loss and backward of one batched training step, captured once
    and replayed.

    The parameter tensors (means [P,3], scales [P,3], rotations [P,4],
    opacities [P,1], emission_mlps [P*194]) and the target [R,H,W] are read in
    place at every replay: update them in place (an optimizer step) between
    replays.  After replay(): `loss` [1], `out_real` / `out_imag` [R,H,W],
    `radii` [R,P] and the gradients `grads` (means3D, emission_mlps, opacities,
    scales, rotations) hold the step's results.  `capacity` bounds the tile
    lists; `overflowed()` tells whether the last replay needed more (then its
    spectra and gradients are not valid)."""

    def __init__(self, means, scales, rotations, opacities, emission_mlps, viewmatrices, tx_pos, rx_pos, target,
                 image_height, image_width, capacity, scale_modifier=1.0, grad_mode="reference",
                 lambda_ssim=LAMBDA_SSIM, confidence=None, confidence_hidden=32, confidence_alpha=CONFIDENCE_ALPHA,
                 deterministic=False):
        """confidence: None (the plain loss) or the packed ConfidenceMLP
        [5 hidden + 1] as a live tensor (ConfidenceMLP.packed() copied into
        it; update it in place): the loss is then the confidence-weighted one
        and grads["confidence"] its gradient, `rec_loss` / `conf` the per
        spectrum L_r and C_r.  deterministic: fixed-order gradient sums
        (bit-reproducible replays) instead of float atomics."""
        self.conf = None if confidence is None else _require_param(confidence, "confidence")
        self.conf_hidden = int(confidence_hidden)
        self.conf_alpha = float(confidence_alpha)
        self.params = dict(means=_require_param(means, "means"), scales=_require_param(scales, "scales"),
                           rotations=_require_param(rotations, "rotations"),
                           opacities=_require_param(opacities, "opacities"),
                           emission_mlps=_require_param(emission_mlps, "emission_mlps"))
        if self.params["emission_mlps"].data_ptr() % 16:
            raise RuntimeError("emission_mlps must be 16-byte aligned")
        dev = means.device
        P = int(means.size(0))
        self.views = gh.f32(viewmatrices).reshape(-1, 16)
        self.rx = gh.f32(rx_pos).reshape(-1, 3)
        self.tx = gh.f32(tx_pos).reshape(3)
        R = int(self.rx.size(0))
        H, W = int(image_height), int(image_width)
        self.target = _require_param(target, "target")
        if tuple(self.target.shape) != (R, H, W):
            raise RuntimeError("target must be [R,H,W]")
        if self.params["emission_mlps"].numel() != P * 194:
            raise RuntimeError("emission_mlps must hold 194 floats per Gaussian")
        self.P, self.R, self.H, self.W = P, R, H, W
        self.capacity = int(capacity)
        self.mode = gh.grad_mode_code(grad_mode, deterministic)
        self.scale_modifier = float(scale_modifier)
        self.lambda_ssim = float(lambda_ssim)
        L = gh.lib()
        f32 = dict(dtype=torch.float32, device=dev)
        u8 = lambda n: torch.empty(int(n), dtype=torch.uint8, device=dev)
        self.geom = u8(L.gsparc_geom_buffer_size(P, R, W, H))
        self.img = u8(L.gsparc_image_buffer_size(P, R, W, H))
        self.bin = u8(L.gsparc_binning_buffer_size(self.capacity))
        gsize = L.gsparc_grad_buffer_size(P, R)
        self.gradbuf = u8(L.gsparc_deterministic_grad_size(gsize, self.capacity) if deterministic else gsize)
        self.work = _work(R, H, W, dev)
        self.radii = torch.empty((R, P), dtype=torch.int32, device=dev)
        self.out_real = torch.empty((R, H, W), **f32)
        self.out_imag = torch.empty((R, H, W), **f32)
        self.loss = torch.empty(1, **f32)
        self.dimg = torch.empty((R, H, W), **f32)
        self.zero_imag = torch.zeros((R, H, W), **f32)
        self.grads = dict(means=torch.empty((P, 3), **f32), emission_mlps=torch.empty(P * 194, **f32),
                          opacities=torch.empty((P, 1), **f32), scales=torch.empty((P, 3), **f32),
                          rotations=torch.empty((P, 4), **f32))
        if self.conf is not None:
            if self.conf.numel() != 5 * self.conf_hidden + 1:
                raise RuntimeError("confidence must hold 5 * confidence_hidden + 1 floats")
            self.work = torch.empty(L.gsparc_conf_loss_work_size(R, H, W), dtype=torch.uint8, device=dev)
            self.grads["confidence"] = torch.empty_like(self.conf)
            self.rec_loss = torch.empty(R, **f32)
            self.confidence = torch.empty(R, **f32)
        self._capture(dev)

    def _loss(self, stream):
        L = gh.lib()
        R, W, H = self.R, self.W, self.H
        if self.conf is None:
            gh.check(L.gsparc_loss_l1_ssim(R, H, W, self.lambda_ssim, SSIM_C1, SSIM_C2, gh.ptr(self.out_real),
                                           gh.ptr(self.target), gh.ptr(self.work), gh.ptr(self.loss),
                                           gh.ptr(self.dimg), stream), "graphed step: loss")
        else:
            gh.check(L.gsparc_loss_l1_ssim_conf(R, H, W, self.lambda_ssim, SSIM_C1, SSIM_C2, self.conf_alpha,
                                                self.conf_hidden, gh.ptr(self.conf), gh.ptr(self.rx),
                                                gh.ptr(self.out_real), gh.ptr(self.target), gh.ptr(self.work),
                                                gh.ptr(self.loss), gh.ptr(self.rec_loss), gh.ptr(self.confidence),
                                                gh.ptr(self.dimg), gh.ptr(self.grads["confidence"]), stream),
                     "graphed step: confidence-weighted loss")

    def _record(self, stream):
        L = gh.lib()
        p = self.params
        P, R, W, H = self.P, self.R, self.W, self.H
        gh.check(L.gsparc_rasterize_forward(P, R, W, H, gh.ptr(p["means"]), gh.ptr(p["scales"]),
                                            gh.ptr(p["rotations"]), gh.ptr(p["opacities"]),
                                            gh.ptr(p["emission_mlps"]), None, gh.ptr(self.views), gh.ptr(self.tx),
                                            gh.ptr(self.rx), self.scale_modifier, gh.ptr(self.geom),
                                            gh.ptr(self.img), gh.ptr(self.bin), self.capacity, gh.ptr(self.radii),
                                            gh.ptr(self.out_real), gh.ptr(self.out_imag), 0, stream),
                 "graphed step: forward")
        self._loss(stream)
        g = self.grads
        gh.check(L.gsparc_rasterize_backward(self.mode, P, R, W, H, gh.ptr(p["means"]), gh.ptr(p["scales"]),
                                             gh.ptr(p["rotations"]), gh.ptr(p["emission_mlps"]), None,
                                             gh.ptr(self.views), gh.ptr(self.tx), gh.ptr(self.rx),
                                             self.scale_modifier, gh.ptr(self.radii), gh.ptr(self.geom),
                                             gh.ptr(self.img), gh.ptr(self.bin), self.capacity, gh.ptr(self.gradbuf),
                                             gh.ptr(self.dimg), gh.ptr(self.zero_imag), None, gh.ptr(g["means"]),
                                             gh.ptr(g["emission_mlps"]), gh.ptr(g["opacities"]), None,
                                             gh.ptr(g["scales"]), gh.ptr(g["rotations"]), None, stream),
                 "graphed step: backward")

    def instance_count(self) -> int:
        """K of the last step (synchronises the current stream)."""
        return wr.instance_count(self.P, self.R, self.H, self.W, self.img)


class GraphedCsiStep(_GraphedStep):
    """GraphedTrainingStep for the multi-frequency maps: the CSI forward of R
    receivers with a tile-list capacity, the CSI loss (L1 on the complex map
    plus SSIM on the 52 magnitude maps) and the exact adjoint, captured once
    and replayed.  Parameters (means [P,3], scales [P,3], rotations [P,4],
    opacities [P,1], emission_mlps [P*194]) and target [R,H,W,52,2] are read
    in place at every replay; after replay() `loss` [1], `out` [R,H,W,52,2],
    `radii` [R,P] and `grads` hold the step's results.  No counterpart in the
    reference, which renders one channel (config.h:15)."""

    def __init__(self, means, scales, rotations, opacities, emission_mlps, viewmatrices, tx_pos, rx_pos, target,
                 image_height, image_width, capacity, scale_modifier=1.0, lambda_ssim=LAMBDA_SSIM,
                 f0=None, df=None):
        self.params = dict(means=_require_param(means, "means"), scales=_require_param(scales, "scales"),
                           rotations=_require_param(rotations, "rotations"),
                           opacities=_require_param(opacities, "opacities"),
                           emission_mlps=_require_param(emission_mlps, "emission_mlps"))
        if self.params["emission_mlps"].data_ptr() % 16:
            raise RuntimeError("emission_mlps must be 16-byte aligned")
        dev = means.device
        P = int(means.size(0))
        self.views = gh.f32(viewmatrices).reshape(-1, 16)
        self.rx = gh.f32(rx_pos).reshape(-1, 3)
        self.tx = gh.f32(tx_pos).reshape(3)
        R = int(self.rx.size(0))
        H, W = int(image_height), int(image_width)
        NC = wr.CSI_NC
        self.target = _require_param(target, "target")
        if tuple(self.target.shape) != (R, H, W, NC, 2):
            raise RuntimeError(f"target must be [R,H,W,{NC},2]")
        if self.views.size(0) != R:
            raise RuntimeError("one view matrix per receiver position is required")
        if self.params["emission_mlps"].numel() != P * 194:
            raise RuntimeError("emission_mlps must hold 194 floats per Gaussian")
        self.P, self.R, self.H, self.W = P, R, H, W
        self.f0 = float(wr.CSI_F0 if f0 is None else f0)
        self.df = float(wr.CSI_DF if df is None else df)
        self.capacity = int(capacity)
        self.scale_modifier = float(scale_modifier)
        self.lambda_ssim = float(lambda_ssim)
        L = gh.lib()
        f32 = dict(dtype=torch.float32, device=dev)
        u8 = lambda n: torch.empty(int(n), dtype=torch.uint8, device=dev)
        self.geom = u8(L.gsparc_geom_buffer_size(P, R, W, H))
        self.img = u8(L.gsparc_image_buffer_size(P, R, W, H))
        self.bin = u8(L.gsparc_binning_buffer_size(self.capacity))
        self.gradbuf = u8(L.gsparc_grad_buffer_size(P, R))
        self.work = u8(L.gsparc_csi_loss_work_size(R, H, W))
        self.radii = torch.empty((R, P), dtype=torch.int32, device=dev)
        self.out = torch.empty((R, H, W, NC, 2), **f32)
        self.loss = torch.empty(1, **f32)
        self.dout = torch.empty((R, H, W, NC, 2), **f32)
        self.grads = dict(means=torch.empty((P, 3), **f32), emission_mlps=torch.empty(P * 194, **f32),
                          opacities=torch.empty((P, 1), **f32), scales=torch.empty((P, 3), **f32),
                          rotations=torch.empty((P, 4), **f32))
        self._capture(dev)

    def _record(self, stream):
        L = gh.lib()
        p, g = self.params, self.grads
        P, R, W, H = self.P, self.R, self.W, self.H
        gh.check(L.gsparc_rasterize_forward_csi(P, R, W, H, self.f0, self.df, gh.ptr(p["means"]),
                                                gh.ptr(p["scales"]), gh.ptr(p["rotations"]),
                                                gh.ptr(p["opacities"]), gh.ptr(p["emission_mlps"]), None,
                                                gh.ptr(self.views), gh.ptr(self.tx), gh.ptr(self.rx),
                                                self.scale_modifier, gh.ptr(self.geom), gh.ptr(self.img),
                                                gh.ptr(self.bin), self.capacity, gh.ptr(self.radii),
                                                gh.ptr(self.out), stream), "graphed csi step: forward")
        gh.check(L.gsparc_csi_loss(R, H, W, self.lambda_ssim, SSIM_C1, SSIM_C2, gh.ptr(self.out),
                                   gh.ptr(self.target), gh.ptr(self.work), gh.ptr(self.loss), gh.ptr(self.dout),
                                   stream), "graphed csi step: loss")
        gh.check(L.gsparc_rasterize_backward_csi(
            P, R, W, H, self.f0, self.df, gh.ptr(p["means"]), gh.ptr(p["scales"]), gh.ptr(p["rotations"]),
            gh.ptr(p["emission_mlps"]), None, gh.ptr(self.views), gh.ptr(self.tx), gh.ptr(self.rx),
            self.scale_modifier, gh.ptr(self.radii), gh.ptr(self.geom), gh.ptr(self.img), gh.ptr(self.bin),
            self.capacity, gh.ptr(self.gradbuf), gh.ptr(self.dout), None, gh.ptr(g["means"]),
            gh.ptr(g["emission_mlps"]), gh.ptr(g["opacities"]), None, gh.ptr(g["scales"]), gh.ptr(g["rotations"]),
            None, stream), "graphed csi step: backward")

    def instance_count(self) -> int:
        """K of the last step (synchronises the current stream)."""
        return wr.instance_count(self.P, self.R, self.H, self.W, self.img)


class GraphedSharedMlpStep(_GraphedStep):
    """The single-receiver training configuration's step with the shared
    emission MLP (gsparc_emission.SharedEmissionMLP), captured once and
    replayed: the MLP forward for every (receiver, Gaussian) pair, the
    rasterizer forward with the emission input and a tile-list capacity, the
    fused loss, the exact backward, and the MLP backward, which adds its
    means gradient to the rasterizer's.  Parameters (means [P,3], scales
    [P,3], rotations [P,4], opacities [P,1]), the MLP weights (W1 [128,68],
    b1, W2 [128,128], b2, w3 [1,128], b3) and target [R,H,W] are read in place
    at every replay; after replay() `loss`, `out_real`, `radii` and `grads`
    (means, opacities, scales, rotations, and `mlp`: dW1 | db1 | dW2 | db2 |
    dw3 | db3 flat) hold the step's results.  No counterpart in the
    reference, whose emission MLP is per Gaussian (gaussian.py:47)."""

    def __init__(self, means, scales, rotations, opacities, mlp_weights, viewmatrices, tx_pos, rx_pos, target,
                 image_height, image_width, capacity, scale_modifier=1.0, lambda_ssim=LAMBDA_SSIM):
        self.params = dict(means=_require_param(means, "means"), scales=_require_param(scales, "scales"),
                           rotations=_require_param(rotations, "rotations"),
                           opacities=_require_param(opacities, "opacities"))
        self.weights = [_require_param(w, "mlp weight") for w in mlp_weights]
        if len(self.weights) != 6:
            raise RuntimeError("mlp_weights: W1, b1, W2, b2, w3, b3")
        dev = means.device
        P = int(means.size(0))
        self.views = gh.f32(viewmatrices).reshape(-1, 16)
        self.rx = gh.f32(rx_pos).reshape(-1, 3)
        self.tx = gh.f32(tx_pos).reshape(3)
        R = int(self.rx.size(0))
        H, W = int(image_height), int(image_width)
        self.target = _require_param(target, "target")
        if tuple(self.target.shape) != (R, H, W):
            raise RuntimeError("target must be [R,H,W]")
        self.P, self.R, self.H, self.W = P, R, H, W
        self.capacity = int(capacity)
        self.scale_modifier = float(scale_modifier)
        self.lambda_ssim = float(lambda_ssim)
        L = gh.lib()
        f32 = dict(dtype=torch.float32, device=dev)
        u8 = lambda n: torch.empty(int(n), dtype=torch.uint8, device=dev)
        self.geom = u8(L.gsparc_geom_buffer_size(P, R, W, H))
        self.img = u8(L.gsparc_image_buffer_size(P, R, W, H))
        self.bin = u8(L.gsparc_binning_buffer_size(self.capacity))
        self.gradbuf = u8(L.gsparc_grad_buffer_size(P, R))
        self.mlp_work = u8(L.gsparc_mlp1_work_size(P, R))
        self.work = _work(R, H, W, dev)
        self.emission = torch.empty((R, P, 2), **f32)
        self.demission = torch.empty((R, P, 2), **f32)
        self.radii = torch.empty((R, P), dtype=torch.int32, device=dev)
        self.out_real = torch.empty((R, H, W), **f32)
        self.out_imag = torch.empty((R, H, W), **f32)
        self.loss = torch.empty(1, **f32)
        self.dimg = torch.empty((R, H, W), **f32)
        self.zero_imag = torch.zeros((R, H, W), **f32)
        n_mlp = sum(int(w.numel()) for w in self.weights)
        self.grads = dict(means=torch.empty((P, 3), **f32), opacities=torch.empty((P, 1), **f32),
                          scales=torch.empty((P, 3), **f32), rotations=torch.empty((P, 4), **f32),
                          mlp=torch.empty(n_mlp, **f32))
        self._capture(dev)

    def _record(self, stream):
        L = gh.lib()
        p, g, w = self.params, self.grads, [gh.ptr(x) for x in self.weights]
        P, R, W, H = self.P, self.R, self.W, self.H
        gh.check(L.gsparc_mlp1_forward(P, R, gh.ptr(p["means"]), gh.ptr(self.rx), gh.ptr(self.tx), *w,
                                       gh.ptr(self.mlp_work), gh.ptr(self.emission), stream),
                 "graphed shared-MLP step: MLP forward")
        gh.check(L.gsparc_rasterize_forward(P, R, W, H, gh.ptr(p["means"]), gh.ptr(p["scales"]),
                                            gh.ptr(p["rotations"]), gh.ptr(p["opacities"]), None,
                                            gh.ptr(self.emission), gh.ptr(self.views), gh.ptr(self.tx),
                                            gh.ptr(self.rx), self.scale_modifier, gh.ptr(self.geom),
                                            gh.ptr(self.img), gh.ptr(self.bin), self.capacity, gh.ptr(self.radii),
                                            gh.ptr(self.out_real), gh.ptr(self.out_imag), 0, stream),
                 "graphed shared-MLP step: forward")
        gh.check(L.gsparc_loss_l1_ssim(R, H, W, self.lambda_ssim, SSIM_C1, SSIM_C2, gh.ptr(self.out_real),
                                       gh.ptr(self.target), gh.ptr(self.work), gh.ptr(self.loss), gh.ptr(self.dimg),
                                       stream), "graphed shared-MLP step: loss")
        gh.check(L.gsparc_rasterize_backward(gh.GRAD_EXACT, P, R, W, H, gh.ptr(p["means"]), gh.ptr(p["scales"]),
                                             gh.ptr(p["rotations"]), None, gh.ptr(self.emission),
                                             gh.ptr(self.views), gh.ptr(self.tx), gh.ptr(self.rx),
                                             self.scale_modifier, gh.ptr(self.radii), gh.ptr(self.geom),
                                             gh.ptr(self.img), gh.ptr(self.bin), self.capacity, gh.ptr(self.gradbuf),
                                             gh.ptr(self.dimg), gh.ptr(self.zero_imag), None, gh.ptr(g["means"]),
                                             None, gh.ptr(g["opacities"]), None, gh.ptr(g["scales"]),
                                             gh.ptr(g["rotations"]), gh.ptr(self.demission), stream),
                 "graphed shared-MLP step: backward")
        gh.check(L.gsparc_mlp1_backward_accumulate(P, R, gh.ptr(p["means"]), gh.ptr(self.rx), gh.ptr(self.tx), *w,
                                                   gh.ptr(self.mlp_work), gh.ptr(self.demission), gh.ptr(g["means"]),
                                                   gh.ptr(g["mlp"]), stream),
                 "graphed shared-MLP step: MLP backward")

    def instance_count(self) -> int:
        """K of the last step (synchronises the current stream)."""
        return wr.instance_count(self.P, self.R, self.H, self.W, self.img)


class GraphedSceneStep(_GraphedStep):
    """GraphedTrainingStep for a batch of uneven scenes (wireless_rasterizer.
    SceneBatch): forward of every scene's receivers, the fused loss over all
    R maps, backward, captured once and replayed.  Parameters are the packed
    [Ptot,...] tensors (emission_mlps packed by wireless_rasterizer.pack_mlps),
    viewmatrices [R,4,4] and rx_pos [R,3] packed by scene, tx_pos [S,3],
    target [R,H,W]; they are read in place at every replay.  No counterpart in
    the reference."""

    def __init__(self, means, scales, rotations, opacities, emission_mlps, viewmatrices, tx_pos, rx_pos, target,
                 image_height, image_width, scenes, capacity, scale_modifier=1.0, grad_mode="reference",
                 lambda_ssim=LAMBDA_SSIM):
        self.params = dict(means=_require_param(means, "means"), scales=_require_param(scales, "scales"),
                           rotations=_require_param(rotations, "rotations"),
                           opacities=_require_param(opacities, "opacities"),
                           emission_mlps=_require_param(emission_mlps, "emission_mlps"))
        if self.params["emission_mlps"].data_ptr() % 16:
            raise RuntimeError("emission_mlps must be 16-byte aligned")
        dev = means.device
        self.scenes = scenes
        P, R = scenes.Ptot, scenes.R
        if int(means.size(0)) != P or self.params["emission_mlps"].numel() != P * 194:
            raise RuntimeError("means and emission_mlps must hold every scene's Gaussians")
        self.views = gh.f32(viewmatrices).reshape(-1, 16)
        self.rx = gh.f32(rx_pos).reshape(-1, 3)
        self.tx = gh.f32(tx_pos).reshape(-1, 3)
        if self.views.size(0) != R or self.rx.size(0) != R or self.tx.size(0) != scenes.S:
            raise RuntimeError("one view matrix and position per receiver, one TX position per scene")
        H, W = int(image_height), int(image_width)
        self.target = _require_param(target, "target")
        if tuple(self.target.shape) != (R, H, W):
            raise RuntimeError("target must be [R,H,W]")
        self.P, self.R, self.H, self.W = P, R, H, W
        self.capacity = int(capacity)
        self.mode = gh.GRAD_MODES[grad_mode] if isinstance(grad_mode, str) else int(grad_mode)
        self.scale_modifier = float(scale_modifier)
        self.lambda_ssim = float(lambda_ssim)
        f32 = dict(dtype=torch.float32, device=dev)
        u8 = lambda n: torch.empty(int(n), dtype=torch.uint8, device=dev)
        geom_b, img_b, bin_b, grad_b = scenes.sizes(H, W, self.capacity)
        self.geom, self.img, self.bin, self.gradbuf = u8(geom_b), u8(img_b), u8(bin_b), u8(grad_b)
        self.work = _work(R, H, W, dev)
        self.radii = torch.empty(scenes.Nrec, dtype=torch.int32, device=dev)
        self.out_real = torch.empty((R, H, W), **f32)
        self.out_imag = torch.empty((R, H, W), **f32)
        self.loss = torch.empty(1, **f32)
        self.dimg = torch.empty((R, H, W), **f32)
        self.zero_imag = torch.zeros((R, H, W), **f32)
        self.grads = dict(means=torch.empty((P, 3), **f32), emission_mlps=torch.empty(P * 194, **f32),
                          opacities=torch.empty((P, 1), **f32), scales=torch.empty((P, 3), **f32),
                          rotations=torch.empty((P, 4), **f32))
        self._capture(dev)

    def _record(self, stream):
        L = gh.lib()
        p, g = self.params, self.grads
        bp, tab = C.byref(self.scenes.batch), gh.ptr(self.scenes.table)
        R, W, H = self.R, self.W, self.H
        gh.check(L.gsparc_rasterize_forward_scenes(bp, tab, W, H, gh.ptr(p["means"]), gh.ptr(p["scales"]),
                                                  gh.ptr(p["rotations"]), gh.ptr(p["opacities"]),
                                                  gh.ptr(p["emission_mlps"]), None, gh.ptr(self.views),
                                                  gh.ptr(self.tx), gh.ptr(self.rx), self.scale_modifier,
                                                  gh.ptr(self.geom), gh.ptr(self.img), gh.ptr(self.bin),
                                                  self.capacity, gh.ptr(self.radii), gh.ptr(self.out_real),
                                                  gh.ptr(self.out_imag), 0, stream), "graphed scene step: forward")
        gh.check(L.gsparc_loss_l1_ssim(R, H, W, self.lambda_ssim, SSIM_C1, SSIM_C2, gh.ptr(self.out_real),
                                       gh.ptr(self.target), gh.ptr(self.work), gh.ptr(self.loss), gh.ptr(self.dimg),
                                       stream), "graphed scene step: loss")
        gh.check(L.gsparc_rasterize_backward_scenes(
            self.mode, bp, tab, W, H, gh.ptr(p["means"]), gh.ptr(p["scales"]), gh.ptr(p["rotations"]),
            gh.ptr(p["emission_mlps"]), None, gh.ptr(self.views), gh.ptr(self.tx), gh.ptr(self.rx),
            self.scale_modifier, gh.ptr(self.radii), gh.ptr(self.geom), gh.ptr(self.img), gh.ptr(self.bin),
            self.capacity, gh.ptr(self.gradbuf), gh.ptr(self.dimg), gh.ptr(self.zero_imag), None,
            gh.ptr(g["means"]), gh.ptr(g["emission_mlps"]), gh.ptr(g["opacities"]), None, gh.ptr(g["scales"]),
            gh.ptr(g["rotations"]), None, stream), "graphed scene step: backward")

    def instance_count(self) -> int:
        """K of the last step (synchronises the current stream)."""
        K = C.c_uint64(0)
        gh.check(gh.lib().gsparc_instance_count_scenes(C.byref(self.scenes.batch), gh.ptr(self.scenes.table), self.W,
                                                       self.H, gh.ptr(self.img), C.byref(K),
                                                       gh.stream_handle(self.target.device)), "instance count")
        return int(K.value)
