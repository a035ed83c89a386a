"""Shared emission MLP of the single-receiver training configuration.

GSpaRC's own emission is a per-Gaussian 3->32->2 MLP (gaussian.py:47,
forward.cu:17-75), handled inside the rasterizer kernels.  The training
configuration measured here replaces it by ONE MLP shared by all Gaussians,
68 -> 128 -> 128 -> 1 with ReLU, evaluated for every (receiver, Gaussian)
pair.  It has no counterpart in the reference; its definition is this:

  features of (receiver r, Gaussian g), 68 values:
    PE(mu_g), 63:  mu, then sin(2^k pi mu), cos(2^k pi mu) for k = 0..9
                   (each a 3-vector), in that order
    d = |rx_r - mu_g|, 1 / max(d^2, 0.04)  (0.04 m^2 is the rasterizer's cull
                   distance, auxiliary.h:143), (tx - mu_g) / |tx - mu_g|
  amplitude a = w3 . relu(W2 relu(W1 x + b1) + b2) + b3
  emission (real, imag) = (a, 0), the rasterizer's `emission=` input [R,P,2]

Layer 1 is evaluated split: W1 = [W1_pe | W1_rx], with PE(mu) W1_pe^T formed
once per Gaussian and shared by all receivers, so only the 5 receiver
features are multiplied per pair.

Two backends compute the same function:
  "mfma"   (default on the GPU) the per-pair part -- layer 1's receiver term,
           the 128x128 layer, the ReLUs and the output dot -- runs in this
           build's fused kernels, the 128x128 layer as FP32 MFMA tiles
           (gsparc_mlp1_forward / _backward); torch keeps the plain GEMM
           PE(mu) W1_pe^T, the feature functions and their chain rule.
  "torch"  every contraction as a torch FP32 GEMM (hipBLASLt on the GPU),
           autograd for the backward; also the CPU / FP64 reference.
"""
from __future__ import annotations

import math

import torch

N_FREQ = 10
N_PE = 3 + 6 * N_FREQ          # 63
N_RX = 5                       # d, 1/d^2, TX direction
N_IN = N_PE + N_RX             # 68
HIDDEN = 128


def positional_encoding(means: torch.Tensor) -> torch.Tensor:
    """[P,3] -> [P,63]: mu, sin(2^k pi mu), cos(2^k pi mu), k = 0..9.

    2^k mu is exact in binary floating point and sin/cos(pi x) have period 2
    in x, so x is first reduced exactly to [-1, 1]: the argument handed to
    sin/cos is then accurate to the format's precision (an FP32 argument of
    2^9 pi mu ~ 1e4 rad would carry an absolute error of ~1e-3)."""
    out = [means]
    for k in range(N_FREQ):
        x = means * (2.0 ** k)
        x = x - 2.0 * torch.round(0.5 * x)
        a = math.pi * x
        out.append(torch.sin(a))
        out.append(torch.cos(a))
    return torch.cat(out, dim=-1)


def receiver_features(means: torch.Tensor, rx_pos: torch.Tensor, tx_pos: torch.Tensor) -> torch.Tensor:
    """[P,3], [R,3], [3] -> [R,P,5]: d, 1/max(d^2, 0.04), unit TX direction."""
    diff = rx_pos[:, None, :] - means[None, :, :]
    d2 = (diff * diff).sum(-1, keepdim=True)
    d = torch.sqrt(d2)
    inv = 1.0 / torch.clamp(d2, min=0.04)
    t = tx_pos.reshape(1, 3) - means
    tdir = t / torch.linalg.vector_norm(t, dim=-1, keepdim=True)
    R = rx_pos.shape[0]
    return torch.cat([d, inv, tdir[None].expand(R, -1, -1)], dim=-1)


class _MlpCoreMfma(torch.autograd.Function):
    """a [R*P] from pe1 [P,128] (= PE W1_pe^T + b1) and rxf [R*P,5] through the
    fused MFMA kernels; backward returns the gradients of all seven inputs."""

    @staticmethod
    def forward(ctx, pe1, rxf, w1rx, w2, b2, w3, b3):
        import gsparc_hip as gh
        P = pe1.shape[0]
        R = rxf.shape[0] // max(P, 1)
        ins = [gh.f32(x) for x in (pe1, rxf, w1rx, w2, b2, w3, b3)]
        gh.require_cuda(*ins)
        em = torch.empty((R * P, 2), dtype=torch.float32, device=pe1.device)
        gh.check(gh.lib().gsparc_mlp1_forward(P, R, *[gh.ptr(x) for x in ins], gh.ptr(em),
                                              gh.stream_handle(pe1.device)), "mlp1 forward")
        ctx.save_for_backward(*ins)
        ctx.PR = (P, R)
        return em[:, 0]

    @staticmethod
    def backward(ctx, da):
        import gsparc_hip as gh
        pe1, rxf, w1rx, w2, b2, w3, b3 = ctx.saved_tensors
        P, R = ctx.PR
        dev = pe1.device
        dem = torch.zeros((R * P, 2), dtype=torch.float32, device=dev)
        dem[:, 0] = da
        dh1 = torch.empty((R * P, HIDDEN), dtype=torch.float32, device=dev)
        L = gh.lib()
        work = torch.empty(L.gsparc_mlp1_work_size(P, R), dtype=torch.uint8, device=dev)
        grads = torch.empty(HIDDEN * HIDDEN + 2 * HIDDEN + 1, dtype=torch.float32, device=dev)
        gh.check(L.gsparc_mlp1_backward(P, R, *[gh.ptr(x) for x in (pe1, rxf, w1rx, w2, b2, w3, b3)], gh.ptr(dem),
                                        gh.ptr(dh1), gh.ptr(work), gh.ptr(grads), gh.stream_handle(dev)),
                 "mlp1 backward")
        dw2 = grads[:HIDDEN * HIDDEN].view(HIDDEN, HIDDEN)
        dw3 = grads[HIDDEN * HIDDEN:HIDDEN * HIDDEN + HIDDEN].view(1, HIDDEN)
        db2 = grads[HIDDEN * HIDDEN + HIDDEN:HIDDEN * HIDDEN + 2 * HIDDEN]
        db3 = grads[-1:]
        dpe1 = dh1.view(R, P, HIDDEN).sum(0)
        dw1rx = dh1.t() @ rxf
        drxf = dh1 @ w1rx
        return dpe1, drxf, dw1rx, dw2, db2, dw3, db3


class SharedEmissionMLP(torch.nn.Module):
    """68 -> 128 -> 128 -> 1 (ReLU) amplitude MLP shared by all Gaussians."""

    def __init__(self, generator: torch.Generator | None = None, dtype=torch.float32, device=None,
                 backend: str | None = None):
        super().__init__()
        self.backend = backend

        def lin(n_out, n_in):
            bound = 1.0 / math.sqrt(n_in)
            w = (torch.rand((n_out, n_in), generator=generator, dtype=dtype) * 2 - 1) * bound
            b = (torch.rand((n_out,), generator=generator, dtype=dtype) * 2 - 1) * bound
            return torch.nn.Parameter(w.to(device)), torch.nn.Parameter(b.to(device))

        self.W1, self.b1 = lin(HIDDEN, N_IN)
        self.W2, self.b2 = lin(HIDDEN, HIDDEN)
        self.w3, self.b3 = lin(1, HIDDEN)

    def _use_mfma(self, means):
        b = self.backend or ("mfma" if means.is_cuda and means.dtype == torch.float32 else "torch")
        if b == "mfma" and not (means.is_cuda and means.dtype == torch.float32):
            raise RuntimeError("the mfma backend needs float32 GPU tensors")
        return b == "mfma"

    def amplitude(self, means: torch.Tensor, rx_pos: torch.Tensor, tx_pos: torch.Tensor) -> torch.Tensor:
        """[R,P] amplitude for every (receiver, Gaussian) pair."""
        W1_pe, W1_rx = self.W1[:, :N_PE], self.W1[:, N_PE:]
        if self._use_mfma(means):
            pe1 = positional_encoding(means) @ W1_pe.t() + self.b1                # [P,128]
            rxf = receiver_features(means, rx_pos, tx_pos).reshape(-1, N_RX)      # [R*P,5]
            a = _MlpCoreMfma.apply(pe1, rxf, W1_rx, self.W2, self.b2, self.w3, self.b3)
            return a.view(rx_pos.shape[0], means.shape[0])
        pe = positional_encoding(means) @ W1_pe.t()                          # [P,128], once per Gaussian
        rxf = receiver_features(means, rx_pos, tx_pos)                     # [R,P,5]
        h1 = torch.relu(pe[None] + rxf @ W1_rx.t() + self.b1)             # [R,P,128]
        h2 = torch.relu(h1 @ self.W2.t() + self.b2)                       # [R,P,128]
        return (h2 @ self.w3.t()).squeeze(-1) + self.b3                   # [R,P]

    def forward(self, means: torch.Tensor, rx_pos: torch.Tensor, tx_pos: torch.Tensor) -> torch.Tensor:
        """Emission [R,P,2] = (amplitude, 0) for the rasterizer's emission input."""
        a = self.amplitude(means, rx_pos, tx_pos)
        return torch.stack([a, torch.zeros_like(a)], dim=-1)
