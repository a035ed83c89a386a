"""Shared emission MLP of the single-receiver training configuration.

GSpaRC's own emission is a per-Gaussian 3->32->2 MLP (gaussian.py:47,
forward.cu:17-75), handled inside the rasterizer kernels.  The training
configuration measured here replaces it by ONE MLP shared by all Gaussians,
68 -> 128 -> 128 -> 1 with ReLU, evaluated for every (receiver, Gaussian)
pair.  It has no counterpart in the reference; its definition is this:

  features of (receiver r, Gaussian g), 68 values:
    PE(mu_g), 63:  mu, then sin(2^k pi mu), cos(2^k pi mu) for k = 0..9
                   (each a 3-vector), in that order; 2^k mu is reduced
                   exactly to [-1, 1] first (sin / cos(pi x) have period 2)
    d = |rx_r - mu_g|, 1 / max(d^2, 0.04)  (0.04 m^2 is the rasterizer's cull
                   distance, auxiliary.h:143), (tx - mu_g) / |tx - mu_g|
  amplitude a = w3 . relu(W2 relu(W1 x + b1) + b2) + b3
  emission (real, imag) = (a, 0), the rasterizer's `emission=` input [R,P,2]

Everything runs in this build's kernels (gsparc_mlp1_forward / _backward):
the first layer split into its PE part, once per Gaussian, and the 5
receiver features per pair; the layers as FP32 MFMA tiles; the backward
through the layers, the positional encoding and the receiver features to
the means.  There is no other backend: CPU tensors raise.
"""
from __future__ import annotations

import math

import torch

import gsparc_hip as gh

N_FREQ = 10
N_PE = 3 + 6 * N_FREQ          # 63
N_RX = 5                       # d, 1/d^2, TX direction
N_IN = N_PE + N_RX             # 68
HIDDEN = 128


class _SharedMlp(torch.autograd.Function):
    """a [R,P] from means [P,3], rx_pos [R,3], tx_pos [3] and the six weight
    tensors; backward returns the gradients of the means and the weights."""

    @staticmethod
    def forward(ctx, means, rx_pos, tx_pos, W1, b1, W2, b2, w3, b3):
        ins = [gh.f32(x) for x in (means, rx_pos, tx_pos, W1, b1, W2, b2, w3, b3)]
        gh.require_cuda(*ins)
        mu, rx, tx = ins[0].reshape(-1, 3), ins[1].reshape(-1, 3), ins[2].reshape(3)
        P, R = int(mu.shape[0]), int(rx.shape[0])
        dev = mu.device
        L = gh.lib()
        work = torch.empty(L.gsparc_mlp1_work_size(P, R), dtype=torch.uint8, device=dev)
        em = torch.empty((R, P, 2), dtype=torch.float32, device=dev)
        gh.check(L.gsparc_mlp1_forward(P, R, gh.ptr(mu), gh.ptr(rx), gh.ptr(tx), *[gh.ptr(x) for x in ins[3:]],
                                       gh.ptr(work), gh.ptr(em), gh.stream_handle(dev)), "mlp1 forward")
        ctx.save_for_backward(mu, rx, tx, *ins[3:], work)
        ctx.PR = (P, R)
        return em[..., 0]

    @staticmethod
    def backward(ctx, da):
        mu, rx, tx, W1, b1, W2, b2, w3, b3, work = ctx.saved_tensors
        P, R = ctx.PR
        dev = mu.device
        dem = torch.zeros((R, P, 2), dtype=torch.float32, device=dev)
        dem[..., 0] = da
        dmu = torch.empty((P, 3), dtype=torch.float32, device=dev)
        n = HIDDEN * N_IN + 3 * HIDDEN + HIDDEN * HIDDEN + 1
        grads = torch.empty(n, dtype=torch.float32, device=dev)
        gh.check(gh.lib().gsparc_mlp1_backward(P, R, gh.ptr(mu), gh.ptr(rx), gh.ptr(tx),
                                               *[gh.ptr(x) for x in (W1, b1, W2, b2, w3, b3)], gh.ptr(work),
                                               gh.ptr(dem), gh.ptr(dmu), gh.ptr(grads), gh.stream_handle(dev)),
                 "mlp1 backward")
        o = 0
        out = []
        for shape in ((HIDDEN, N_IN), (HIDDEN,), (HIDDEN, HIDDEN), (HIDDEN,), (1, HIDDEN), (1,)):
            k = 1
            for s_ in shape:
                k *= s_
            out.append(grads[o:o + k].view(shape))
            o += k
        dW1, db1, dW2, db2, dw3, db3 = out
        return dmu, None, None, dW1, db1, dW2, db2, dw3, db3


class SharedEmissionMLP(torch.nn.Module):
    """68 -> 128 -> 128 -> 1 (ReLU) amplitude MLP shared by all Gaussians."""

    def __init__(self, generator: torch.Generator | None = None, dtype=torch.float32, device=None):
        super().__init__()

        def lin(n_out, n_in):
            bound = 1.0 / math.sqrt(n_in)
            w = (torch.rand((n_out, n_in), generator=generator, dtype=dtype) * 2 - 1) * bound
            b = (torch.rand((n_out,), generator=generator, dtype=dtype) * 2 - 1) * bound
            return torch.nn.Parameter(w.to(device)), torch.nn.Parameter(b.to(device))

        self.W1, self.b1 = lin(HIDDEN, N_IN)
        self.W2, self.b2 = lin(HIDDEN, HIDDEN)
        self.w3, self.b3 = lin(1, HIDDEN)

    def amplitude(self, means: torch.Tensor, rx_pos: torch.Tensor, tx_pos: torch.Tensor) -> torch.Tensor:
        """[R,P] amplitude for every (receiver, Gaussian) pair."""
        return _SharedMlp.apply(means, rx_pos, tx_pos, self.W1, self.b1, self.W2, self.b2, self.w3, self.b3)

    def forward(self, means: torch.Tensor, rx_pos: torch.Tensor, tx_pos: torch.Tensor) -> torch.Tensor:
        """Emission [R,P,2] = (amplitude, 0) for the rasterizer's emission input."""
        a = self.amplitude(means, rx_pos, tx_pos)
        return torch.stack([a, torch.zeros_like(a)], dim=-1)
