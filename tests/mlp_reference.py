"""The shared emission MLP of the single-receiver training configuration
written with torch ops: the checker of the library's kernels (CPU FP64 and
GPU FP32).  Same parameters, names, shapes and initialisation as
gsparc_emission.SharedEmissionMLP, whose docstring defines the function."""
from __future__ import annotations

import math

import torch

N_FREQ = 10
N_PE = 3 + 6 * N_FREQ
N_RX = 5
N_IN = N_PE + N_RX
HIDDEN = 128


def positional_encoding(means: torch.Tensor) -> torch.Tensor:
    """[P,3] -> [P,63]: mu, sin(2^k pi mu), cos(2^k pi mu), k = 0..9, with
    2^k mu reduced exactly to [-1, 1] first."""
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


class TorchSharedMLP(torch.nn.Module):
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

    def amplitude(self, means, rx_pos, tx_pos):
        W1_pe, W1_rx = self.W1[:, :N_PE], self.W1[:, N_PE:]
        pe = positional_encoding(means) @ W1_pe.t()                  # [P,128], once per Gaussian
        rxf = receiver_features(means, rx_pos, tx_pos)             # [R,P,5]
        h1 = torch.relu(pe[None] + rxf @ W1_rx.t() + self.b1)
        h2 = torch.relu(h1 @ self.W2.t() + self.b2)
        return (h2 @ self.w3.t()).squeeze(-1) + self.b3

    def forward(self, means, rx_pos, tx_pos):
        a = self.amplitude(means, rx_pos, tx_pos)
        return torch.stack([a, torch.zeros_like(a)], dim=-1)
