"""The multi-frequency kernels' phase factors (the hardware v_cos_f32 /
v_sin_f32 of a cycle fraction, k_render_csi.hip) against FP64 cos / sin of
2 pi x, over every float x in [-1/2, 1/2]: the domain the kernels feed them
(the exact cycle fraction of L f / c, reduced to [-1/2, 1/2])."""
import math

import numpy as np
import pytest
import torch

import gsparc_hip as gh

pytestmark = pytest.mark.gpu

# the measured maximum absolute error over the whole domain, 1.254e-7 for
# cos and sin alike (at x = 0.0078015 and x + 1/4), rounded up
MAX_ABS_ERR = 1.3e-7


def test_phase_factors_over_every_cycle_fraction():
    dev = "cuda"
    L = gh.lib()
    chunk = 1 << 26
    top = 0x3F000000                      # bits of 0.5
    worst = {"cos": (0.0, 0.0), "sin": (0.0, 0.0)}
    count = 0
    for sign in (0, 0x80000000):
        for start in range(0, top + 1, chunk):
            n = min(chunk, top + 1 - start)
            bits = (torch.arange(start, start + n, dtype=torch.int64, device=dev) | sign).to(torch.int32)
            x = bits.view(torch.float32)
            c = torch.empty_like(x)
            s = torch.empty_like(x)
            gh.check(L.gsparc_csi_phase_factors(n, gh.ptr(x), gh.ptr(c), gh.ptr(s), gh.stream_handle(dev)),
                     "phase factors")
            a = (2.0 * math.pi) * x.double()
            for name, got, ref in (("cos", c, torch.cos(a)), ("sin", s, torch.sin(a))):
                err = (got.double() - ref).abs()
                i = int(torch.argmax(err))
                e = float(err[i])
                if e > worst[name][0]:
                    worst[name] = (e, float(x[i]))
            count += n
    print("floats checked", count, "worst |error| (error, at x):", worst, flush=True)
    assert count == 2 * (top + 1)
    for name, (e, at) in worst.items():
        assert e <= MAX_ABS_ERR, (name, e, at)
    # the values the blend relies on exactly: x = 0 gives (1, 0)
    x = torch.zeros(1, device=dev)
    c = torch.empty_like(x)
    s = torch.empty_like(x)
    gh.check(L.gsparc_csi_phase_factors(1, gh.ptr(x), gh.ptr(c), gh.ptr(s), gh.stream_handle(dev)), "phase")
    assert float(c) == 1.0 and float(s) == 0.0
