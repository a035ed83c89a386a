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


# the multi-frequency backward's phasors: the exact phase at subcarriers 0 and
# 26 (hardware sin/cos), then the complex recurrence over the subcarrier
# spacing (25 steps, step factor from the polynomial gs_sincos2pi).  Bounds: a
# float32 emulation of the recurrence on the host reaches 1.42e-6 below 40 m
# and 4.24e-6 below 500 m; the anchors' hardware sin/cos add at most 1.3e-7.
PHASOR_ERR = {40.0: 1.8e-6, 500.0: 5.0e-6}


def test_backward_phasors_against_fp64():
    dev = "cuda"
    L = gh.lib()
    f0, df, c = 3.5e9 - 25.5 * 312.5e3, 312.5e3, 299792458.0
    g = torch.Generator().manual_seed(11)
    for top, bound in PHASOR_ERR.items():
        n = 1 << 19
        plen = (torch.rand(n, generator=g, dtype=torch.float64) * top).float()
        plen[:4] = torch.tensor([0.0, 1e-6, top / 2, top])
        d_plen = plen.to(dev)
        out = torch.empty((n, 52, 2), dtype=torch.float32, device=dev)
        gh.check(L.gsparc_csi_phasors(n, gh.ptr(d_plen), f0, df, gh.ptr(out), gh.stream_handle(dev)), "phasors")
        torch.cuda.synchronize()
        k = torch.arange(52, dtype=torch.float64)
        cyc = (plen.double()[:, None] * ((f0 + k * df) / c)[None, :]) % 1.0
        ref = torch.stack([torch.cos(2 * math.pi * cyc), -torch.sin(2 * math.pi * cyc)], -1)
        err = (out.cpu().double() - ref).abs()
        worst = float(err.max())
        print("phasors below", top, "m: worst |error|", worst, "at subcarrier",
              int(torch.argmax(err.amax(dim=(0, 2)))), flush=True)
        assert worst <= bound
        # the anchors of the two halves are the exact-phase factors
        assert float(err[:, 0].max()) <= 2.6e-7 and float(err[:, 26].max()) <= 2.6e-7
