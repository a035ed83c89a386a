"""Autotuner of the blend mapping: it picks a valid (forward ppl, backward
ppl, order), later calls of that batch shape use it, and the results stay
what they were (forward bit-identical, backward within 1e-4)."""
import numpy as np
import pytest
import torch

import gsparc_hip as gh
import gsparc_scenes as S
import wireless_rasterizer as wr
from helpers import gpu_forward

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float((a - b).abs().max() / max(float(b.abs().max()), 1e-30))


def _bwd(sc, d, out, dre, dim):
    K, ore, oim, radii, geom, binbuf, img = out
    g = wr.backward_batched(d["means"], d["mlp"], radii, d["opac"], d["scales"], d["rots"], 1.0, d["views"], dre,
                            dim, d["tx"], d["rx"], geom, K, binbuf, img)
    torch.cuda.synchronize()
    return g


def test_autotune_uniform_batch():
    wr.autotune_clear()
    sc = S.small_scene(91, 2500, R=18, walls=True)
    print("autotune P=2500 R=18", flush=True)
    d, out = gpu_forward(sc)
    g = torch.Generator().manual_seed(3)
    dre = torch.randn((sc.R, sc.H, sc.W), generator=g).cuda()
    dim = torch.randn((sc.R, sc.H, sc.W), generator=g).cuda()
    ref = _bwd(sc, d, out, dre, dim)
    res = wr.autotune(out)
    torch.cuda.synchronize()
    assert res["fwd_ppl"] in (1, 2, 4) and res["bwd_ppl"] in (1, 2, 4) and res["order"] in (1, 2, 3)
    assert res["fwd_depth"] in (32, 64, 128) and res["bwd_depth"] in (32, 64, 128)
    assert res["fwd_ms"] > 0 and res["bwd_ms"] > 0
    # every candidate was timed: 3 orders x (3 pixel mappings x 3 batch depths), forward and backward
    assert len(res["times_ms"]) == 54 and all(t > 0 for t in res["times_ms"].values())
    assert res["fwd_ms"] == min(v for k, v in res["times_ms"].items() if "_fwd_" in k and
                                k.startswith(f"order{res['order']}_"))
    for name in ("render_fwd_ppl", "render_bwd_ppl", "render_order", "render_fwd_depth", "render_bwd_depth"):
        assert gh.get_option(name) == 0
    # the tuned shape renders the same spectra and gradients
    d2, out2 = gpu_forward(sc)
    assert torch.equal(out2[1], out[1]) and torch.equal(out2[2], out[2])
    g2 = _bwd(sc, d2, out2, dre, dim)
    for a, b in zip(g2[:7], ref[:7]):
        if a is not None:
            assert _rel(a, b) < 1e-4
    wr.autotune_clear()


def test_autotune_scene_batch():
    wr.autotune_clear()
    scs = [S.small_scene(92 + i, P, R=R) for i, (P, R) in enumerate([(800, 3), (2000, 9)])]
    import test_gpu_scenes as TS
    p, _ = TS._packed(scs)
    batch = wr.SceneBatch([sc.P for sc in scs], [sc.R for sc in scs], "cuda")
    out = TS._forward_scenes(scs, p, batch)
    res = wr.autotune(out, scenes=batch)
    torch.cuda.synchronize()
    assert res["order"] in (1, 2, 3)
    out2 = TS._forward_scenes(scs, p, batch)
    assert torch.equal(out2[1], out[1]) and torch.equal(out2[2], out[2])
    wr.autotune_clear()
