"""CPU tests: accuracy of the shared float32 math functions (gs_math.h) and
the C ABI library's exported symbols (no GPU calls)."""
import ctypes
import os
import re

import numpy as np

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _ulps(got, ref):
    got = got.astype(np.float64)
    rf = ref.astype(np.float32)
    spacing = np.spacing(np.abs(rf)).astype(np.float64)
    spacing = np.maximum(spacing, 2.0 ** -149)
    return np.abs(got - ref) / spacing


def test_expf_accuracy():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-103.0, 88.7, 2_000_000), rng.uniform(-12, 0, 2_000_000)]).astype(np.float32)
    got = O.math_vec("exp", x)
    assert _ulps(got, np.exp(x.astype(np.float64))).max() <= 2.0
    assert O.math_vec("exp", np.array([0.0, -200.0, 100.0, np.nan], np.float32)).tolist()[:3] == [1.0, 0.0, np.inf]


def test_expf_nonpos_is_bit_identical():
    """The blend kernels use the branch-free variant for power <= 0."""
    rng = np.random.default_rng(3)
    x = np.concatenate([-rng.uniform(0, 120, 4_000_000), -np.exp(rng.uniform(-30, 5, 1_000_000)),
                        np.array([0.0, -0.0, -104.0, -103.99999, -104.00001, -1e30, -np.inf])]).astype(np.float32)
    a = O.math_vec("exp", x)
    b = O.math_vec("exp_nonpos", x)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_logf_accuracy():
    rng = np.random.default_rng(1)
    x = np.exp(rng.uniform(-80, 80, 3_000_000)).astype(np.float32)
    x = np.concatenate([x, rng.uniform(0.2, 12.0, 1_000_000).astype(np.float32),
                        np.array([1e-40, 1.0, 2.0], np.float32)])
    got = O.math_vec("log", x)
    assert _ulps(got, np.log(x.astype(np.float64))).max() <= 2.0
    assert O.math_vec("log", np.array([1.0], np.float32))[0] == 0.0


def test_asinf_accuracy():
    x = np.linspace(-1.0, 1.0, 4_000_001, dtype=np.float32)
    got = O.math_vec("asin", x)
    assert _ulps(got, np.arcsin(x.astype(np.float64))).max() <= 2.0


def test_atan2f_accuracy():
    rng = np.random.default_rng(2)
    y = (rng.normal(size=3_000_000) * np.exp(rng.uniform(-5, 5, 3_000_000))).astype(np.float32)
    x = (rng.normal(size=3_000_000) * np.exp(rng.uniform(-5, 5, 3_000_000))).astype(np.float32)
    got = O.math_vec("atan2", x, y)
    assert _ulps(got, np.arctan2(y.astype(np.float64), x.astype(np.float64))).max() <= 2.0
    special = O.math_vec("atan2", np.array([1, -1, 0, 0], np.float32), np.array([0, 0, 1, -1], np.float32))
    np.testing.assert_allclose(special, [0.0, np.pi, np.pi / 2, -np.pi / 2], rtol=1e-7)


def test_expf_blend_is_bit_identical_on_its_domain():
    """gs_expf_blend (the blend kernels' exp) equals gs_expf bit for bit on
    its domain [-86, 0], including the ties of rint(x log2 e) and the edges."""
    rng = np.random.default_rng(11)
    x = np.concatenate([-rng.uniform(0, 86, 4_000_000), -np.exp(rng.uniform(-30, np.log(86), 1_000_000)),
                        np.array([0.0, -0.0, -86.0, -85.99999, -5.545177, -1e-30])]).astype(np.float32)
    # arguments whose x log2 e rounds to exactly k + 1/2 (ties of the rounding)
    k = np.arange(-124, 0, dtype=np.float64)
    ties = ((k + 0.5) / np.log2(np.e)).astype(np.float32)
    x = np.concatenate([x, ties, np.nextafter(ties, 0), np.nextafter(ties, -1)]).astype(np.float32)
    x = x[x >= -86.0]
    np.testing.assert_array_equal(O.math_vec("exp_blend", x).view(np.uint32), O.math_vec("exp", x).view(np.uint32))


def test_expf_blend_is_monotone_on_its_whole_domain():
    """Exhaustive over the 1.1e9 floats of [-86, 0]: the blend exponential
    never increases as its argument decreases.  That is what makes the skip
    decision o * exp(power) >= 1/255 a single threshold on the power."""
    assert O.expf_blend_monotone_violations() == 0


def test_alpha_threshold_decides_exactly_like_the_alpha_test():
    """power >= gs_alpha_thr(o) (both blend kernels) against the reference's
    !(min(0.99, o exp(power)) < 1/255) (forward.cu:219-221) with the oracle's
    exponential: every float within 256 of each threshold and random powers,
    for random and edge-case opacities."""
    rng = np.random.default_rng(5)
    o = 1.0 / (1.0 + np.exp(-rng.normal(size=4000)))
    a = np.float32(1.0 / 255.0)
    edge = [a, np.nextafter(a, np.float32(1)), np.nextafter(a, np.float32(0)), 1.0, 0.99, 0.5, 1e-3, 0.0, -0.5,
            1e30, 1e36, 2.0, 254.9, 255.0, 1e-30]
    o = np.concatenate([o, np.exp(rng.uniform(np.log(a), 0.0, 4000)), edge]).astype(np.float32)
    checked, bad = O.check_alpha_thr(o, span=256, nrand=256, seed=7)
    assert checked > 3_000_000 and bad == 0
    thr = O.alpha_thr(o)
    # the threshold is the boundary itself: thr passes, the next float below
    # fails (a threshold one float off in either direction breaks one of the two)
    ok = (o >= a) & (o < 1e34)
    t = thr[ok]
    below = np.nextafter(t, np.float32(-np.inf))
    ob = o[ok]
    e_at = O.math_vec("exp", t)
    e_below = O.math_vec("exp", below)
    assert np.all(np.minimum(np.float32(0.99), ob * e_at) >= a)
    assert np.all(np.minimum(np.float32(0.99), ob * e_below) < a)
    # special values
    s = O.alpha_thr(np.array([np.nan, 0.0, -1.0, 1e-3, 1e36], np.float32))
    assert np.isnan(s[0]) and s[1] == 1.0 and s[2] == 1.0 and s[3] == 1.0 and s[4] == -86.0


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "gsparc.h")).read()
    declared = set(re.findall(r"\b(gsparc_\w+)\s*\(", header))
    import gsparc_hip as gh
    lib = ctypes.CDLL(gh.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(gh.exported_symbols())
    assert gh.lib().gsparc_version().startswith(b"gsparc")


def test_buffer_layout_sizes():
    import gsparc_hip as gh
    L = gh.layout(50000, 64, 360, 90, 1 << 20)
    assert L.splat == 0 and L.sig >= 64 * 50000 * 32 and L.geom_bytes > L.nvis > L.sort_hist
    assert L.sort_val1 - L.sort_key1 >= 4 * 64 * 50000
    assert L.chunk_cnt - L.segbase >= 4 * 64 * 138 and L.total - L.chunk_cnt >= 4 * 64 * 138
    for off in (L.sig, L.depth, L.cov3, L.sort_key0, L.sort_val0, L.sort_key1, L.sort_val1, L.sort_hist,
                L.n_contrib, L.ranges, L.segcnt, L.segbase, L.chunk_cnt, L.plist, L.keys_sorted):
        assert off % 256 == 0
    # the tile lists need 4 bytes per instance, 8 with the sorted keys kept
    assert 8 * (1 << 20) <= gh.lib().gsparc_binning_buffer_size(1 << 20) <= 8 * (1 << 20) + 1024


def test_tuning_options():
    import gsparc_hip as gh
    assert gh.get_option("render_bwd_ppl") == 0
    gh.set_option("render_bwd_ppl", 4)
    assert gh.get_option("render_bwd_ppl") == 4
    gh.set_option("render_bwd_ppl", 0)
    import pytest
    with pytest.raises(RuntimeError, match="render_fwd_ppl"):
        gh.set_option("render_fwd_ppl", 3)
    with pytest.raises(RuntimeError, match="unknown option"):
        gh.set_option("no_such_option", 1)
    for name in ("render_fwd_depth", "render_bwd_depth"):
        assert gh.get_option(name) == 0
        for v in (32, 64, 128, 0):
            gh.set_option(name, v)
            assert gh.get_option(name) == v
        for bad in (16, 48, 256, -1):
            with pytest.raises(RuntimeError, match=name):
                gh.set_option(name, bad)


def test_deterministic_workspace_size():
    """A deterministic backward's workspace: the normal one rounded up to
    256 bytes, 8 floats per tile-list instance, and a word per record and per
    1024 records (bounded by an eighth of the normal size)."""
    import gsparc_hip as gh
    L = gh.lib()
    assert L.gsparc_deterministic_grad_size(1000, 10) == 1024 + 512 + 256 + 4096
    assert L.gsparc_deterministic_grad_size(1024, 0) == 1024 + 256 + 256 + 4096
    P, R, K = 50000, 64, 15_740_375
    g = L.gsparc_grad_buffer_size(P, R)
    assert g >= 32 * P * R
    assert L.gsparc_deterministic_grad_size(g, K) >= g + 32 * K + 4 * (P * R + P * R // 1024 + 1)
    assert gh.grad_mode_code("exact", True) == 16 and gh.grad_mode_code("reference", True) == 17
    assert gh.grad_mode_code("reference") == 1


def test_reference_import_paths():
    """GSpaRC's own import lines resolve to this build (renderer.py:5, train.py:40)."""
    from cuda_kernel_rasterize.cuda_rasterizer import GaussianRasterizer, GaussianRasterizerSetting
    import cuda_rasterizer
    from fused_ssim import fused_ssim
    assert GaussianRasterizer is cuda_rasterizer.GaussianRasterizer
    assert GaussianRasterizerSetting is cuda_rasterizer.GaussianRasterizerSetting
    assert callable(fused_ssim)


def test_scene_table_build():
    """Host-side scene table of an uneven batch: per-receiver record offsets,
    receiver groups of at most 16 (the preprocess kernels' group) and
    per-scene spans."""
    import ctypes as C
    import gsparc_hip as gh
    L = gh.lib()
    Ps, Rs = [5, 0, 7, 3], [20, 2, 0, 1]
    S, R = len(Ps), sum(Rs)
    words = L.gsparc_scene_table_words(S, R)
    tab = (C.c_uint32 * words)()
    b = gh.Batch()
    assert L.gsparc_scene_table_build(S, (C.c_int * S)(*Ps), (C.c_int * S)(*Rs), tab, words, C.byref(b)) == 0
    G = 2 + 1 + 0 + 1
    assert (b.S, b.R, b.G, b.Pmax, b.Ptot, b.Nrec, b.Npart) == (4, 23, G, 7, 15, 20 * 5 + 1 * 3, 2 * 5 + 3)
    t = np.frombuffer(tab, dtype=np.uint32)
    assert list(t[:7]) == [4, 23, G, 7, 15, 103, 13]
    rx = t[8:8 + 4 * R].reshape(R, 4)
    # receivers 0..19 -> scene 0 (5 records each), 20..21 -> scene 1 (no Gaussians), 22 -> scene 3
    np.testing.assert_array_equal(rx[:, 0], [5 * i for i in range(20)] + [100, 100, 100])
    np.testing.assert_array_equal(rx[:, 1], [0] * 20 + [5, 5, 12])
    np.testing.assert_array_equal(rx[:, 2], [5] * 20 + [0, 0, 3])
    np.testing.assert_array_equal(rx[:, 3], [0] * 20 + [1, 1, 3])
    gr = t[8 + 4 * R:8 + 4 * R + 4 * G].reshape(G, 4)
    np.testing.assert_array_equal(gr, [[0, 16, 0, 0], [16, 4, 0, 5], [20, 2, 1, 10], [22, 1, 3, 10]])
    sc = t[8 + 4 * R + 4 * G:8 + 4 * R + 4 * G + 8 * S].reshape(S, 8)
    np.testing.assert_array_equal(sc[:, :6], [[0, 5, 0, 20, 0, 2], [5, 0, 20, 2, 2, 1], [5, 7, 22, 0, 3, 0],
                                              [12, 3, 22, 1, 3, 1]])
    sizes = (C.c_size_t * 4)()
    assert L.gsparc_scene_buffer_sizes(C.byref(b), 360, 90, 100, sizes) == 0
    assert all(x > 0 for x in sizes)
    # too small a table and negative sizes are refused
    assert L.gsparc_scene_table_build(S, (C.c_int * S)(*Ps), (C.c_int * S)(*Rs), tab, words - 1, C.byref(b)) != 0
    assert L.gsparc_scene_table_build(1, (C.c_int * 1)(-1), (C.c_int * 1)(1), tab, words, C.byref(b)) != 0


def test_argument_checks_before_any_device_work():
    """Size checks run on the host before any HIP call: a map of 2^20 or more
    16x16 tiles (the float tile-row division of the ordered scatter is exact
    below that) and a column-plus-row count over 8192 are refused."""
    import ctypes as C
    import gsparc_hip as gh
    L = gh.lib()
    K = C.c_uint64(0)
    img = C.c_void_p(1)
    assert L.gsparc_instance_count(10, 1, 16 * 1024, 16 * 1024, img, C.byref(K), None) != 0
    assert b"2^20 tiles" in L.gsparc_last_error()
    assert L.gsparc_instance_count(10, 1, 16 * 8000, 16 * 200, img, C.byref(K), None) != 0
    assert b"W + H too large" in L.gsparc_last_error()
    assert L.gsparc_instance_count(10, 1, 360, 90, None, C.byref(K), None) != 0
    assert b"null pointer" in L.gsparc_last_error()
    # the depth sort's look-back words count a receiver's Gaussians in 28 bits
    assert L.gsparc_instance_count(1 << 28, 1, 360, 90, img, C.byref(K), None) != 0
    assert b"below 2^28" in L.gsparc_last_error()
    # the device keeps K saturated at 2^32 - 1: a capacity must stay below it
    dummy = C.c_void_p(256)
    assert L.gsparc_rasterize_forward_render(10, 1, 360, 90, None, dummy, dummy, dummy, (1 << 32) - 1, dummy, dummy,
                                             dummy, 0, None) != 0
    assert b"K_capacity must be below 2^32 - 1" in L.gsparc_last_error()


def test_autograd_forms_refuse_torch_graph_capture(monkeypatch):
    """The batched, scene and CSI autograd forwards refuse torch.cuda.graph
    capture before any device work, pointing to the library-owned graph."""
    import pytest
    import torch
    import cuda_rasterizer as CR
    monkeypatch.setattr(torch.cuda, "is_available", lambda: True)
    monkeypatch.setattr(torch.cuda, "is_current_stream_capturing", lambda: True)
    for fn in (CR.rasterize_gaussians_batched, CR.rasterize_scenes, CR.rasterize_csi):
        with pytest.raises(RuntimeError, match="GraphedTrainingStep"):
            fn(torch.zeros(1, 3), None, torch.zeros(1), torch.ones(1, 3), torch.ones(1, 4), None)
