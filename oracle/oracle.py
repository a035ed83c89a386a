"""ctypes front end of the CPU oracle (oracle/gs_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Nothing in the product imports
it.  Build with `make -C oracle` (done by __graft_entry__.build()).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libgs_oracle.so")
_lib = None

GRAD_EXACT = 0
GRAD_REFERENCE = 1
BLOCK = 16


def build():
    import subprocess
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _lib.orc_bin.restype = C.c_int64
        for n in ("orc_gs_expf", "orc_gs_logf", "orc_gs_asinf"):
            getattr(_lib, n).restype = C.c_float
            getattr(_lib, n).argtypes = [C.c_float]
        _lib.orc_gs_atan2f.restype = C.c_float
        _lib.orc_gs_atan2f.argtypes = [C.c_float, C.c_float]
    return _lib


def math_vec(which, x, y=None):
    """gs_math.h functions over arrays: which in {"exp", "log", "asin", "atan2"}."""
    code = {"exp": 0, "log": 1, "asin": 2, "atan2": 3, "exp_nonpos": 4, "cos2pi": 5, "sin2pi": 6, "exp_blend": 7}[which]
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(x if y is None else y, np.float32)
    out = np.empty_like(x)
    lib().orc_gs_math_vec(C.c_int(code), _p(x), _p(y), _p(out), C.c_int64(x.size))
    return out


def set_threads(n=0):
    """Threads of the oracle's OpenMP loops (n <= 0: unchanged); returns the
    count in effect."""
    f = lib().orc_set_threads
    f.restype = C.c_int
    return int(f(C.c_int(int(n))))


def alpha_thr(o):
    """gs_alpha_thr over an array of opacities (float32)."""
    o = np.ascontiguousarray(o, np.float32)
    out = np.empty_like(o)
    lib().orc_alpha_thr(C.c_int64(o.size), _p(o), _p(out))
    return out


def expf_blend_monotone_violations():
    """Exhaustive: increases of gs_expf_blend as its argument walks down
    through every float from -0 to -86 (0 required)."""
    f = lib().orc_expf_blend_monotone
    f.restype = C.c_int64
    return int(f())


def check_alpha_thr(o, span=256, nrand=256, seed=0):
    """(pairs checked, disagreements) between power >= gs_alpha_thr(o) and
    the alpha >= 1/255 test, on the `span` floats either side of each
    threshold and `nrand` random powers."""
    o = np.ascontiguousarray(o, np.float32)
    stats = np.zeros(2, np.int64)
    lib().orc_check_alpha_thr(C.c_int64(o.size), _p(o), C.c_int(span), C.c_int(nrand), C.c_uint64(seed), _p(stats))
    return int(stats[0]), int(stats[1])


def pair_power(pre, px, py):
    """Power of each Gaussian of `pre` at pixel (px[i], py[i]), as the FP32
    mirror and the kernels evaluate it."""
    P = pre.radii.shape[0]
    out = np.zeros(P, np.float32)
    lib().orc_pair_power_f32(C.c_int64(P), _p(_c(pre.conic_o, np.float32)), _p(_c(pre.xy, np.float32)),
                             _p(np.ascontiguousarray(px, np.int32)), _p(np.ascontiguousarray(py, np.int32)), _p(out))
    return out


def cull_check(n, seed=0, margin=4.0e-6):
    """Brute-force check of the blend kernels' wave-level culling test:
    [cases, culled, violations, cases where every pixel skips]."""
    stats = np.zeros(4, np.int64)
    lib().orc_cull_check(C.c_int64(n), C.c_uint64(seed), C.c_float(margin), _p(stats))
    return stats


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _ft(dtype):
    return C.c_float if dtype == np.float32 else C.c_double


class Pre:
    """Per-Gaussian preprocess outputs for one receiver."""

    def __init__(self, P, dtype):
        self.radii = np.zeros(P, np.int32)
        self.xy = np.zeros((P, 2), dtype)
        self.depth = np.zeros(P, dtype)
        self.cov3 = np.zeros((P, 6), dtype)
        self.conic_o = np.zeros((P, 4), dtype)
        self.sig_re = np.zeros(P, dtype)
        self.sig_im = np.zeros(P, dtype)
        self.tiles = np.zeros(P, np.int32)
        self.rect = np.zeros((P, 4), np.int32)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def preprocess(scene, r=0, dtype=np.float32, mod=1.0):
    """forward.cu:252-327 for receiver r."""
    P, W, H = scene.P, scene.W, scene.H
    args = [_c(scene.means, dtype), _c(scene.scales, dtype), _c(scene.rotations, dtype),
            _c(scene.opacities.reshape(-1), dtype), _c(scene.mlp, dtype), _c(scene.views[r].reshape(-1), dtype),
            _c(scene.tx_pos, dtype), _c(scene.rx_pos[r], dtype)]
    o = Pre(P, dtype)
    f = lib().orc_preprocess_f32 if dtype == np.float32 else lib().orc_preprocess_f64
    f(C.c_int(P), C.c_int(W), C.c_int(H), *[_p(a) for a in args], _ft(dtype)(mod),
      _p(o.radii), _p(o.xy), _p(o.depth), _p(o.cov3), _p(o.conic_o), _p(o.sig_re), _p(o.sig_im),
      _p(o.tiles), _p(o.rect))
    return o


def n_tiles(W, H):
    return ((W + BLOCK - 1) // BLOCK) * ((H + BLOCK - 1) // BLOCK)


def bin_gaussians(pre, W, H):
    """Keys, values (point list) and tile ranges (rasterizer_impl.cu:102-171)."""
    P = pre.radii.shape[0]
    depth32 = np.ascontiguousarray(pre.depth, dtype=np.float32)
    rect = np.ascontiguousarray(pre.rect)
    K = lib().orc_bin(C.c_int(P), C.c_int(W), C.c_int(H), _p(rect), _p(depth32), None, None, None)
    keys = np.zeros(max(K, 1), np.uint64)
    vals = np.zeros(max(K, 1), np.uint32)
    ranges = np.zeros((n_tiles(W, H), 2), np.uint32)
    lib().orc_bin(C.c_int(P), C.c_int(W), C.c_int(H), _p(rect), _p(depth32), _p(keys), _p(vals), _p(ranges))
    return keys[:K], vals[:K], ranges


def render_fwd(pre, vals, ranges, W, H, dtype=None, want_hash=False):
    dtype = dtype or pre.xy.dtype
    N = W * H
    out_re = np.zeros(N, dtype)
    out_im = np.zeros(N, dtype)
    T = np.zeros(N, dtype)
    nc = np.zeros(N, np.uint32)
    stats = np.zeros(4, np.int64)
    h = np.zeros(N, np.uint64) if want_hash else None
    vals = np.ascontiguousarray(vals if len(vals) else np.zeros(1, np.uint32), dtype=np.uint32)
    f = lib().orc_render_fwd_f32 if dtype == np.float32 else lib().orc_render_fwd_f64
    f(C.c_int(W), C.c_int(H), _p(np.ascontiguousarray(ranges)), _p(vals), _p(_c(pre.xy, dtype)),
      _p(_c(pre.conic_o, dtype)), _p(_c(pre.sig_re, dtype)), _p(_c(pre.sig_im, dtype)),
      _p(out_re), _p(out_im), _p(T), _p(nc), _p(stats), _p(h) if h is not None else None)
    res = dict(out_re=out_re.reshape(H, W), out_im=out_im.reshape(H, W), final_T=T.reshape(H, W),
               n_contrib=nc.reshape(H, W), stats=stats)
    if want_hash:
        res["hash"] = h
    return res


def render_bwd(pre, vals, ranges, fwd, dre, dim, W, H, mode, dtype=None):
    """Returns [P,8] float64: dpx, dpy, dqa, dqb, dqc, dopacity, dsig_re, dsig_im."""
    dtype = dtype or pre.xy.dtype
    P = pre.radii.shape[0]
    g8 = np.zeros((P, 8), np.float64)
    vals = np.ascontiguousarray(vals if len(vals) else np.zeros(1, np.uint32), dtype=np.uint32)
    f = lib().orc_render_bwd_f32 if dtype == np.float32 else lib().orc_render_bwd_f64
    f(C.c_int(mode), C.c_int(W), C.c_int(H), _p(np.ascontiguousarray(ranges)), _p(vals),
      _p(_c(pre.xy, dtype)), _p(_c(pre.conic_o, dtype)), _p(_c(pre.sig_re, dtype)), _p(_c(pre.sig_im, dtype)),
      _p(_c(fwd["final_T"].reshape(-1), dtype)), _p(np.ascontiguousarray(fwd["n_contrib"].reshape(-1))),
      _p(_c(np.asarray(dre).reshape(-1), dtype)), _p(_c(np.asarray(dim).reshape(-1), dtype)), _p(g8))
    return g8


class Grads:
    def __init__(self, P):
        self.means = np.zeros((P, 3))
        self.scales = np.zeros((P, 3))
        self.rotations = np.zeros((P, 4))
        self.opacities = np.zeros((P, 1))
        self.mlp = np.zeros(P * 194)
        self.cov3 = np.zeros((P, 6))


def preprocess_bwd(scene, pre, g8, mode, r=0, dtype=np.float32, mod=1.0, grads=None):
    """backward.cu:124-437 for receiver r; accumulates into `grads`."""
    P, W, H = scene.P, scene.W, scene.H
    g = grads or Grads(P)
    args = [_c(scene.means, dtype), _c(scene.scales, dtype), _c(scene.rotations, dtype), _c(scene.mlp, dtype),
            _c(scene.views[r].reshape(-1), dtype), _c(scene.tx_pos, dtype), _c(scene.rx_pos[r], dtype)]
    f = lib().orc_preprocess_bwd_f32 if dtype == np.float32 else lib().orc_preprocess_bwd_f64
    f(C.c_int(mode), C.c_int(P), C.c_int(W), C.c_int(H), *[_p(a) for a in args], _ft(dtype)(mod),
      _p(np.ascontiguousarray(pre.radii)), _p(np.ascontiguousarray(g8, dtype=np.float64)),
      _p(g.means), _p(g.scales), _p(g.rotations), _p(g.opacities), _p(g.mlp), _p(g.cov3))
    return g


def render(scene, r=0, dtype=np.float32):
    """Full forward for one receiver: preprocess, binning, blend."""
    pre = preprocess(scene, r, dtype)
    keys, vals, ranges = bin_gaussians(pre, scene.W, scene.H)
    fwd = render_fwd(pre, vals, ranges, scene.W, scene.H, dtype)
    return pre, keys, vals, ranges, fwd


def backward(scene, pre, vals, ranges, fwd, dre, dim, mode, r=0, dtype=np.float32, grads=None):
    g8 = render_bwd(pre, vals, ranges, fwd, dre, dim, scene.W, scene.H, mode, dtype)
    return preprocess_bwd(scene, pre, g8, mode, r, dtype, grads=grads), g8


# ---------------------------------------------------------------- SSIM
C1 = 0.01 ** 2
C2 = 0.03 ** 2


def ssim_fwd(img1, img2, dtype=np.float64, train=True):
    """fused-ssim/ssim.cu:187-286.  img: [B,CH,H,W]."""
    B, CH, H, W = img1.shape
    a = _c(img1, dtype)
    b = _c(img2, dtype)
    m = np.zeros_like(a)
    d1 = np.zeros_like(a) if train else None
    d2 = np.zeros_like(a) if train else None
    d3 = np.zeros_like(a) if train else None
    f = lib().orc_ssim_fwd_f32 if dtype == np.float32 else lib().orc_ssim_fwd_f64
    f(C.c_int(B * CH), C.c_int(H), C.c_int(W), _ft(dtype)(C1), _ft(dtype)(C2), _p(a), _p(b), _p(m),
      _p(d1) if train else None, _p(d2) if train else None, _p(d3) if train else None)
    return m, d1, d2, d3


def ssim_bwd(img1, img2, dmap, d1, d2, d3, dtype=np.float64):
    """fused-ssim/ssim.cu:288-366."""
    B, CH, H, W = img1.shape
    out = np.zeros(img1.shape, dtype)
    f = lib().orc_ssim_bwd_f32 if dtype == np.float32 else lib().orc_ssim_bwd_f64
    f(C.c_int(B * CH), C.c_int(H), C.c_int(W), _p(_c(img1, dtype)), _p(_c(img2, dtype)), _p(_c(dmap, dtype)),
      _p(_c(d1, dtype)), _p(_c(d2, dtype)), _p(_c(d3, dtype)), _p(out))
    return out


# ---------------------------------------------------------------- CSI
CSI_F0 = 3.5e9 - 25.5 * 312.5e3
CSI_DF = 312.5e3
CSI_NC = 52
LIGHT_SPEED = 299792458.0


def sincos2pi(x):
    """gs_sincos2pi over an array: (cos 2 pi x, sin 2 pi x) in float32."""
    return math_vec("cos2pi", x), math_vec("sin2pi", x)


def path_length(scene, r=0, dtype=np.float32):
    """|mu - tx| + |mu - rx_r| per Gaussian (mirror: the kernels' op order)."""
    P = scene.P
    out = np.zeros(P, dtype)
    f = lib().orc_path_length_f32 if dtype == np.float32 else lib().orc_path_length_f64
    f(C.c_int(P), _p(_c(scene.means, dtype)), _p(_c(scene.tx_pos, dtype)), _p(_c(scene.rx_pos[r], dtype)), _p(out))
    return out


def render_fwd_csi(pre, plen, vals, ranges, W, H, dtype=None, f0=CSI_F0, df=CSI_DF, sig=None):
    """CSI blend of one receiver: out [H,W,NC,2].  sig (optional) replaces
    the preprocess emission: (sig_re, sig_im) arrays of length P."""
    dtype = dtype or pre.xy.dtype
    N = W * H
    out = np.zeros(N * CSI_NC * 2, dtype)
    T = np.zeros(N, dtype)
    nc = np.zeros(N, np.uint32)
    sre, sim = (pre.sig_re, pre.sig_im) if sig is None else sig
    vals = np.ascontiguousarray(vals if len(vals) else np.zeros(1, np.uint32), dtype=np.uint32)
    f = lib().orc_render_fwd_csi_f32 if dtype == np.float32 else lib().orc_render_fwd_csi_f64
    f(C.c_int(W), C.c_int(H), C.c_double(f0), C.c_double(df), _p(np.ascontiguousarray(ranges)), _p(vals),
      _p(_c(pre.xy, dtype)), _p(_c(pre.conic_o, dtype)), _p(_c(sre, dtype)), _p(_c(sim, dtype)),
      _p(_c(plen, dtype)), _p(out), _p(T), _p(nc))
    return dict(out=out.reshape(H, W, CSI_NC, 2), final_T=T.reshape(H, W), n_contrib=nc.reshape(H, W))


def render_bwd_csi(pre, plen, vals, ranges, fwd, dout, W, H, dtype=None, f0=CSI_F0, df=CSI_DF, sig=None):
    """Exact adjoint of render_fwd_csi: [P,9] float64 = dpx, dpy, dqa, dqb,
    dqc, dopacity, dsig_re, dsig_im, dL/d(path length)."""
    dtype = dtype or pre.xy.dtype
    P = pre.radii.shape[0]
    g9 = np.zeros((P, 9), np.float64)
    sre, sim = (pre.sig_re, pre.sig_im) if sig is None else sig
    vals = np.ascontiguousarray(vals if len(vals) else np.zeros(1, np.uint32), dtype=np.uint32)
    f = lib().orc_render_bwd_csi_f32 if dtype == np.float32 else lib().orc_render_bwd_csi_f64
    f(C.c_int(W), C.c_int(H), C.c_double(f0), C.c_double(df), _p(np.ascontiguousarray(ranges)), _p(vals),
      _p(_c(pre.xy, dtype)), _p(_c(pre.conic_o, dtype)), _p(_c(sre, dtype)), _p(_c(sim, dtype)),
      _p(_c(plen, dtype)), _p(_c(fwd["final_T"].reshape(-1), dtype)),
      _p(np.ascontiguousarray(fwd["n_contrib"].reshape(-1))), _p(_c(np.asarray(dout).reshape(-1), dtype)), _p(g9))
    return g9


def path_length_grad(scene, r, gL, radii):
    """d/dmu of L = |mu - tx| + |mu - rx| times gL (FP64), visible Gaussians only."""
    m = np.asarray(scene.means, np.float64)
    a = m - np.asarray(scene.tx_pos, np.float64)
    b = m - np.asarray(scene.rx_pos[r], np.float64)
    u = a / np.linalg.norm(a, axis=1, keepdims=True) + b / np.linalg.norm(b, axis=1, keepdims=True)
    return np.where((radii > 0)[:, None], gL[:, None] * u, 0.0)


def render_csi(scene, r=0, dtype=np.float32, f0=CSI_F0, df=CSI_DF):
    pre = preprocess(scene, r, dtype)
    plen = path_length(scene, r, dtype)
    keys, vals, ranges = bin_gaussians(pre, scene.W, scene.H)
    fwd = render_fwd_csi(pre, plen, vals, ranges, scene.W, scene.H, dtype, f0, df)
    return pre, plen, vals, ranges, fwd


def backward_csi(scene, pre, plen, vals, ranges, fwd, dout, r=0, dtype=np.float32, f0=CSI_F0, df=CSI_DF,
                 grads=None):
    """Exact gradients of the CSI render of receiver r (accumulated into grads)."""
    g9 = render_bwd_csi(pre, plen, vals, ranges, fwd, dout, scene.W, scene.H, dtype, f0, df)
    g = preprocess_bwd(scene, pre, g9[:, :8], GRAD_EXACT, r, dtype, grads=grads)
    g.means += path_length_grad(scene, r, g9[:, 8], pre.radii)
    return g, g9
