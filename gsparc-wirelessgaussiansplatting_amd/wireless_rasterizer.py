"""Native-level rasterizer API: drop-in for the reference's `wireless_rasterizer`
torch extension (cuda_kernel_rasterize/ext.cpp:7-11, rasterize_points.h:8-46).

    rasterize_gaussians(means3D, emission_mlps, opacities, scales, rotations,
                        image_height, image_width, scale_modifier, viewmatrix,
                        projmatrix, tx_pos, rx_pos)
        -> (num_rendered, out_real [1,H,W], out_imag [1,H,W], radii [P] int32,
            geomBuffer, binningBuffer, imgBuffer)            (rasterize_points.cu:32-102)

    rasterize_gaussians_backward(means3D, emission_mlps, radii, out_real,
                                 out_imag, opacities, scales, rotations,
                                 scale_modifier, viewmatrix, projmatrix,
                                 dL_dout_real, dL_dout_imag, tx_pos, rx_pos,
                                 geomBuffer, num_rendered, binningBuffer,
                                 imageBuffer)
        -> (dL_dmeans2D [P,3], dL_dmeans3D [P,3], dL_demission_mlps [M],
            dL_dopacity [P,1], dL_dcov3D [P,6], dL_dscales [P,3],
            dL_drotations [P,4])                                (rasterize_points.cu:106-199)

    mark_visible(means3D, viewmatrix, projmatrix) -> bool [P]   (rasterize_points.cu:203-221)

Extensions beyond the reference (keyword-only, defaults keep the reference
behaviour): `grad_mode` ("reference" reproduces the reference backward,
"exact" is the true adjoint of the forward) and `emission` (per-Gaussian
(real, imag) emission supplied directly instead of the per-Gaussian MLP).
The batched functions render R receivers in one launch sequence.

The byte buffers hold private state of this implementation (their layout is
described in csrc/gs_device.h); like the reference's they are kept by the
autograd function between forward and backward.  projmatrix is accepted and
ignored, as in the reference (it is never read by its kernels).
"""
from __future__ import annotations

import ctypes as C

import torch

import gsparc_hip as gh

U8 = torch.uint8


def _empty_bytes(device, n=0):
    return torch.empty(n, dtype=U8, device=device)


def _check_means(means3D):
    if means3D.ndimension() != 2 or means3D.size(1) != 3:
        raise RuntimeError("means3D must have dimensions (num_points, 3)")


def forward_batched(means3D, emission_mlps, opacities, scales, rotations, image_height, image_width,
                    scale_modifier, viewmatrices, tx_pos, rx_pos, emission=None, keep_keys=False, capacity=None):
    """Forward for R receivers.  viewmatrices [R,4,4] (column-major storage as
    the reference's viewmatrix), rx_pos [R,3], emission None or [R,P,2].
    capacity None: the tile lists are sized by K read back from the device;
    an int: they hold up to `capacity` instances and nothing synchronises
    (K stays on the device, see instance_count).
    Returns (K or capacity, out_real [R,H,W], out_imag [R,H,W], radii [R,P],
    geom, bin, img)."""
    _check_means(means3D)
    P = int(means3D.size(0))
    H, W = int(image_height), int(image_width)
    dev = means3D.device
    views = gh.f32(viewmatrices).reshape(-1, 16)
    rxs = gh.f32(rx_pos).reshape(-1, 3)
    R = int(rxs.size(0))
    if views.size(0) != R:
        raise RuntimeError("one view matrix per receiver position is required")
    if P == 0 or R == 0:
        z = torch.zeros((R, H, W), dtype=torch.float32, device=dev)
        return (0, z, z.clone(), torch.zeros((R, P), dtype=torch.int32, device=dev), _empty_bytes(dev),
                _empty_bytes(dev), _empty_bytes(dev))
    if scales is None or rotations is None or scales.numel() == 0 or rotations.numel() == 0:
        raise RuntimeError("scales and rotations are required")
    means = gh.f32(means3D)
    sc = gh.f32(scales)
    rot = gh.f32(rotations)
    op = gh.f32(opacities).reshape(-1)
    tx = gh.f32(tx_pos).reshape(-1)
    mlp = gh.f32_aligned(emission_mlps).reshape(-1) if emission is None else None
    em = None
    if emission is not None:
        em = gh.f32(emission).reshape(R, P, 2)
    elif mlp.numel() != P * 194:
        raise RuntimeError("emission_mlps must hold 194 floats per Gaussian")
    gh.require_cuda(means, sc, rot, op, tx, views, rxs, mlp, em)
    L = gh.lib()
    geom = _empty_bytes(dev, L.gsparc_geom_buffer_size(P, R, W, H))
    img = _empty_bytes(dev, L.gsparc_image_buffer_size(P, R, W, H))
    radii = torch.empty((R, P), dtype=torch.int32, device=dev)
    stream = gh.stream_handle(dev)
    flags = gh.FLAG_KEEP_KEYS if keep_keys else 0
    if capacity is not None:
        Kc = int(capacity)
        binbuf = _empty_bytes(dev, L.gsparc_binning_buffer_size(Kc))
        out_re = torch.empty((R, H, W), dtype=torch.float32, device=dev)
        out_im = torch.empty((R, H, W), dtype=torch.float32, device=dev)
        gh.check(L.gsparc_rasterize_forward(P, R, W, H, gh.ptr(means), gh.ptr(sc), gh.ptr(rot), gh.ptr(op),
                                           gh.ptr(mlp), gh.ptr(em), gh.ptr(views), gh.ptr(tx), gh.ptr(rxs),
                                           float(scale_modifier), gh.ptr(geom), gh.ptr(img), gh.ptr(binbuf), Kc,
                                           gh.ptr(radii), gh.ptr(out_re), gh.ptr(out_im), flags, stream),
                 "rasterize forward")
        return Kc, out_re, out_im, radii, geom, binbuf, img
    K = C.c_uint64(0)
    gh.check(L.gsparc_rasterize_forward_geom(P, R, W, H, gh.ptr(means), gh.ptr(sc), gh.ptr(rot), gh.ptr(op),
                                            gh.ptr(mlp), gh.ptr(views), gh.ptr(tx), gh.ptr(rxs),
                                            float(scale_modifier), gh.ptr(geom), gh.ptr(img), gh.ptr(radii),
                                            C.byref(K), stream), "rasterize forward (preprocess)")
    K = int(K.value)
    binbuf = _empty_bytes(dev, L.gsparc_binning_buffer_size(K))
    out_re = torch.empty((R, H, W), dtype=torch.float32, device=dev)
    out_im = torch.empty((R, H, W), dtype=torch.float32, device=dev)
    gh.check(L.gsparc_rasterize_forward_render(P, R, W, H, gh.ptr(em), gh.ptr(geom), gh.ptr(img), gh.ptr(binbuf),
                                              K, gh.ptr(radii), gh.ptr(out_re), gh.ptr(out_im), flags, stream),
             "rasterize forward (render)")
    return K, out_re, out_im, radii, geom, binbuf, img


def instance_count(P, R, image_height, image_width, imageBuffer):
    """K (tile-list instances) of the forward that filled imageBuffer;
    synchronises the current stream.  With a capacity forward, K > capacity
    means the outputs of that step were not computed."""
    K = C.c_uint64(0)
    dev = imageBuffer.device
    gh.check(gh.lib().gsparc_instance_count(int(P), int(R), int(image_width), int(image_height), gh.ptr(imageBuffer),
                                            C.byref(K), gh.stream_handle(dev)), "instance count")
    return int(K.value)


def backward_batched(means3D, emission_mlps, radii, opacities, scales, rotations, scale_modifier,
                     viewmatrices, dL_dout_real, dL_dout_imag, tx_pos, rx_pos, geomBuffer, num_rendered,
                     binningBuffer, imageBuffer, grad_mode="reference", emission=None, deterministic=False):
    """Backward for R receivers; gradients are summed over receivers.
    Returns (dmeans2D [R,P,3], dmeans3D [P,3], demission_mlps [M] or None,
    dopacity [P,1], dcov3D [P,6], dscales [P,3], drotations [P,4],
    demission [R,P,2] or None).  deterministic=True sums every gradient in a
    fixed order (bit-reproducible; 32 extra bytes of workspace per tile-list
    instance) instead of with the reference's float atomics."""
    P = int(means3D.size(0))
    R = int(dL_dout_real.size(0))
    H, W = int(dL_dout_real.size(1)), int(dL_dout_real.size(2))
    dev = means3D.device
    opts = dict(dtype=torch.float32, device=dev)
    mode = gh.grad_mode_code(grad_mode, deterministic)
    M = int(emission_mlps.numel()) if (emission_mlps is not None and emission is None) else 0
    dmeans2D = torch.zeros((R, P, 3), **opts)
    dmeans3D = torch.zeros((P, 3), **opts)
    dmlp = torch.zeros(M, **opts) if emission is None else None
    dop = torch.zeros((P, 1), **opts)
    dcov = torch.zeros((P, 6), **opts)
    dsc = torch.zeros((P, 3), **opts)
    drot = torch.zeros((P, 4), **opts)
    dem = torch.zeros((R, P, 2), **opts) if emission is not None else None
    if P == 0 or R == 0:
        return dmeans2D, dmeans3D, dmlp, dop, dcov, dsc, drot, dem
    means = gh.f32(means3D)
    sc = gh.f32(scales)
    rot = gh.f32(rotations)
    views = gh.f32(viewmatrices).reshape(-1, 16)
    rxs = gh.f32(rx_pos).reshape(-1, 3)
    tx = gh.f32(tx_pos).reshape(-1)
    mlp = gh.f32_aligned(emission_mlps).reshape(-1) if emission is None else None
    em = gh.f32(emission).reshape(R, P, 2) if emission is not None else None
    rad = radii.reshape(R, P).to(torch.int32).contiguous()
    gre = gh.f32(dL_dout_real)
    gim = gh.f32(dL_dout_imag)
    gh.require_cuda(means, sc, rot, views, rxs, tx, mlp, em, rad, gre, gim)
    L = gh.lib()
    gsize = L.gsparc_grad_buffer_size(P, R)
    if deterministic:
        gsize = L.gsparc_deterministic_grad_size(gsize, int(num_rendered))
    grad = _empty_bytes(dev, gsize)
    gh.check(L.gsparc_rasterize_backward(mode, P, R, W, H, gh.ptr(means), gh.ptr(sc), gh.ptr(rot), gh.ptr(mlp),
                                        gh.ptr(em), gh.ptr(views), gh.ptr(tx), gh.ptr(rxs), float(scale_modifier),
                                        gh.ptr(rad), gh.ptr(geomBuffer), gh.ptr(imageBuffer), gh.ptr(binningBuffer),
                                        int(num_rendered), gh.ptr(grad), gh.ptr(gre), gh.ptr(gim),
                                        gh.ptr(dmeans2D), gh.ptr(dmeans3D), gh.ptr(dmlp), gh.ptr(dop), gh.ptr(dcov),
                                        gh.ptr(dsc), gh.ptr(drot), gh.ptr(dem), gh.stream_handle(dev)),
             "rasterize backward")
    return dmeans2D, dmeans3D, dmlp, dop, dcov, dsc, drot, dem


def rasterize_gaussians(means3D, emission_mlps, opacities, scales, rotations, image_height, image_width,
                        scale_modifier, viewmatrix, projmatrix, tx_pos, rx_pos, *, emission=None):
    """One receiver, reference tuple layout (rasterize_points.cu:32-102)."""
    _check_means(means3D)
    em = None if emission is None else emission.reshape(1, -1, 2)
    K, ore, oim, radii, geom, binbuf, img = forward_batched(
        means3D, emission_mlps, opacities, scales, rotations, image_height, image_width, scale_modifier,
        viewmatrix.reshape(1, 4, 4), tx_pos, rx_pos.reshape(1, 3), emission=em)
    return K, ore, oim, radii.reshape(-1), geom, binbuf, img


def rasterize_gaussians_backward(means3D, emission_mlps, radii, out_signal_real, out_signal_imag, opacities, scales,
                                 rotations, scale_modifier, viewmatrix, projmatrix, dL_dout_signal_real,
                                 dL_dout_signal_imag, tx_pos, rx_pos, geomBuffer, R, binningBuffer, imageBuffer,
                                 *, grad_mode="reference", emission=None, deterministic=False):
    """One receiver, reference tuple layout (rasterize_points.cu:106-199); R is
    num_rendered from the forward.  deterministic: as in backward_batched."""
    P = int(means3D.size(0))
    em = None if emission is None else emission.reshape(1, -1, 2)
    d2, d3, dmlp, dop, dcov, dsc, drot, dem = backward_batched(
        means3D, emission_mlps, radii.reshape(1, -1) if P else radii, opacities, scales, rotations, scale_modifier,
        viewmatrix.reshape(1, 4, 4), dL_dout_signal_real.reshape(1, *dL_dout_signal_real.shape[-2:]),
        dL_dout_signal_imag.reshape(1, *dL_dout_signal_imag.shape[-2:]), tx_pos, rx_pos.reshape(1, 3), geomBuffer, R,
        binningBuffer, imageBuffer, grad_mode=grad_mode, emission=em, deterministic=deterministic)
    if dmlp is None:
        dmlp = dem.reshape(P, 2)
    return d2.reshape(P, 3), d3, dmlp, dop, dcov, dsc, drot


def mark_visible(means3D, viewmatrix, projmatrix):
    """bool [P]: r^2 > 0.04 in receiver space (rasterize_points.cu:203-221)."""
    P = int(means3D.size(0))
    present = torch.zeros(P, dtype=torch.bool, device=means3D.device)
    if P == 0:
        return present
    means = gh.f32(means3D)
    view = gh.f32(viewmatrix).reshape(-1)
    gh.require_cuda(means, view)
    out = torch.empty(P, dtype=torch.uint8, device=means3D.device)
    gh.check(gh.lib().gsparc_mark_visible(P, gh.ptr(means), gh.ptr(view), gh.ptr(out), gh.stream_handle(means.device)),
             "mark_visible")
    return out.bool()


def mark_visisble(means3D, viewmatrix, projmatrix):
    """The name the reference's Python wrapper calls (cuda_rasterizer/__init__.py:115)."""
    return mark_visible(means3D, viewmatrix, projmatrix)


# ---------------------------------------------------------------- uneven scenes
class SceneBatch:
    """S independent scenes in one launch sequence (no reference counterpart).
    Scene s has scene_P[s] Gaussians and scene_R[s] receivers, packed one
    scene after the other; holds the device scene table and the batch totals."""

    def __init__(self, scene_P, scene_R, device):
        scene_P = [int(p) for p in scene_P]
        scene_R = [int(r) for r in scene_R]
        if len(scene_P) != len(scene_R) or not scene_P:
            raise RuntimeError("one Gaussian count and one receiver count per scene")
        S = len(scene_P)
        L = gh.lib()
        words = L.gsparc_scene_table_words(S, sum(scene_R))
        tab = (C.c_uint32 * words)()
        self.batch = gh.Batch()
        gh.check(L.gsparc_scene_table_build(S, (C.c_int * S)(*scene_P), (C.c_int * S)(*scene_R), tab, words,
                                            C.byref(self.batch)), "scene table")
        import numpy as np
        self.table = torch.from_numpy(np.frombuffer(tab, dtype=np.int32).copy()).to(device)
        self.scene_P, self.scene_R = scene_P, scene_R
        self.S, self.R, self.Ptot, self.Nrec = S, self.batch.R, self.batch.Ptot, int(self.batch.Nrec)

    def sizes(self, H, W, K=0):
        out = (C.c_size_t * 4)()
        gh.check(gh.lib().gsparc_scene_buffer_sizes(C.byref(self.batch), int(W), int(H), int(K), out), "sizes")
        return list(out)


def pack_mlps(mlps):
    """Per-scene flat [P_s*194] MLP arrays -> the packed [Ptot*194] layout
    (each of the W1 | b1 | W2 | b2 sections holds all scenes in order)."""
    cuts = ((0, 96), (96, 128), (128, 192), (192, 194))
    parts = []
    for a, b in cuts:
        for m in mlps:
            P = m.numel() // 194
            parts.append(m.reshape(-1)[P * a:P * b])
    return torch.cat(parts)


def forward_scenes(means3D, emission_mlps, opacities, scales, rotations, image_height, image_width, scale_modifier,
                   viewmatrices, tx_pos, rx_pos, scenes: SceneBatch, emission=None, capacity=None):
    """Forward of a batch of uneven scenes.  means etc. [Ptot,...] packed by
    scene; viewmatrices [R,4,4], rx_pos [R,3] packed by scene; tx_pos [S,3];
    emission None or [Nrec,2].  Returns (K or capacity, out_real [R,H,W],
    out_imag [R,H,W], radii [Nrec], geom, bin, img)."""
    _check_means(means3D)
    H, W = int(image_height), int(image_width)
    dev = means3D.device
    if int(means3D.size(0)) != scenes.Ptot:
        raise RuntimeError("means3D must hold every scene's Gaussians")
    views = gh.f32(viewmatrices).reshape(-1, 16)
    rxs = gh.f32(rx_pos).reshape(-1, 3)
    tx = gh.f32(tx_pos).reshape(-1, 3)
    if views.size(0) != scenes.R or rxs.size(0) != scenes.R or tx.size(0) != scenes.S:
        raise RuntimeError("one view matrix and position per receiver, one TX position per scene")
    means, sc, rot = gh.f32(means3D), gh.f32(scales), gh.f32(rotations)
    op = gh.f32(opacities).reshape(-1)
    mlp = gh.f32_aligned(emission_mlps).reshape(-1) if emission is None else None
    em = gh.f32(emission).reshape(-1, 2) if emission is not None else None
    if em is None and mlp.numel() != scenes.Ptot * 194:
        raise RuntimeError("emission_mlps must hold 194 floats per Gaussian")
    gh.require_cuda(means, sc, rot, op, tx, views, rxs, mlp, em)
    L = gh.lib()
    geom_b, img_b, _, _ = scenes.sizes(H, W)
    geom = _empty_bytes(dev, geom_b)
    img = _empty_bytes(dev, img_b)
    radii = torch.empty(scenes.Nrec, dtype=torch.int32, device=dev)
    out_re = torch.empty((scenes.R, H, W), dtype=torch.float32, device=dev)
    out_im = torch.empty((scenes.R, H, W), dtype=torch.float32, device=dev)
    stream = gh.stream_handle(dev)
    bp = C.byref(scenes.batch)
    if capacity is not None:
        Kc = int(capacity)
        binbuf = _empty_bytes(dev, scenes.sizes(H, W, Kc)[2])
        gh.check(L.gsparc_rasterize_forward_scenes(bp, gh.ptr(scenes.table), W, H, gh.ptr(means), gh.ptr(sc),
                                                  gh.ptr(rot), gh.ptr(op), gh.ptr(mlp), gh.ptr(em), gh.ptr(views),
                                                  gh.ptr(tx), gh.ptr(rxs), float(scale_modifier), gh.ptr(geom),
                                                  gh.ptr(img), gh.ptr(binbuf), Kc, gh.ptr(radii), gh.ptr(out_re),
                                                  gh.ptr(out_im), 0, stream), "rasterize forward (scenes)")
        return Kc, out_re, out_im, radii, geom, binbuf, img
    K = C.c_uint64(0)
    gh.check(L.gsparc_rasterize_forward_geom_scenes(bp, gh.ptr(scenes.table), W, H, gh.ptr(means), gh.ptr(sc),
                                                   gh.ptr(rot), gh.ptr(op), gh.ptr(mlp), gh.ptr(views), gh.ptr(tx),
                                                   gh.ptr(rxs), float(scale_modifier), gh.ptr(geom), gh.ptr(img),
                                                   gh.ptr(radii), C.byref(K), stream), "rasterize forward (scenes)")
    K = int(K.value)
    binbuf = _empty_bytes(dev, scenes.sizes(H, W, K)[2])
    gh.check(L.gsparc_rasterize_forward_render_scenes(bp, gh.ptr(scenes.table), W, H, gh.ptr(em), gh.ptr(geom),
                                                     gh.ptr(img), gh.ptr(binbuf), K, gh.ptr(radii), gh.ptr(out_re),
                                                     gh.ptr(out_im), 0, stream), "rasterize forward (scenes)")
    return K, out_re, out_im, radii, geom, binbuf, img


def backward_scenes(means3D, emission_mlps, radii, opacities, scales, rotations, scale_modifier, viewmatrices,
                    dL_dout_real, dL_dout_imag, tx_pos, rx_pos, scenes: SceneBatch, geomBuffer, num_rendered,
                    binningBuffer, imageBuffer, grad_mode="reference", emission=None, deterministic=False):
    """Backward of forward_scenes.  Returns (dmeans2D [Nrec,3], dmeans3D
    [Ptot,3], demission_mlps [Ptot*194] or None, dopacity [Ptot,1], dcov3D
    [Ptot,6], dscales [Ptot,3], drotations [Ptot,4], demission [Nrec,2] or None).
    deterministic: as in backward_batched."""
    H, W = int(dL_dout_real.size(1)), int(dL_dout_real.size(2))
    dev = means3D.device
    opts = dict(dtype=torch.float32, device=dev)
    mode = gh.grad_mode_code(grad_mode, deterministic)
    P = scenes.Ptot
    dmeans2D = torch.zeros((scenes.Nrec, 3), **opts)
    dmeans3D = torch.zeros((P, 3), **opts)
    dmlp = torch.zeros(P * 194, **opts) if emission is None else None
    dop = torch.zeros((P, 1), **opts)
    dcov = torch.zeros((P, 6), **opts)
    dsc = torch.zeros((P, 3), **opts)
    drot = torch.zeros((P, 4), **opts)
    dem = torch.zeros((scenes.Nrec, 2), **opts) if emission is not None else None
    means, sc, rot = gh.f32(means3D), gh.f32(scales), gh.f32(rotations)
    views = gh.f32(viewmatrices).reshape(-1, 16)
    rxs = gh.f32(rx_pos).reshape(-1, 3)
    tx = gh.f32(tx_pos).reshape(-1, 3)
    mlp = gh.f32_aligned(emission_mlps).reshape(-1) if emission is None else None
    em = gh.f32(emission).reshape(-1, 2) if emission is not None else None
    rad = radii.reshape(-1).to(torch.int32).contiguous()
    gre, gim = gh.f32(dL_dout_real), gh.f32(dL_dout_imag)
    gh.require_cuda(means, sc, rot, views, rxs, tx, mlp, em, rad, gre, gim)
    gsize = scenes.sizes(H, W)[3]
    if deterministic:
        gsize = gh.lib().gsparc_deterministic_grad_size(gsize, int(num_rendered))
    grad = _empty_bytes(dev, gsize)
    gh.check(gh.lib().gsparc_rasterize_backward_scenes(
        mode, C.byref(scenes.batch), gh.ptr(scenes.table), W, H, gh.ptr(means), gh.ptr(sc), gh.ptr(rot),
        gh.ptr(mlp), gh.ptr(em), gh.ptr(views), gh.ptr(tx), gh.ptr(rxs), float(scale_modifier), gh.ptr(rad),
        gh.ptr(geomBuffer), gh.ptr(imageBuffer), gh.ptr(binningBuffer), int(num_rendered), gh.ptr(grad),
        gh.ptr(gre), gh.ptr(gim), gh.ptr(dmeans2D), gh.ptr(dmeans3D), gh.ptr(dmlp), gh.ptr(dop), gh.ptr(dcov),
        gh.ptr(dsc), gh.ptr(drot), gh.ptr(dem), gh.stream_handle(dev)), "rasterize backward (scenes)")
    return dmeans2D, dmeans3D, dmlp, dop, dcov, dsc, drot, dem


def unpack_mlps(packed, scene_P):
    """Inverse of pack_mlps: [Ptot*194] -> one flat [P_s*194] array per scene."""
    Ptot = sum(int(p) for p in scene_P)
    out = [[] for _ in scene_P]
    for a, b in ((0, 96), (96, 128), (128, 192), (192, 194)):
        off = Ptot * a
        for s, P in enumerate(scene_P):
            n = int(P) * (b - a)
            out[s].append(packed[off:off + n])
            off += n
    return [torch.cat(o) for o in out]


def autotune(forward_out, grad_mode="reference", iters=5, emission=None, scenes: SceneBatch | None = None):
    """Tune the blend kernels' mapping (pixels per lane and LDS batch depth
    of the forward and backward blend, work order) on the buffers of a
    finished forward, the tuple returned by forward_batched or
    forward_scenes.  Later calls with a batch of the same shape use the
    choice.  Returns {"fwd_ppl", "bwd_ppl", "order", "fwd_depth",
    "bwd_depth", "fwd_ms", "bwd_ms", "times_ms"} (every candidate's median
    time)."""
    K, ore, oim, radii, geom, binbuf, img = forward_out
    R, H, W = int(ore.size(0)), int(ore.size(1)), int(ore.size(2))
    dev = ore.device
    L = gh.lib()
    mode = gh.GRAD_MODES[grad_mode] if isinstance(grad_mode, str) else int(grad_mode)
    g = torch.Generator(device=dev).manual_seed(0)
    dre = torch.randn((R, H, W), generator=g, device=dev)
    dim = torch.randn((R, H, W), generator=g, device=dev)
    out_re, out_im = torch.empty_like(ore), torch.empty_like(oim)
    em = gh.f32(emission).reshape(-1, 2) if emission is not None else None
    choice = (C.c_int * 5)()
    ms = (C.c_float * 2)()
    table = (C.c_float * 54)()
    rad = radii.reshape(-1).to(torch.int32).contiguous()
    stream = gh.stream_handle(dev)
    if scenes is None:
        P = int(radii.shape[-1]) if radii.dim() == 2 else int(radii.numel()) // max(R, 1)
        grad = _empty_bytes(dev, L.gsparc_grad_buffer_size(P, R))
        gh.check(L.gsparc_autotune(mode, P, R, W, H, gh.ptr(em), gh.ptr(rad), gh.ptr(geom), gh.ptr(img),
                                   gh.ptr(binbuf), int(K), gh.ptr(grad), gh.ptr(dre), gh.ptr(dim), gh.ptr(out_re),
                                   gh.ptr(out_im), int(iters), choice, ms, table, stream), "autotune")
    else:
        grad = _empty_bytes(dev, scenes.sizes(H, W)[3])
        gh.check(L.gsparc_autotune_scenes(mode, C.byref(scenes.batch), gh.ptr(scenes.table), W, H, gh.ptr(em),
                                          gh.ptr(rad), gh.ptr(geom), gh.ptr(img), gh.ptr(binbuf), int(K),
                                          gh.ptr(grad), gh.ptr(dre), gh.ptr(dim), gh.ptr(out_re), gh.ptr(out_im),
                                          int(iters), choice, ms, table, stream), "autotune")
    times = {}
    for o in range(3):
        for k in range(9):
            ppl, depth = (1, 2, 4)[k // 3], (32, 64, 128)[k % 3]
            times[f"order{o + 1}_fwd_ppl{ppl}_d{depth}"] = round(table[o * 18 + k], 4)
            times[f"order{o + 1}_bwd_ppl{ppl}_d{depth}"] = round(table[o * 18 + 9 + k], 4)
    return {"fwd_ppl": choice[0], "bwd_ppl": choice[1], "order": choice[2], "fwd_depth": choice[3],
            "bwd_depth": choice[4], "fwd_ms": round(ms[0], 4), "bwd_ms": round(ms[1], 4), "times_ms": times}


def autotune_clear():
    """Forget every tuned mapping (the built-in rule applies again)."""
    gh.check(gh.lib().gsparc_autotune_clear(), "autotune clear")


# ------------------------------------------------------- multi-frequency CSI
CSI_NC = 52
CSI_F0 = 3.5e9 - 25.5 * 312.5e3      # Hz, subcarrier 0
CSI_DF = 312.5e3                     # Hz, spacing


def forward_csi(means3D, emission_mlps, opacities, scales, rotations, image_height, image_width, scale_modifier,
                viewmatrices, tx_pos, rx_pos, emission=None, capacity=None, f0=CSI_F0, df=CSI_DF):
    """Multi-frequency forward for R receivers: 52 subcarriers, record (r, g)
    emitting s * exp(-j 2 pi f_k L / c) with s its MLP output (or `emission`
    [R,P,2]) and L = |tx - mu| + |mu - rx_r|.  Returns (K or capacity, out
    [R,H,W,52,2], radii [R,P], geom, bin, img)."""
    _check_means(means3D)
    P = int(means3D.size(0))
    H, W = int(image_height), int(image_width)
    dev = means3D.device
    views = gh.f32(viewmatrices).reshape(-1, 16)
    rxs = gh.f32(rx_pos).reshape(-1, 3)
    R = int(rxs.size(0))
    if views.size(0) != R:
        raise RuntimeError("one view matrix per receiver position is required")
    out = torch.zeros((R, H, W, CSI_NC, 2), dtype=torch.float32, device=dev)
    if P == 0 or R == 0:
        return (0, out, torch.zeros((R, P), dtype=torch.int32, device=dev), _empty_bytes(dev), _empty_bytes(dev),
                _empty_bytes(dev))
    means, sc, rot = gh.f32(means3D), gh.f32(scales), gh.f32(rotations)
    op = gh.f32(opacities).reshape(-1)
    tx = gh.f32(tx_pos).reshape(-1)
    mlp = gh.f32_aligned(emission_mlps).reshape(-1) if emission is None else None
    em = gh.f32(emission).reshape(R, P, 2) if emission is not None else None
    if em is None and mlp.numel() != P * 194:
        raise RuntimeError("emission_mlps must hold 194 floats per Gaussian")
    gh.require_cuda(means, sc, rot, op, tx, views, rxs, mlp, em)
    L = gh.lib()
    geom = _empty_bytes(dev, L.gsparc_geom_buffer_size(P, R, W, H))
    img = _empty_bytes(dev, L.gsparc_image_buffer_size(P, R, W, H))
    radii = torch.empty((R, P), dtype=torch.int32, device=dev)
    stream = gh.stream_handle(dev)
    if capacity is not None:
        Kc = int(capacity)
        binbuf = _empty_bytes(dev, L.gsparc_binning_buffer_size(Kc))
        gh.check(L.gsparc_rasterize_forward_csi(P, R, W, H, float(f0), float(df), gh.ptr(means), gh.ptr(sc),
                                               gh.ptr(rot), gh.ptr(op), gh.ptr(mlp), gh.ptr(em), gh.ptr(views),
                                               gh.ptr(tx), gh.ptr(rxs), float(scale_modifier), gh.ptr(geom),
                                               gh.ptr(img), gh.ptr(binbuf), Kc, gh.ptr(radii), gh.ptr(out), stream),
                 "rasterize forward (csi)")
        return Kc, out, radii, geom, binbuf, img
    K = C.c_uint64(0)
    gh.check(L.gsparc_rasterize_forward_geom(P, R, W, H, gh.ptr(means), gh.ptr(sc), gh.ptr(rot), gh.ptr(op),
                                            gh.ptr(mlp), gh.ptr(views), gh.ptr(tx), gh.ptr(rxs),
                                            float(scale_modifier), gh.ptr(geom), gh.ptr(img), gh.ptr(radii),
                                            C.byref(K), stream), "rasterize forward (csi preprocess)")
    K = int(K.value)
    binbuf = _empty_bytes(dev, L.gsparc_binning_buffer_size(K))
    gh.check(L.gsparc_rasterize_forward_render_csi(P, R, W, H, float(f0), float(df), gh.ptr(em), gh.ptr(geom),
                                                  gh.ptr(img), gh.ptr(binbuf), K, gh.ptr(radii), gh.ptr(out), stream),
             "rasterize forward (csi render)")
    return K, out, radii, geom, binbuf, img


def backward_csi(means3D, emission_mlps, radii, opacities, scales, rotations, scale_modifier, viewmatrices,
                 dL_dout, tx_pos, rx_pos, geomBuffer, num_rendered, binningBuffer, imageBuffer, emission=None,
                 f0=CSI_F0, df=CSI_DF):
    """Exact adjoint of forward_csi.  dL_dout [R,H,W,52,2].  Returns (dmeans2D
    [R,P,3], dmeans3D [P,3], demission_mlps [M] or None, dopacity [P,1],
    dcov3D [P,6], dscales [P,3], drotations [P,4], demission [R,P,2] or None)."""
    P = int(means3D.size(0))
    R = int(dL_dout.size(0))
    dev = means3D.device
    opts = dict(dtype=torch.float32, device=dev)
    dmeans2D = torch.zeros((R, P, 3), **opts)
    dmeans3D = torch.zeros((P, 3), **opts)
    dmlp = torch.zeros(P * 194, **opts) if emission is None else None
    dop = torch.zeros((P, 1), **opts)
    dcov = torch.zeros((P, 6), **opts)
    dsc = torch.zeros((P, 3), **opts)
    drot = torch.zeros((P, 4), **opts)
    dem = torch.zeros((R, P, 2), **opts) if emission is not None else None
    if P == 0 or R == 0:
        return dmeans2D, dmeans3D, dmlp, dop, dcov, dsc, drot, dem
    H, W = int(dL_dout.size(1)), int(dL_dout.size(2))
    means, sc, rot = gh.f32(means3D), gh.f32(scales), gh.f32(rotations)
    views = gh.f32(viewmatrices).reshape(-1, 16)
    rxs = gh.f32(rx_pos).reshape(-1, 3)
    tx = gh.f32(tx_pos).reshape(-1)
    mlp = gh.f32_aligned(emission_mlps).reshape(-1) if emission is None else None
    em = gh.f32(emission).reshape(R, P, 2) if emission is not None else None
    rad = radii.reshape(R, P).to(torch.int32).contiguous()
    g = gh.f32(dL_dout)
    gh.require_cuda(means, sc, rot, views, rxs, tx, mlp, em, rad, g)
    L = gh.lib()
    grad = _empty_bytes(dev, L.gsparc_grad_buffer_size(P, R))
    gh.check(L.gsparc_rasterize_backward_csi(
        P, R, W, H, float(f0), float(df), gh.ptr(means), gh.ptr(sc), gh.ptr(rot), gh.ptr(mlp), gh.ptr(em),
        gh.ptr(views), gh.ptr(tx), gh.ptr(rxs), float(scale_modifier), gh.ptr(rad), gh.ptr(geomBuffer),
        gh.ptr(imageBuffer), gh.ptr(binningBuffer), int(num_rendered), gh.ptr(grad), gh.ptr(g), gh.ptr(dmeans2D),
        gh.ptr(dmeans3D), gh.ptr(dmlp), gh.ptr(dop), gh.ptr(dcov), gh.ptr(dsc), gh.ptr(drot), gh.ptr(dem),
        gh.stream_handle(dev)), "rasterize backward (csi)")
    return dmeans2D, dmeans3D, dmlp, dop, dcov, dsc, drot, dem
