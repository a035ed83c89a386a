"""Python rasterizer API: drop-in for the reference's
cuda_kernel_rasterize/cuda_rasterizer/__init__.py (lines 12-131).

Same names, argument order, outputs and gradient order:
  rasterize_gaussians(means3D, means2D, emission_mlps, opacities, scales,
                      rotations, raster_settings) -> (out_real, out_imag, radii)
  _RasterizeGaussians.backward returns grads for (means3D, means2D,
  emission_mlps, opacities, scales, rotations, raster_settings)
  GaussianRasterizerSetting(image_height, image_width, scale_modifier,
                            viewmatrix, projmatrix, tx_pos, rx_pos)
  GaussianRasterizer(raster_settings)(means3D, means2D, emission_mlps,
                                      opacities, scales=None, rotations=None)

Additions: GaussianRasterizerSetting.grad_mode ("reference", the default,
reproduces the reference backward; "exact" is the true adjoint),
GaussianRasterizerSetting.deterministic (False, the default, sums the
gradients with float atomics as the reference does; True sums them in a
fixed order, bit-reproducible from run to run), the
`rasterize_gaussians_batched` function for many receivers per launch, and
`emission=` for precomputed per-Gaussian emission.
"""
from __future__ import annotations

from typing import NamedTuple, Optional

import torch
import torch.nn as nn

import wireless_rasterizer as wr


def rasterize_gaussians(means3D, means2D, emission_mlps, opacities, scales, rotations, raster_settings):
    return _RasterizeGaussians.apply(means3D, means2D, emission_mlps, opacities, scales, rotations,
                                     raster_settings)


class _RasterizeGaussians(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means3D, means2D, emission_mlps, opacities, scales, rotations, raster_settings):
        s = raster_settings
        num_rendered, out_real, out_imag, radii, geom, binbuf, img = wr.rasterize_gaussians(
            means3D, emission_mlps, opacities, scales, rotations, s.image_height, s.image_width, s.scale_modifier,
            s.viewmatrix, s.projmatrix, s.tx_pos, s.rx_pos)
        ctx.raster_settings = s
        ctx.num_rendered = num_rendered
        ctx.save_for_backward(emission_mlps, out_real, out_imag, means3D, scales, rotations, opacities, radii, geom,
                              binbuf, img)
        ctx.mark_non_differentiable(radii)
        return out_real, out_imag, radii

    @staticmethod
    def backward(ctx, grad_out_real, grad_out_imag, grad_radii):
        s = ctx.raster_settings
        (emission_mlps, out_real, out_imag, means3D, scales, rotations, opacities, radii, geom, binbuf,
         img) = ctx.saved_tensors
        g2, g3, gmlp, gop, _gcov, gsc, grot = wr.rasterize_gaussians_backward(
            means3D, emission_mlps, radii, out_real, out_imag, opacities, scales, rotations, s.scale_modifier,
            s.viewmatrix, s.projmatrix, grad_out_real, grad_out_imag, s.tx_pos, s.rx_pos, geom, ctx.num_rendered,
            binbuf, img, grad_mode=getattr(s, "grad_mode", "reference"),
            deterministic=bool(getattr(s, "deterministic", False)))
        return g3, g2, gmlp.reshape(emission_mlps.shape), gop.reshape(opacities.shape), gsc, grot, None


class GaussianRasterizerSetting(NamedTuple):
    image_height: int
    image_width: int
    scale_modifier: float
    viewmatrix: torch.Tensor
    projmatrix: torch.Tensor
    tx_pos: torch.Tensor
    rx_pos: torch.Tensor
    grad_mode: str = "reference"
    deterministic: bool = False


class GaussianRasterizer(nn.Module):
    def __init__(self, raster_settings):
        super().__init__()
        self.raster_settings = raster_settings

    def markVisible(self, positions):
        with torch.no_grad():
            s = self.raster_settings
            return wr.mark_visible(positions, s.viewmatrix, s.projmatrix)

    def forward(self, means3D, means2D, emission_mlps, opacities, scales=None, rotations=None):
        s = self.raster_settings
        if scales is None:
            scales = torch.Tensor([])
        if rotations is None:
            rotations = torch.Tensor([])
        return rasterize_gaussians(means3D, means2D, emission_mlps, opacities, scales, rotations, s)


def _refuse_torch_capture(what):
    """The autograd forms below are not for torch.cuda.graph capture: the one
    capture of a whole autograd step through them ended in a host
    segmentation fault inside torch's CUDAGraph.capture_end, with its cause
    not established.  A captured training step is built by
    gsparc_train.GraphedTrainingStep (GraphedSceneStep for uneven scenes),
    whose library-owned HIP graph records forward, loss and backward and is
    tested against the eager step."""
    if torch.cuda.is_available() and torch.cuda.is_current_stream_capturing():
        raise RuntimeError(f"{what} cannot be recorded by torch.cuda.graph; capture a training step with "
                           "gsparc_train.GraphedTrainingStep (or GraphedSceneStep) instead")


# ------------------------------------------------------------------ batched
class BatchedSetting(NamedTuple):
    """R receivers rendered in one launch sequence (no reference counterpart:
    the reference renders one receiver per call, renderer.py:39-88)."""
    image_height: int
    image_width: int
    scale_modifier: float
    viewmatrices: torch.Tensor      # [R,4,4]
    tx_pos: torch.Tensor            # [3]
    rx_pos: torch.Tensor            # [R,3]
    grad_mode: str = "reference"
    # 0: size the tile lists exactly (one host synchronisation per forward);
    # > 0: a fixed instance capacity and no host synchronisation.  Check the
    # capacity afterwards with wireless_rasterizer.instance_count(); on an
    # overflow the spectra are NaN and the gradients zero.  These autograd
    # forms refuse torch.cuda.graph capture: a captured step is
    # gsparc_train.GraphedTrainingStep.
    capacity: int = 0
    # fixed-order gradient sums (bit-reproducible) instead of float atomics
    deterministic: bool = False


class _RasterizeBatched(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means3D, emission_mlps, emission, opacities, scales, rotations, settings):
        _refuse_torch_capture("rasterize_gaussians_batched")
        s = settings
        K, ore, oim, radii, geom, binbuf, img = wr.forward_batched(
            means3D, emission_mlps, opacities, scales, rotations, s.image_height, s.image_width, s.scale_modifier,
            s.viewmatrices, s.tx_pos, s.rx_pos, emission=emission, capacity=s.capacity or None)
        ctx.s = s
        ctx.K = K
        ctx.has_emission = emission is not None
        ctx.save_for_backward(means3D, emission_mlps if emission is None else emission, opacities, scales, rotations,
                              radii, geom, binbuf, img)
        ctx.mark_non_differentiable(radii)
        return ore, oim, radii

    @staticmethod
    def backward(ctx, gre, gim, _gr):
        s = ctx.s
        means3D, em_or_mlp, opacities, scales, rotations, radii, geom, binbuf, img = ctx.saved_tensors
        mlp = None if ctx.has_emission else em_or_mlp
        em = em_or_mlp if ctx.has_emission else None
        d2, d3, dmlp, dop, _dcov, dsc, drot, dem = wr.backward_batched(
            means3D, mlp, radii, opacities, scales, rotations, s.scale_modifier, s.viewmatrices, gre, gim, s.tx_pos,
            s.rx_pos, geom, ctx.K, binbuf, img, grad_mode=s.grad_mode, emission=em, deterministic=s.deterministic)
        return (d3, None if dmlp is None else dmlp.reshape(mlp.shape), None if dem is None else dem.reshape(em.shape),
                dop.reshape(opacities.shape), dsc, drot, None)


def rasterize_gaussians_batched(means3D, emission_mlps, opacities, scales, rotations, settings: BatchedSetting,
                                emission: Optional[torch.Tensor] = None):
    """Render R receivers: returns (out_real [R,H,W], out_imag [R,H,W], radii [R,P]).
    Pass emission_mlps (per-Gaussian 3->32->2 MLPs) or emission [R,P,2]."""
    return _RasterizeBatched.apply(means3D, emission_mlps, emission, opacities, scales, rotations, settings)


# ------------------------------------------------------------------ uneven scenes
class SceneSetting(NamedTuple):
    """A batch of uneven scenes in one launch sequence (no reference
    counterpart).  scenes: wireless_rasterizer.SceneBatch; viewmatrices
    [R,4,4] and rx_pos [R,3] packed by scene; tx_pos [S,3]."""
    image_height: int
    image_width: int
    scale_modifier: float
    scenes: object
    viewmatrices: torch.Tensor
    tx_pos: torch.Tensor
    rx_pos: torch.Tensor
    grad_mode: str = "reference"
    capacity: int = 0
    deterministic: bool = False


class _RasterizeScenes(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means3D, emission_mlps, emission, opacities, scales, rotations, settings):
        _refuse_torch_capture("rasterize_scenes")
        s = settings
        K, ore, oim, radii, geom, binbuf, img = wr.forward_scenes(
            means3D, emission_mlps, opacities, scales, rotations, s.image_height, s.image_width, s.scale_modifier,
            s.viewmatrices, s.tx_pos, s.rx_pos, s.scenes, emission=emission, capacity=s.capacity or None)
        ctx.s = s
        ctx.K = K
        ctx.has_emission = emission is not None
        ctx.save_for_backward(means3D, emission_mlps if emission is None else emission, opacities, scales, rotations,
                              radii, geom, binbuf, img)
        ctx.mark_non_differentiable(radii)
        return ore, oim, radii

    @staticmethod
    def backward(ctx, gre, gim, _gr):
        s = ctx.s
        means3D, em_or_mlp, opacities, scales, rotations, radii, geom, binbuf, img = ctx.saved_tensors
        mlp = None if ctx.has_emission else em_or_mlp
        em = em_or_mlp if ctx.has_emission else None
        d2, d3, dmlp, dop, _dcov, dsc, drot, dem = wr.backward_scenes(
            means3D, mlp, radii, opacities, scales, rotations, s.scale_modifier, s.viewmatrices, gre, gim, s.tx_pos,
            s.rx_pos, s.scenes, geom, ctx.K, binbuf, img, grad_mode=s.grad_mode, emission=em,
            deterministic=s.deterministic)
        return (d3, None if dmlp is None else dmlp.reshape(mlp.shape), None if dem is None else dem.reshape(em.shape),
                dop.reshape(opacities.shape), dsc, drot, None)


def rasterize_scenes(means3D, emission_mlps, opacities, scales, rotations, settings: SceneSetting,
                     emission: Optional[torch.Tensor] = None):
    """Render a batch of uneven scenes: (out_real [R,H,W], out_imag [R,H,W],
    radii [Nrec]).  Pass emission_mlps (packed per-Gaussian MLPs, see
    wireless_rasterizer.pack_mlps) or emission [Nrec,2]."""
    return _RasterizeScenes.apply(means3D, emission_mlps, emission, opacities, scales, rotations, settings)


# ----------------------------------------------------- multi-frequency CSI
class CsiSetting(NamedTuple):
    """Multi-frequency render of R receivers (no reference counterpart):
    viewmatrices [R,4,4], tx_pos [3], rx_pos [R,3]; 52 subcarriers from f0
    spaced df (Hz).  The backward is the exact adjoint."""
    image_height: int
    image_width: int
    scale_modifier: float
    viewmatrices: torch.Tensor
    tx_pos: torch.Tensor
    rx_pos: torch.Tensor
    f0: float = wr.CSI_F0
    df: float = wr.CSI_DF
    capacity: int = 0


class _RasterizeCsi(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means3D, emission_mlps, emission, opacities, scales, rotations, settings):
        _refuse_torch_capture("rasterize_csi")
        s = settings
        K, out, radii, geom, binbuf, img = wr.forward_csi(
            means3D, emission_mlps, opacities, scales, rotations, s.image_height, s.image_width, s.scale_modifier,
            s.viewmatrices, s.tx_pos, s.rx_pos, emission=emission, capacity=s.capacity or None, f0=s.f0, df=s.df)
        ctx.s = s
        ctx.K = K
        ctx.has_emission = emission is not None
        ctx.save_for_backward(means3D, emission_mlps if emission is None else emission, opacities, scales, rotations,
                              radii, geom, binbuf, img)
        ctx.mark_non_differentiable(radii)
        return out, radii

    @staticmethod
    def backward(ctx, gout, _gr):
        s = ctx.s
        means3D, em_or_mlp, opacities, scales, rotations, radii, geom, binbuf, img = ctx.saved_tensors
        mlp = None if ctx.has_emission else em_or_mlp
        em = em_or_mlp if ctx.has_emission else None
        d2, d3, dmlp, dop, _dcov, dsc, drot, dem = wr.backward_csi(
            means3D, mlp, radii, opacities, scales, rotations, s.scale_modifier, s.viewmatrices, gout, s.tx_pos,
            s.rx_pos, geom, ctx.K, binbuf, img, emission=em, f0=s.f0, df=s.df)
        return (d3, None if dmlp is None else dmlp.reshape(mlp.shape), None if dem is None else dem.reshape(em.shape),
                dop.reshape(opacities.shape), dsc, drot, None)


def rasterize_csi(means3D, emission_mlps, opacities, scales, rotations, settings: CsiSetting,
                  emission: Optional[torch.Tensor] = None):
    """CSI maps (out [R,H,W,52,2], radii [R,P]) with autograd."""
    return _RasterizeCsi.apply(means3D, emission_mlps, emission, opacities, scales, rotations, settings)
