"""ctypes binding of libgsparc.so (the C ABI declared in include/gsparc.h).

Loads the in-tree gfx950 library and fails loudly when it is missing: there is
no CPU or eager fallback anywhere in the product path.  PyTorch provides device
memory and the current HIP stream; the kernels are the library's.
"""
from __future__ import annotations

import ctypes as C
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsparc.so")

GRAD_EXACT = 0
GRAD_REFERENCE = 1
GRAD_MODES = {"exact": GRAD_EXACT, "reference": GRAD_REFERENCE}
GRAD_DETERMINISTIC = 16   # OR-ed into a grad mode: fixed summation order, bit-reproducible gradients
FLAG_KEEP_KEYS = 1


def grad_mode_code(grad_mode, deterministic=False):
    """The C ABI's grad_mode for "exact" / "reference" (or an int), with the
    deterministic bit when asked."""
    m = GRAD_MODES[grad_mode] if isinstance(grad_mode, str) else int(grad_mode)
    return m | GRAD_DETERMINISTIC if deterministic else m

_vp = C.c_void_p
_i = C.c_int
_f = C.c_float
_u64 = C.c_uint64
_sz = C.c_size_t


class Layout(C.Structure):
    _fields_ = [(n, C.c_size_t) for n in (
        "splat", "sig", "depth", "cov3", "sort_key0", "sort_val0", "sort_key1", "sort_val1", "sort_hist", "nvis",
        "plen", "rect", "srect", "geom_bytes",
        "final_t", "n_contrib", "ranges", "segcnt", "segbase", "tile_order", "chunk_cnt", "total", "image_bytes",
        "plist", "keys_sorted", "bin_bytes")]


class Batch(C.Structure):
    """gsparc_batch: totals of a batch of uneven scenes."""
    _fields_ = [("S", C.c_int), ("R", C.c_int), ("G", C.c_int), ("Pmax", C.c_int), ("Ptot", C.c_int),
                ("Nrec", C.c_uint64), ("Npart", C.c_uint64)]


_BP = C.POINTER(Batch)


class GaussiansC(C.Structure):
    """gsparc_gaussians: the device arrays of a training run's Gaussian set."""
    _fields_ = [(n, C.c_void_p) for n in (
        "centre", "log_scale", "rotation", "logit", "emission_mlps", "scale", "rotation_n", "opacity", "state",
        "state_mlp", "accum", "denom", "max_radius")]


class AdamConfig(C.Structure):
    """gsparc_adam_config."""
    _fields_ = [("beta1", C.c_double), ("beta2", C.c_double), ("eps", C.c_double),
                ("lr_centre_init", C.c_double), ("lr_centre_final", C.c_double),
                ("lr_centre_delay_mult", C.c_double), ("lr_centre_delay_steps", C.c_int),
                ("lr_centre_max_steps", C.c_int), ("lr_scaling", C.c_double), ("lr_rotation", C.c_double),
                ("lr_opacity", C.c_double), ("lr_emission", C.c_double), ("lr_confidence", C.c_double),
                ("clamp_centres", C.c_int), ("scene_min", C.c_float * 3), ("scene_max", C.c_float * 3)]


class DensifyConfig(C.Structure):
    """gsparc_densify_config."""
    _fields_ = [("densify", C.c_int), ("grad_threshold", C.c_float), ("clone_max_scale", C.c_float),
                ("min_opacity", C.c_float), ("prune_max_scale", C.c_float), ("use_bounds", C.c_int),
                ("scene_min", C.c_float * 3), ("scene_max", C.c_float * 3)]


GROUP_CENTRE, GROUP_SCALING, GROUP_ROTATION, GROUP_OPACITY, GROUP_EMISSION, GROUP_CONFIDENCE = 1, 2, 4, 8, 16, 32
GROUPS_ALL = 63
TRAIN_STATE_BYTES = 256
ERR_CAPACITY = 3

_GP = C.POINTER(GaussiansC)

_SIGS = {
    "gsparc_version": (C.c_char_p, []),
    "gsparc_last_error": (C.c_char_p, []),
    "gsparc_set_option": (_i, [C.c_char_p, _i]),
    "gsparc_get_option": (_i, [C.c_char_p, C.POINTER(_i)]),
    "gsparc_geom_buffer_size": (_sz, [_i, _i, _i, _i]),
    "gsparc_image_buffer_size": (_sz, [_i, _i, _i, _i]),
    "gsparc_binning_buffer_size": (_sz, [_u64]),
    "gsparc_grad_buffer_size": (_sz, [_i, _i]),
    "gsparc_deterministic_grad_size": (_sz, [_sz, _u64]),
    "gsparc_layout_query": (_i, [_i, _i, _i, _i, _u64, C.POINTER(Layout)]),
    "gsparc_rasterize_forward_geom": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f, _vp, _vp,
                                           _vp, C.POINTER(_u64), _vp]),
    "gsparc_rasterize_forward_render": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _u64, _vp, _vp, _vp, _i, _vp]),
    "gsparc_rasterize_backward": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f, _vp, _vp,
                                       _vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                       _vp]),
    "gsparc_rasterize_forward": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f, _vp, _vp,
                                      _vp, _u64, _vp, _vp, _vp, _i, _vp]),
    "gsparc_instance_count": (_i, [_i, _i, _i, _i, _vp, C.POINTER(_u64), _vp]),
    "gsparc_scene_table_words": (_sz, [_i, _i]),
    "gsparc_scene_table_build": (_i, [_i, C.POINTER(_i), C.POINTER(_i), C.POINTER(C.c_uint32), _sz, _BP]),
    "gsparc_scene_buffer_sizes": (_i, [_BP, _i, _i, _u64, C.POINTER(_sz)]),
    "gsparc_rasterize_forward_scenes": (_i, [_BP, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f,
                                             _vp, _vp, _vp, _u64, _vp, _vp, _vp, _i, _vp]),
    "gsparc_rasterize_forward_geom_scenes": (_i, [_BP, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f,
                                                  _vp, _vp, _vp, C.POINTER(_u64), _vp]),
    "gsparc_rasterize_forward_render_scenes": (_i, [_BP, _vp, _i, _i, _vp, _vp, _vp, _vp, _u64, _vp, _vp, _vp,
                                                    _i, _vp]),
    "gsparc_instance_count_scenes": (_i, [_BP, _vp, _i, _i, _vp, C.POINTER(_u64), _vp]),
    "gsparc_rasterize_backward_scenes": (_i, [_i, _BP, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f,
                                              _vp, _vp, _vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                              _vp, _vp, _vp, _vp]),
    "gsparc_mark_visible": (_i, [_i, _vp, _vp, _vp, _vp]),
    "gsparc_ssim_forward": (_i, [_i, _i, _i, _f, _f, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gsparc_ssim_backward": (_i, [_i, _i, _i, _f, _f, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gsparc_autotune": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, C.c_uint64, _vp, _vp, _vp, _vp, _vp, _i,
                             C.POINTER(_i), C.POINTER(_f), C.POINTER(_f), _vp]),
    "gsparc_autotune_scenes": (_i, [_i, _BP, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp, C.c_uint64, _vp, _vp, _vp, _vp,
                                    _vp, _i, C.POINTER(_i), C.POINTER(_f), C.POINTER(_f), _vp]),
    "gsparc_autotune_clear": (_i, []),
    "gsparc_rasterize_forward_render_csi": (_i, [_i, _i, _i, _i, C.c_double, C.c_double, _vp, _vp, _vp, _vp,
                                                 C.c_uint64, _vp, _vp, _vp]),
    "gsparc_rasterize_forward_csi": (_i, [_i, _i, _i, _i, C.c_double, C.c_double] + [_vp] * 9 + [_f, _vp, _vp, _vp,
                                          C.c_uint64, _vp, _vp, _vp]),
    "gsparc_rasterize_backward_csi": (_i, [_i, _i, _i, _i, C.c_double, C.c_double] + [_vp] * 8 + [_f, _vp, _vp, _vp,
                                           _vp, C.c_uint64, _vp, _vp] + [_vp] * 8 + [_vp]),
    "gsparc_mlp1_forward": (_i, [_i, _i] + [_vp] * 12),
    "gsparc_mlp1_work_size": (C.c_size_t, [_i, _i]),
    "gsparc_mlp1_backward": (_i, [_i, _i] + [_vp] * 14),
    "gsparc_mlp1_backward_accumulate": (_i, [_i, _i] + [_vp] * 14),
    "gsparc_csi_loss_work_size": (C.c_size_t, [_i, _i, _i]),
    "gsparc_csi_loss": (_i, [_i, _i, _i, _f, _f, _f, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gsparc_loss_work_size": (C.c_size_t, [_i, _i, _i]),
    "gsparc_loss_l1_ssim": (_i, [_i, _i, _i, _f, _f, _f, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gsparc_loss_l1_ssim_noalloc": (_i, [_i, _i, _i, _f, _f, _f, _vp, _vp, _vp, _vp, _vp]),
    "gsparc_conf_loss_work_size": (_sz, [_i, _i, _i]),
    "gsparc_loss_l1_ssim_conf": (_i, [_i, _i, _i, _f, _f, _f, _f, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _vp, _vp]),
    "gsparc_gaussians_activate": (_i, [_i, _GP, _vp]),
    "gsparc_adam_step": (_i, [_i, _i, C.POINTER(AdamConfig), _GP] + [_vp] * 5 + [_i, _vp, _vp, _vp, _vp, _vp]),
    "gsparc_densify_stats": (_i, [_i, _i, _f, _vp, _vp, _GP, _vp]),
    "gsparc_opacity_reset": (_i, [_i, _GP, _vp]),
    "gsparc_densify_work_size": (_sz, [_i]),
    "gsparc_densify_and_prune": (_i, [_i, C.POINTER(DensifyConfig), _GP, _vp, _GP, _i, _vp, C.POINTER(_i),
                                      C.POINTER(_i), C.POINTER(_i), _vp]),
    "gsparc_eval_work_size": (_sz, [_i, _i, _i]),
    "gsparc_eval_metrics": (_i, [_i, _i, _i, _f, _f, _f, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gsparc_view_matrices": (_i, [_i, C.POINTER(C.c_double), _vp, _vp, _vp]),
    "gsparc_gather_batch": (_i, [_i, _i, _i, _vp, _i] + [_vp] * 8),
    "gsparc_csi_phase_factors": (_i, [_sz, _vp, _vp, _vp, _vp]),
    "gsparc_csi_phasors": (_i, [_sz, _vp, C.c_double, C.c_double, _vp, _vp]),
    "gsparc_graph_begin": (_i, [_vp]),
    "gsparc_graph_end": (_i, [_vp, C.POINTER(_vp)]),
    "gsparc_graph_launch": (_i, [_vp, _vp]),
    "gsparc_graph_destroy": (_i, [_vp]),
    "gsparc_profile_enable": (_i, [_i]),
    "gsparc_profile_collect": (_i, [C.POINTER(_i)]),
    "gsparc_profile_query": (_i, [_i, C.c_char_p, _i, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "gsparc_profile_counts": (_i, [C.POINTER(_u64)]),
}

_lib = None


def lib():
    """The loaded library (raises ImportError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              "`python gsparc-wirelessgaussiansplatting_amd/csrc/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def check(status: int, what: str):
    if status != 0:
        msg = lib().gsparc_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (status {status}): {msg}")


def ptr(t: torch.Tensor | None):
    return None if t is None else C.c_void_p(t.data_ptr())


def stream_handle(device=None):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def set_option(name: str, value: int):
    check(lib().gsparc_set_option(name.encode(), int(value)), f"gsparc_set_option({name})")


def get_option(name: str) -> int:
    v = C.c_int(0)
    check(lib().gsparc_get_option(name.encode(), C.byref(v)), f"gsparc_get_option({name})")
    return v.value


def profile_enable(on: bool = True, timing: bool = True, counting: bool = True):
    """Per-kernel HIP-event timing and/or the blend kernels' counters (the
    counting variants are slower, so time and count in separate passes)."""
    mode = 0 if not on else (1 if timing and counting else (2 if timing else (3 if counting else 0)))
    check(lib().gsparc_profile_enable(mode), "gsparc_profile_enable")


def profile_collect() -> dict:
    """{kernel name: (total ms, launches)} since profile_enable(True)."""
    nk = C.c_int(0)
    check(lib().gsparc_profile_collect(C.byref(nk)), "gsparc_profile_collect")
    n = nk.value
    out = {}
    buf = C.create_string_buffer(64)
    for i in range(n):
        ms = C.c_double(0.0)
        cnt = C.c_longlong(0)
        check(lib().gsparc_profile_query(i, buf, 64, C.byref(ms), C.byref(cnt)), "gsparc_profile_query")
        out[buf.value.decode()] = (ms.value, cnt.value)
    return out


def profile_counts() -> dict:
    """Blend-kernel counters accumulated since profile_enable(True), in the
    reference's terms: pairs its forward visits and its backward walks, and
    the contributing pairs found by the forward and by the backward."""
    v = (C.c_uint64 * 4)()
    check(lib().gsparc_profile_counts(v), "gsparc_profile_counts")
    return {"fwd_pairs": int(v[0]), "bwd_pairs": int(v[1]), "fwd_contrib": int(v[2]), "bwd_contrib": int(v[3])}


def layout(P, R, W, H, K=0) -> Layout:
    L = Layout()
    check(lib().gsparc_layout_query(P, R, W, H, K, C.byref(L)), "gsparc_layout_query")
    return L


def require_cuda(*tensors):
    for t in tensors:
        if t is not None and t.numel() > 0 and not t.is_cuda:
            raise RuntimeError("gsparc kernels need tensors on a HIP device (got a CPU tensor)")


def f32(t: torch.Tensor) -> torch.Tensor:
    """Contiguous float32 view/copy, as the reference's .contiguous().data_ptr<float>()."""
    if t.dtype != torch.float32:
        t = t.float()
    return t.contiguous()


def f32_aligned(t: torch.Tensor) -> torch.Tensor:
    """As f32, and 16-byte aligned (a copy when a view starts mid-vector)."""
    t = f32(t)
    return t if t.data_ptr() % 16 == 0 else t.clone()
