"""ctypes binding of libmaomi_amd.so (the C ABI declared in include/maomi_amd.h).

The library is loaded lazily on first use; there is no fallback: every GPU
operation of the package raises if the HIP library is missing or a call fails.
"""

import ctypes
import os
from ctypes import c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_void_p
from pathlib import Path

_PATH = Path(os.environ.get("MAOMI_AMD_LIBRARY") or
             Path(__file__).resolve().with_name("libmaomi_amd.so"))
_LIB = None


class SitesDesc(ctypes.Structure):
    """Mirror of maomi_sites_desc."""
    _fields_ = [
        ("bases", c_void_p), ("base_off", c_void_p), ("phase", c_void_p), ("phase16", c_void_p),
        ("slice_scale", c_void_p), ("edge", c_void_p), ("q1d", c_void_p), ("twiddle", c_void_p),
        ("site_x", c_void_p), ("site_y", c_void_p), ("site_n", c_void_p), ("site_s", c_void_p),
        ("order", c_void_p), ("mask", c_void_p), ("mask_off", c_void_p), ("energy", c_void_p),
        ("workspace", c_void_p), ("scratch_stride", c_int64),
        ("P", c_int64), ("H", c_int64), ("W", c_int64), ("Z", c_int64),
        ("grid", c_int64), ("uniform_n", c_int64), ("block_threads", c_int64),
        ("class_edge", c_int64 * 5), ("class_q", c_int64 * 5), ("class_tw", c_int64 * 5),
        ("psf_out", c_void_p), ("psf_z0", c_int64), ("psf_nz", c_int64), ("psf_c0", c_int64),
        ("psf_w", c_int64), ("screen", c_void_p), ("queue", c_void_p), ("audit", c_void_p),
    ]


class SitesAudit(ctypes.Structure):
    """Mirror of maomi_sites_audit."""
    _fields_ = [("pred", c_void_p), ("slot", c_void_p), ("state", c_void_p), ("rel", c_void_p),
                ("done", c_void_p), ("n_audit", c_int64), ("max_fail", c_int64),
                ("tol", c_double)]


class SitesChoice(ctypes.Structure):
    """Mirror of maomi_sites_choice."""
    _fields_ = [("grid", c_int64), ("block_threads", c_int64), ("order", c_int64),
                ("schedule", c_int64), ("ms", c_double)]


ORDERS = {"natural": 0, "lpt": 1, "class_interleaved": 2}
SCHEDULES = {"static": 0, "queue": 1}


class PupilParams(ctypes.Structure):
    """Mirror of maomi_pupil_params."""
    _fields_ = [("kind", ctypes.c_int32), ("profile", ctypes.c_int32), ("clip", ctypes.c_int32),
                ("cylinder", ctypes.c_int32), ("rr", ctypes.c_float), ("obj2", ctypes.c_float),
                ("focus", ctypes.c_float), ("radius", ctypes.c_float), ("width", ctypes.c_float),
                ("offset", ctypes.c_float), ("shift", ctypes.c_float), ("tilt", ctypes.c_float),
                ("scale_x", c_double), ("scale_y", c_double)]


class Segment(ctypes.Structure):
    """Mirror of maomi_segment."""
    _fields_ = [("phase", c_void_p), ("H", c_int64), ("W", c_int64), ("Z", c_int64),
                ("n", c_int64), ("steps", c_int64), ("n1", c_int64), ("n2f", c_int64),
                ("q", c_void_p), ("qidx", c_void_p), ("edge", c_void_p), ("tw", c_void_p)]


class SegmentedDesc(ctypes.Structure):
    """Mirror of maomi_segmented_desc."""
    _fields_ = [
        ("seg", Segment * 3), ("base", c_void_p), ("site_x", c_void_p), ("site_y", c_void_p),
        ("order_list", c_void_p),
        ("P", c_int64), ("cut", c_int64), ("tails", c_int64), ("order", c_int64),
        ("n", c_int64), ("n2", c_int64), ("fx", c_int64), ("fy", c_int64),
        ("X", c_int64), ("Y", c_int64), ("TT", c_int64),
        ("mfx", c_void_p), ("mfy", c_void_p), ("mt", c_void_p),
        ("ax_focus", c_void_p), ("ax_top", c_void_p), ("ax_bot", c_void_p),
        ("ss2", c_double), ("focal_scale", c_double), ("tail_scale", c_double),
        ("prop_um", c_double),
        ("psf", c_void_p), ("top_conv", c_void_p), ("top_prof", c_void_p), ("top_w", c_void_p),
        ("bot_conv", c_void_p), ("bot_prof", c_void_p), ("bot_w", c_void_p),
        ("work", c_void_p), ("work_stride", c_int64), ("grid", c_int64),
        ("mfx_band", c_void_p), ("mfy_band", c_void_p), ("mt_band", c_void_p),
    ]


_SIGS = {
    "maomi_last_error": (c_char_p, []),
    "maomi_version": (c_char_p, []),
    "maomi_persistent_sites": (c_int, [ctypes.POINTER(SitesDesc), c_void_p]),
    "maomi_sites_workspace_bytes": (c_int64, [c_int64, c_int64]),
    "maomi_sites_default_grid": (c_int64, []),
    "maomi_sites_plan_create": (c_int, [ctypes.POINTER(SitesDesc), ctypes.POINTER(c_void_p)]),
    "maomi_sites_plan_capture": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_int64]),
    "maomi_sites_plan_tune": (c_int, [c_void_p, ctypes.POINTER(c_int64), c_int,
                                      ctypes.POINTER(c_int64), c_int, ctypes.POINTER(c_int64),
                                      c_int, ctypes.POINTER(c_int64), c_int, c_int, c_void_p,
                                      ctypes.POINTER(SitesChoice),
                                      ctypes.POINTER(SitesChoice)]),
    "maomi_sites_plan_replay": (c_int, [c_void_p, c_void_p]),
    "maomi_sites_plan_choice": (c_int, [c_void_p, ctypes.POINTER(SitesChoice)]),
    "maomi_sites_plan_destroy": (c_int, [c_void_p]),
    "maomi_phase_prepare": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_int64, c_double,
                                    c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "maomi_twiddles": (c_int, [c_void_p, c_int64, c_void_p]),
    "maomi_pupil_gaussian": (c_int, [c_void_p, c_int64, c_double, c_double, c_double, c_double,
                                     c_double, c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "maomi_pupil_field": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                                  ctypes.POINTER(PupilParams), c_void_p]),
    "maomi_fresnel_step": (c_int, [c_void_p, c_int64, c_int64, c_double, c_double, c_double,
                                   c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p]),
    "maomi_normalize_l2": (c_int, [c_void_p, c_int64, c_void_p, c_void_p]),
    "maomi_edge_window": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "maomi_fresnel_q1d": (c_int, [c_void_p, c_int64, c_double, c_double, c_double, c_void_p]),
    "maomi_fft2_c64": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p]),
    "maomi_resize_phase": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p,
                                   c_int64, c_int64, c_void_p, c_void_p, c_int64, c_int64,
                                   c_double, c_void_p, c_void_p, c_void_p]),
    "maomi_segmented_sites": (c_int, [ctypes.POINTER(SegmentedDesc), c_void_p]),
    "maomi_segmented_workspace_elems": (c_int64, [c_int64, c_int64, c_int64, c_int64, c_int64,
                                                  c_int64]),
    "maomi_gemm_bf16": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64,
                                c_int64, c_int64, c_int64, c_int64, c_int64, c_float, c_float,
                                c_void_p]),
    "maomi_im2col_bf16": (c_int, [c_void_p, c_void_p] + [c_int64] * 7 + [c_void_p]),
    "maomi_col2im": (c_int, [c_void_p, c_void_p] + [c_int64] * 7 + [c_void_p]),
    "maomi_bias_act": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_float, c_void_p]),
    "maomi_act_bwd": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_float, c_void_p, c_void_p]),
    "maomi_plane_bias": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64,
                                 c_void_p]),
    "maomi_plane_bias_grad": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_int64, c_float,
                                      c_void_p]),
    "maomi_to_bf16": (c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    "maomi_upsample2": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64, c_void_p]),
    "maomi_rel_sq_loss": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64,
                                  c_int64, c_void_p]),
    "maomi_energy_sq_loss": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
                                     c_int64, c_int64, c_void_p]),
    "maomi_adam": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_float,
                           c_float, c_float, c_float, c_int64, c_float, c_void_p]),
    "maomi_patch_features": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p,
                                     c_void_p, c_int64, c_int64, c_int64, c_int64, c_double,
                                     c_double, c_void_p, c_void_p, c_void_p]),
    "maomi_psf_stats": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p]),
    "maomi_lfft_gather": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p,
                                  c_void_p, c_int64, c_int64, c_int64, c_int64, c_void_p]),
    "maomi_lfft_product": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_void_p]),
    "maomi_lfft_gather_z": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64,
                                    c_void_p]),
    "maomi_lfft_scatter_z": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_int64, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64,
                                     c_int64, c_int64, c_void_p]),
    "maomi_lfft_product_fields": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_int64,
                                          c_void_p, c_int64, c_int64, c_void_p]),
    "maomi_rana_node_weights": (c_int, [c_void_p, c_int64, c_double, c_double, c_int64,
                                        c_void_p, c_void_p]),
    "maomi_rana_depth_weights": (c_int, [c_void_p, c_int64, c_void_p, c_int64, c_void_p,
                                         c_void_p]),
    "maomi_lfft_scatter": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_int64, c_void_p,
                                   c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64,
                                   c_int64, c_void_p]),
    "maomi_grf_spectrum": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p,
                                   c_void_p, c_int32, c_void_p, c_void_p]),
    "maomi_grf_finish": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_float, c_float, c_float,
                                 c_float, c_void_p, c_float, c_void_p, c_void_p]),
    "maomi_zsum": (c_int, [c_void_p, c_int32, c_int64, c_int64, c_int64, c_int64, c_int64,
                           c_void_p, c_int32, c_void_p, c_int64, c_int64, c_void_p]),
    "maomi_collection_mask_workspace": (c_int64, [c_int64, c_int64, c_int64, c_int64, c_int32]),
    "maomi_collection_mask": (c_int, [c_void_p, c_int32, c_int64, c_int64, c_int64, c_int64,
                                      c_int64, c_void_p, c_void_p, c_void_p, c_int32, c_int64,
                                      c_int64, c_int64, c_int64, c_float, c_float, c_void_p,
                                      c_void_p, c_void_p]),
    "maomi_screen_exp": (c_int, [c_void_p, c_int64, c_void_p, c_void_p]),
    "maomi_fresnel_steps": (c_int, [c_void_p, c_int64, c_int64, c_double, c_double, c_void_p,
                                    c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p]),
    "maomi_v4_green_matrix_f64": (c_int, [c_void_p, c_int64, c_void_p, c_void_p]),
    "maomi_v4_apply_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64,
                                   c_int64, c_void_p]),
    "maomi_v4_apply_split_f64": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
                                         c_int64, c_int64, c_void_p, c_int64, c_void_p]),
    "maomi_v4_apply_split_f64_noalloc": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p,
                                                 c_int64, c_int64, c_int64, c_int64, c_void_p]),
    "maomi_v4_apply_rows": (c_int64, [c_int64, c_int64]),
    "maomi_v4_apply_grid_f64": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64,
                                        c_int64, c_void_p, c_int64, c_void_p]),
    "maomi_cholesky_f64": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p]),
    "maomi_cholesky_f64_workspace": (c_int64, []),
    "maomi_v4_finish_f64": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                                    c_int64, c_void_p, c_void_p, c_void_p]),
    "maomi_v4_finish_f64_workspace": (c_int64, [c_int64, c_int64]),
    "maomi_conv_bf16_fwd": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_int64, c_int64,
                                    c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64,
                                    c_float, c_void_p, c_int64, c_void_p]),
    "maomi_to_channel_last_bf16": (c_int, [c_void_p, c_int64, c_int64, c_int64, c_int64,
                                           c_void_p, c_void_p]),
    "maomi_resample_linear_f64": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_int64,
                                          c_void_p, c_int64, c_void_p, c_void_p]),
    "maomi_pupil_embed_f64": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p,
                                      c_void_p]),
    "maomi_power_shift_f64": (c_int, [c_void_p, c_int64, c_void_p, c_void_p]),
}


def library_path():
    return _PATH


def lib():
    """Load libmaomi_amd.so and declare every C-ABI signature."""
    global _LIB
    if _LIB is None:
        if not _PATH.exists():
            raise RuntimeError(f"HIP library not built: {_PATH} is missing "
                               "(run `python -c 'import __graft_entry__ as g; g.build()'`)")
        # torch ships its own HIP runtime (libamdhip64.so.7).  Loaded first, it also serves the
        # library's libamdhip64.so.7 dependency, so kernels launch in the runtime that owns
        # torch's memory and streams.  Loaded after the library, it would be a second runtime
        # in the process, and the library's launches would fail ("no ROCm-capable device").
        import torch  # noqa: F401
        handle = ctypes.CDLL(str(_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def exported_names():
    return list(_SIGS)


def check(rc, what):
    if rc != 0:
        msg = lib().maomi_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed with code {rc}: {msg}")


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle(device=None):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
