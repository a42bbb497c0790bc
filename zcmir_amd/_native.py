"""ctypes bindings to libzcsmi.so (the HIP kernels and C ABI built from zcmir_amd/csrc).

Signatures and result-struct layouts follow the reference binding (python/zcmir/_native.py:14-133)
so code written against it binds unchanged; the extra names at the end are this library's own
inspection and benchmark hooks. Loading fails loudly when the library is missing: there is no
CPU fallback.
"""
import ctypes
import os
from pathlib import Path

HERE = Path(__file__).resolve().parent

U32, I32, I64, F64, SZ, P = ctypes.c_uint32, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p
F32 = ctypes.c_float
f32p, f64p, u8p, u32p = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_uint32)
Mat = ctypes.c_double * 9


class ShiftResult(ctypes.Structure):
    _fields_ = [
        ("n", U32), ("cw", U32), ("ch", U32), ("canvas_w", U32), ("canvas_h", U32),
        ("ox", I32), ("oy", I32), ("peak_index", U32),
        ("dx", F64), ("dy", F64), ("peak", F64), ("zero", F64), ("corr_area_px", F64),
        ("exact", U32), ("calibrated", U32), ("n_y", U32), ("_pad", U32),
    ]


class RsResult(ctypes.Structure):
    _fields_ = [
        ("n", U32), ("n_th", U32), ("n_lam", U32), ("peak_index", U32),
        ("dlam", F64), ("r0", F64), ("r1", F64), ("cx", F64), ("cy", F64), ("peak", F64), ("zero", F64),
        ("dth", F64), ("dsg", F64), ("symmetric", U32), ("calibrated", U32),
    ]


class GradResult(ctypes.Structure):
    _fields_ = [
        ("score", F64), ("fwd", F64), ("inv", F64), ("n", F64),
        ("grad", F64 * 8), ("hess", F64 * 64), ("has_hess", U32), ("nk", U32),
    ]


class ClimbResult(ctypes.Structure):
    _fields_ = [("H", Mat), ("score", F64), ("iterations", U32), ("moved", U32)]


class DetectResult(ctypes.Structure):
    _fields_ = [("n1", U32), ("n2", U32), ("levels1", U32), ("levels2", U32)]


class MatchResult(ctypes.Structure):
    _fields_ = [
        ("H", Mat), ("ninl", U32), ("n_corr", U32), ("ninl_aff", U32), ("pos_n", U32), ("pos_inl", U32),
        ("has_rot", U32), ("rot_deg", F64), ("pos_kept", U32), ("_pad", U32),
    ]


class SearchResult(ctypes.Structure):
    _fields_ = [("H", Mat), ("score", F64), ("prev", F64), ("improved", U32), ("count_a", U32), ("count_b", U32), ("_pad", U32)]


class SweepInfo(ctypes.Structure):
    _fields_ = [("nc", U32), ("n", U32), ("grid", U32), ("n_peaks", U32), ("cell", F64), ("D", F64), ("ox", F64),
                ("oy", F64)]


class SweepPeak(ctypes.Structure):
    _fields_ = [("score", F64), ("cand", U32), ("it", U32), ("is_", U32), ("ik", U32), ("ia", U32), ("_pad", U32),
                ("th", F64), ("s", F64), ("sx", I32), ("sy", I32), ("H", Mat)]


class Canvas(ctypes.Structure):
    _fields_ = [("ox", I32), ("oy", I32), ("ow", U32), ("oh", U32), ("Hc", Mat)]


_SIGS = {
    # the reference's names (python/zcmir/_native.py:79-133)
    "zc_version": ([], ctypes.c_char_p),
    "zc_create": ([ctypes.c_char_p], P),
    "zc_create_error": ([ctypes.c_char_p, SZ], SZ),
    "zc_destroy": ([P], None),
    "zc_adapter": ([P, ctypes.c_char_p, SZ], SZ),
    "zc_error": ([P, ctypes.c_char_p, SZ], SZ),
    "zc_configure": ([P, ctypes.c_char_p, SZ], I32),
    "zc_settings": ([P, ctypes.c_char_p, SZ], SZ),
    "zc_set_image": ([P, U32, f32p, U32, U32], I32),
    "zc_set_image_gray": ([P, U32, f32p, U32, U32, U32], I32),
    "zc_set_image_rgba": ([P, U32, u8p, U32, U32], I32),
    "zc_work_image": ([P, U32, f32p, SZ], I32),
    "zc_features": ([P, U32, f32p, SZ], I32),
    "zc_compile_wgsl": ([P, ctypes.c_char_p, SZ, ctypes.c_char_p, SZ, ctypes.c_char_p, SZ], I32),
    "zc_moments": ([P, f64p, f64p], I32),
    "zc_pg_info": ([P, f64p], I32),
    "zc_pg_maps": ([P, U32, f32p, SZ], I32),
    "zc_pg_descriptors": ([P, U32, f32p, SZ], I32),
    "zc_image_size": ([P, U32, u32p], None),
    "zc_cancel": ([P], None),
    "zc_profile": ([P, U32], I32),
    "zc_bench_moments": ([P, U32, U32], I32),
    "zc_tune_parts": ([U32, U32, U32, U32], None),
    "zc_gpu_counts": ([P, u32p], None),
    "zc_profile_report": ([P, ctypes.c_char_p, SZ], SZ),
    "zc_swap": ([P], None),
    "zc_match_scores": ([P, f32p, SZ], I32),
    "zc_detect": ([P, ctypes.POINTER(DetectResult)], I32),
    "zc_match": ([P, ctypes.POINTER(MatchResult)], I32),
    "zc_search": ([P, ctypes.POINTER(SearchResult)], I32),
    "zc_auto": ([P, ctypes.POINTER(ClimbResult)], I32),
    "zc_climb": ([P, ctypes.POINTER(ClimbResult)], I32),
    "zc_keypoints": ([P, U32, f32p, SZ], I32),
    "zc_keypoint_frames": ([P, U32, f32p, SZ], I32),
    "zc_matches": ([P, u32p, SZ], I32),
    "zc_set_ffd": ([P, f32p, SZ], I32),
    "zc_get_ffd": ([P, f32p, SZ], SZ),
    "zc_set_pose": ([P, f64p], None),
    "zc_fit_check": ([P, f64p], U32),
    "zc_get_pose": ([P, f64p], None),
    "zc_events": ([P, ctypes.c_char_p, SZ], SZ),
    "zc_score": ([P, f64p, f64p], I32),
    "zc_gradient": ([P, f64p, U32, ctypes.POINTER(GradResult)], I32),
    "zc_ffd_gradient": ([P, f64p, f64p, f64p, SZ], I32),
    "zc_shift_map": ([P, f64p, U32, U32, ctypes.POINTER(ShiftResult), f32p, SZ], I32),
    "zc_shift_map_set": ([P, f64p, ctypes.POINTER(ShiftResult), f32p, SZ], I32),
    "zc_fft_tiling": ([P, U32, u32p], I32),
    "zc_rs_map": ([P, f64p, f64p, U32, ctypes.POINTER(RsResult), f32p, f32p, f32p, SZ], I32),
    "zc_canvas": ([P, f64p, U32, ctypes.POINTER(Canvas)], I32),
    "zc_overlay": ([P, f64p, U32, ctypes.POINTER(Canvas), f32p, f32p, SZ], I32),
    "zc_tile_heat": ([P, f64p, U32, U32, f64p, f32p, f32p, SZ], I32),
    "zc_export_warp": ([P, ctypes.c_char_p, SZ, P, SZ], I64),
    "zc_displacement": ([P, f32p, SZ], I32),
    # exports.zig names the reference binding leaves out
    "zc_fft2": ([P, f32p, U32, U32, U32, U32], I32),
    "zc_fft_tune": ([U32, U32], None),
    "zc_result_sizes": ([u32p], None),
    # this library's own hooks (no reference counterpart)
    "zc_rank": ([P, U32, f32p, SZ, u32p, u32p, f32p], I32),
    "zc_bench_shift_map": ([P, f64p, U32, U32, U32, f64p], I32),
    "zc_sweep": ([P, f64p, P, P, SZ, f32p, SZ], I32),
    "zc_sweep_cells": ([P, f64p, u32p, U32, P, P, SZ, f32p, SZ], I32),
    "zc_sweep_map": ([P, f32p, SZ, u32p], I32),
    "zc_sweep_tune": ([U32], None),
    "zc_map_path": ([U32], None),
    "zc_moments_noalloc": ([P, U32, U32, P, U32, U32, f64p, P, P], I32),
    "zc_stack_set": ([P, U32, P, u32p, P, u32p], I32),
    "zc_stack_register": ([P, f64p, U32, U32, f64p, f64p, f64p], I32),
    "zc_set_image_device": ([P, U32, P, U32, U32], I32),
    "zc_search_cells": ([P, u32p, U32, P, f64p], I32),
    # the learned prior (device pointers as integers, a HIP stream handle)
    "zc_cnn_conv_fwd": ([P, P, P, P, I32, I32, I32, I32, I32, I32, I32, P], I32),
    "zc_cnn_conv_bwd": ([P, P, P, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, P], I32),
    "zc_cnn_fc_fwd": ([P, P, P, P, I32, I32, I32, P], I32),
    "zc_cnn_fc_bwd": ([P, P, P, P, P, P, I32, I32, I32, P], I32),
    "zc_cnn_xent": ([P, P, P, P, I32, I32, P], I32),
    "zc_cnn_adam": ([P, P, P, P, I64, F32, F32, F32, F32, I32, P], I32),
    "zc_lp_features": ([P, I32, I32, I32, I32, I32, I32, I32, I32, F32, F32, P, P, P], I32),
    "zc_warp_gray": ([P, I32, I32, I32, P, P, P, P], I32),
}

LIB_NAME = "libzcsmi.so"


def lib_path() -> Path:
    if os.environ.get("ZCSMI_LIB"):
        return Path(os.environ["ZCSMI_LIB"])
    return HERE / LIB_NAME


def load():
    """Load libzcsmi.so and bind every signature. Raises if it is missing (build it first:
    ``python -m zcmir_amd.build``)."""
    p = lib_path()
    if not p.exists():
        raise FileNotFoundError(f"{p} not found: build the HIP library with `python -m zcmir_amd.build`")
    # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7 (the same soname as the
    # system's), and whichever loads first serves both. PyTorch fails to initialize on a runtime
    # other than its own, so when it is installed it loads first (the learned prior uses it).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(str(p))
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes, fn.restype = args, res
    return lib
