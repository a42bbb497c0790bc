"""The C ABI library (CPU-side checks, no GPU compute): it loads, exports every function the
header declares, binds the reference's full ctypes symbol list, and its structs have the
reference's layouts."""
import ctypes
import re
from pathlib import Path

import pytest

from zcmir_amd import _native

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "zcsmi.h"

# python/zcmir/_native.py:79-133 — every name the reference binding resolves at load time
REFERENCE_SIGS = """zc_version zc_create zc_create_error zc_destroy zc_adapter zc_error zc_configure zc_settings
zc_set_image zc_set_image_gray zc_set_image_rgba zc_work_image zc_features zc_compile_wgsl zc_moments zc_pg_info
zc_pg_maps zc_pg_descriptors zc_image_size zc_cancel zc_profile zc_bench_moments zc_tune_parts zc_gpu_counts
zc_profile_report zc_swap zc_match_scores zc_detect zc_match zc_search zc_auto zc_climb zc_keypoints
zc_keypoint_frames zc_matches zc_set_ffd zc_get_ffd zc_set_pose zc_fit_check zc_get_pose zc_events zc_score
zc_gradient zc_ffd_gradient zc_shift_map zc_shift_map_set zc_rs_map zc_canvas zc_overlay zc_tile_heat
zc_export_warp zc_displacement""".split()


def _declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(zc_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    p = _native.lib_path()
    if not p.exists():
        pytest.fail(f"{p} is not built (python -m zcmir_amd.build)")
    return ctypes.CDLL(str(p))


def test_header_declares_the_reference_surface():
    d = set(_declared())
    missing = [n for n in REFERENCE_SIGS if n not in d]
    assert not missing, missing


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_loads(lib):
    l2 = _native.load()
    for name in _native._SIGS:
        assert hasattr(l2, name)
    assert l2.zc_version().startswith(b"0.1.4")


def test_struct_sizes(lib):
    out = (ctypes.c_uint32 * 4)()
    l2 = _native.load()
    l2.zc_result_sizes(out)
    assert out[0] == ctypes.sizeof(_native.ShiftResult) == 88
    assert out[1] == ctypes.sizeof(_native.RsResult) == 96
    assert out[2] == ctypes.sizeof(_native.GradResult) == 616


def test_create_without_a_device_fails_with_a_reason():
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    l2 = _native.load()
    e = l2.zc_create(b"0")
    assert not e
    buf = ctypes.create_string_buffer(512)
    n = l2.zc_create_error(buf, 512)
    assert n > 0
