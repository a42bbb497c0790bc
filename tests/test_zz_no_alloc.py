"""zc_moments_noalloc (the allocation-free variant of zc_moments) against zc_moments through the
C ABI: the same 45 sums, bit for bit (the same kernels and reduction order), on two calls in a
row, the second with a smaller pair (the fixed scratch array keeps the first call's rows, which
the second must not read)."""
import ctypes

import numpy as np
import pytest

import zcmir_amd
from zcmir_amd import synth

pytestmark = pytest.mark.gpu


def _noalloc(lib, torch, fa, fb, H):
    da = torch.tensor(fa, device="cuda")
    db = torch.tensor(fb, device="cuda")
    out = torch.full((45,), float("nan"), dtype=torch.float64, device="cuda")
    h9 = np.ascontiguousarray(np.asarray(H, np.float64).reshape(9))
    torch.cuda.synchronize()
    rc = lib.zc_moments_noalloc(ctypes.c_void_p(da.data_ptr()), fa.shape[2], fa.shape[1],
                                ctypes.c_void_p(db.data_ptr()), fb.shape[2], fb.shape[1],
                                h9.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.c_void_p(out.data_ptr()),
                                None)
    torch.cuda.synchronize()
    assert rc == 0
    return out.cpu().numpy()


def test_moments_noalloc_matches_engine_twice(reg_factory):
    torch = pytest.importorskip("torch")
    lib = zcmir_amd._load()
    cases = [((300, 340), (7, -4), np.array([[1.02, 0.03, -5.0], [-0.02, 0.99, 3.0], [1e-5, 0.0, 1.0]])),
             ((96, 80), (3, 2), np.array([[1.0, 0.0, 2.5], [0.0, 1.0, -1.5], [0.0, 0.0, 1.0]]))]
    for size, shift, H in cases:
        mv, fx, _ = synth.shift_pair(size, shift, seed=size[0])
        reg = reg_factory(half=False)
        reg.set_images(mv, fx, preprocess=False)
        want = reg.moments(H)
        got = _noalloc(lib, torch, reg.features(0), reg.features(1), H)
        assert want[0] > 0 and np.array_equal(got, want), (size, np.abs(got - want).max())
    # an invalid argument is refused without touching the device
    assert lib.zc_moments_noalloc(None, 1, 1, None, 1, 1, None, None, None) == -1
