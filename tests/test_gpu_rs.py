"""GPU parity of the roto-scale (log-polar) map against the float64 oracle, through the C ABI.

Bars: the forward and inverse-frame maps within max |Δ| / max |map| <= 1e-4 on lags whose
overlap is clear of the floor and whose unscaled overlap weight is at least 10 % of the largest;
within 1e-2 on the other lags clear of the floor; the finished map likewise; the same peak
(dth, dsg); the same grid (n, n_lam) and radii. Grids: powers of two (the FFT length itself) and
2^a 3^b lengths (angle axis padded and folded), both feature modes, exact and global scores,
calibrated.

Why the two bands: every log-polar cell carries the area weight rho (per side), so the planes span
(r1 / r0)^2 (up to 400x) in magnitude and a float32 FFT's absolute error follows the largest of
them. A lag pairing one image's inner rings with the other's outer rings has a small raw overlap
sum, which the extract then scales up by e^(k dlam): its moments carry float32 error that the
per-lag whitening amplifies. The reference computes these maps in float32 as well. Replacing the
oracle's float64 correlations by float32 FFTs of the same planes gives, at 192^2 (1024 angles),
max relative errors of 2.9e-3 over all clear lags and 1.6e-5 over lags with raw overlap >= 10 %.
"""
import math

import numpy as np
import pytest

from oracle import smi_ref as O
from zcmir_amd import synth

pytestmark = pytest.mark.gpu

MAP_TOL = 1e-4


def sim_about(c, th, sg):
    s = math.exp(sg)
    a, b, d, e = s * math.cos(th), -s * math.sin(th), s * math.sin(th), s * math.cos(th)
    return np.array([[a, b, c[0] - a * c[0] - b * c[1]], [d, e, c[1] - d * c[0] - e * c[1]], [0, 0, 1.0]])


def _clear(plan, calibrated, band=1e-3):
    """Lags whose overlap count is clear of the score's floor thresholds."""
    n0 = plan["corr"][0]
    if calibrated:
        ths = [float(O.MIN_N), 10.0 * max(plan["area"], 1.0)]
    else:
        ths = [max(float(O.MIN_N), 0.08 * max(float(n0[0, 0]), 0.0))]
    ok = np.ones(n0.shape, bool)
    for t in ths:
        ok &= np.abs(n0 - t) > band * max(t, 1.0) + 1e-3 * max(t, 1.0)
    return ok


def _raw_ok(plan, frac=0.10):
    """Lags whose unscaled overlap weight (before the e^(k dlam) dlam dth factor) is at least frac
    of the largest."""
    n = plan["n"]
    k = np.arange(n)
    k = np.where(k > n // 2, k - n, k)
    s = np.exp(k * plan["dlam"]) * plan["dlam"] * (2 * math.pi / plan["n_th"])
    raw = plan["corr"][0] / s[:, None]
    return raw >= frac * raw.max()


def _err(g, o, clear):
    return float(np.abs(g - o)[clear].max() / max(np.abs(o).max(), 1e-30)) if clear.any() else 0.0


def _check(g, o, clear, well):
    strict, loose = _err(g, o, clear & well), _err(g, o, clear)
    assert strict <= MAP_TOL and loose <= 1e-2, (strict, loose)


CASES = [
    # size, degrees, scale, mode, exact, calibrated, symmetric, map_res
    ((100, 100), 10.0, 0.93, "laplacian", True, False, True, 1024),    # 486 angles: padded, folded
    ((96, 120), -8.0, 1.06, "three", True, False, False, 1024),        # 576 angles, three features
    ((128, 128), 20.0, 1.10, "laplacian", True, False, True, 1024),    # 864 angles
    ((192, 192), 6.0, 1.08, "laplacian", True, False, True, 1024),     # 1024 angles: power of two
    ((100, 100), 10.0, 0.93, "laplacian", False, False, True, 1024),   # global score
    ((100, 100), 10.0, 0.93, "laplacian", True, True, True, 1024),     # calibrated
    ((100, 100), -25.0, 1.15, "laplacian", True, False, True, 0),      # automatic grid (512)
]


@pytest.mark.parametrize("size,deg,scale,mode,exact,calibrated,symmetric,res", CASES)
def test_rs_map_matches_oracle(reg_factory, size, deg, scale, mode, exact, calibrated, symmetric, res):
    h, w = size
    print(f"rs_map case {size} {deg} {scale} {mode} exact={exact} calib={calibrated} sym={symmetric} res={res}",
          flush=True)
    mv, fx, _ = synth.shift_pair((h, w), (0, 0), seed=4)
    c0 = ((w + 1) / 2, (h + 1) / 2)
    H1 = sim_about(c0, math.radians(deg), math.log(scale))
    reg = reg_factory(half=False, smi_features=mode, exact=exact, calibrated=calibrated, map_res=res)
    reg.set_images(mv, fx, preprocess=False)
    rm = reg.rs_map(H1, symmetric=symmetric)
    pa, pb = O.prepare(mv, mode), O.prepare(fx, mode)
    om, info = O.rs_map(pa, pb, H1, symmetric=symmetric, exact=exact, calibrated=calibrated, cap=res)
    assert (rm.n, rm.n_th, rm.n_lam) == (info["n"], info["n_th"], info["n_lam"])
    assert abs(rm.r0 - info["r0"]) < 1e-9 and abs(rm.r1 - info["r1"]) < 1e-9 and abs(rm.dlam - info["dlam"]) < 1e-12
    assert np.allclose(rm.center, info["center"], atol=1e-9)
    calib = exact and calibrated
    cf, wf = _clear(info, calib), _raw_ok(info)
    _check(rm.fwd, info["fwd"], cf, wf)
    clear, well = cf, wf
    if symmetric:
        ci, wi = _clear(info["inv_plan"], calib), _raw_ok(info["inv_plan"])
        _check(rm.inv, info["inv"], ci, wi)
        n = info["n"]
        mir = (n - np.arange(n)) % n
        xm = mir if H1[0, 0] * H1[1, 1] - H1[0, 1] * H1[1, 0] > 0 else np.arange(n)
        clear = cf & ci[mir][:, xm]
        well = wf & wi[mir][:, xm]
    _check(rm.map, om, clear, well)
    assert (rm.dth, rm.dsg) == (info["dth"], info["dsg"]), ((rm.dth, rm.dsg), (info["dth"], info["dsg"]))
    if exact and not calibrated:
        assert abs(rm.dth + math.radians(deg)) <= 2 * math.pi / rm.n_th and abs(rm.dsg + math.log(scale)) <= rm.dlam
        Hc = rm.corrected(H1)
        pts = np.array([[1, 1, 1], [w, 1, 1], [1, h, 1], [w, h, 1]], float).T
        q = Hc @ pts
        assert np.abs(q[:2] / q[2] - pts[:2]).max() < 1.5


def test_rs_map_errors_fail_loudly(reg_factory):
    reg = reg_factory(half=False)
    with pytest.raises(RuntimeError):
        reg.rs_map(np.eye(3))                 # no images
