"""Pins of the oracle's roto-scale (log-polar) map (oracle/smi_ref.py rs_*), CPU only.

* circular-in-angle, linear-in-log-radius correlations: the FFT path against brute-force sums;
* log-polar sampling: scipy.ndimage.map_coordinates (order 1) wherever all four taps are inside;
* rs_finish: the scale mask and the mirrored average on hand-computed rows and cells;
* the whole map: pairs built with a known rotation and scale about the centre give the
  correction (dth, dsg) = (-rotation, -log scale) within one grid cell, on power-of-two and
  3-smooth grids, with both feature modes, forward-only and symmetric (the reading of
  examples/03_smi.py:215-275: the start pose is the truth turned and scaled about the centre,
  and the map's peak is the offset back to the truth; corrected = simAbout(c, dth, dsg) · H).
"""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import smi_ref as O
from zcmir_amd import synth


def sim_about(c, th, sg):
    s = math.exp(sg)
    a, b, d, e = s * math.cos(th), -s * math.sin(th), s * math.sin(th), s * math.cos(th)
    return np.array([[a, b, c[0] - a * c[0] - b * c[1]], [d, e, c[1] - d * c[0] - e * c[1]], [0, 0, 1.0]])


def test_rs_correlations_circular_in_angle_match_direct_sums():
    rng = np.random.default_rng(0)
    n, h = 16, 8
    ra = rng.standard_normal((3, h, n))
    rb = rng.standard_normal((3, h, n))
    pairs = [(0, 0), (1, 2), (2, 1)]
    f = O.correlations_fft(ra, rb, n, n, pairs)
    d = O.correlations_direct_circ_x(ra, rb, n, n, pairs)
    assert np.abs(f - d).max() < 1e-12 * np.abs(d).max()


def test_logpolar_sampling_matches_map_coordinates():
    rng = np.random.default_rng(1)
    canvas = rng.standard_normal((5, 60, 70))
    p = O.rs_plan(70, 60, 0, 0, (36.0, 31.0), cap=256)
    lp = O.spatial_logpolar(canvas, p)
    y, x = np.mgrid[0:p["n_lam"], 0:p["n_th"]].astype(np.float64)
    rho = np.exp(p["log0"] + (y + 0.5) / p["n_lam"] * p["span"])
    th = (x + 0.5) / p["n_th"] * 2 * math.pi
    xs, ys = p["cx"] + rho * np.cos(th), p["cy"] + rho * np.sin(th)
    inside = (xs >= 0) & (ys >= 0) & (np.floor(xs) + 1 < 70) & (np.floor(ys) + 1 < 60)
    assert inside.mean() > 0.3 and (~inside).any()
    for k in range(5):
        ref = ndimage.map_coordinates(canvas[k], [ys, xs], order=1, mode="constant", cval=0.0)
        assert np.abs(lp[k] - ref)[inside].max() < 1e-12
        assert np.all(lp[k][~inside] == 0)


def test_rs_plan_radii_and_grid():
    # canvas 101 x 101 about its centre: r1 = hypot(50, 50), r0 = r1 / 16, 2 pi r1 = 444.3 -> 445
    # angles -> the next 2^a 3^b length 486 (cap 1024); auto: the next power of two 512
    p = O.rs_plan(101, 101, 0, 0, (51.0, 51.0), cap=1024)
    assert abs(p["r1"] - math.hypot(50, 50)) < 1e-12 and abs(p["r0"] - math.hypot(50, 50) / 16) < 1e-12
    assert (p["n"], p["n_lam"]) == (486, 243)
    assert O.rs_plan(101, 101, 0, 0, (51.0, 51.0), cap=0)["n"] == 512
    assert O.rs_plan(101, 101, 0, 0, (51.0, 51.0), cap=256)["n"] == 256


def test_rs_finish_mask_and_mirror():
    n, dlam = 8, 0.5
    fwd = np.arange(64, dtype=np.float64).reshape(8, 8) + 1
    # rows k = 0, 1, 2, 3, 4, -3, -2, -1; applied log-scale -k dlam inside [log 0.5, log 2]: rows 0, 1, 7
    m = O.rs_finish(fwd, None, True, dlam, 0.5, 2.0)
    keep = np.zeros(8, bool)
    keep[[0, 1, 7]] = True
    assert np.array_equal(m[keep], fwd[keep]) and not m[~keep].any()
    inv = np.zeros((8, 8))
    inv[7, 5] = 100.0                 # the mirror of lag (y, x) = (1, 3) is (-1, -3) = (7, 5)
    m = O.rs_finish(fwd, inv, True, dlam, 0.5, 2.0)
    assert m[1, 3] == 0.5 * (fwd[1, 3] + 100.0) and m[1, 2] == 0.5 * fwd[1, 2]
    inv = np.zeros((8, 8))
    inv[7, 3] = 100.0                 # a reflecting pose mirrors the angle lag in place: (7, 3)
    m = O.rs_finish(fwd, inv, False, dlam, 0.5, 2.0)
    assert m[1, 3] == 0.5 * (fwd[1, 3] + 100.0)


@pytest.mark.parametrize("size,deg,scale,mode,symmetric", [
    ((100, 100), 10.0, 0.93, "laplacian", True),     # 486 angles (3-smooth)
    ((96, 120), -8.0, 1.06, "three", False),         # 576 angles
    ((128, 128), 20.0, 1.10, "laplacian", False),    # 864 angles
])
def test_rs_map_recovers_known_rotation_and_scale(size, deg, scale, mode, symmetric):
    h, w = size
    mv, fx, _ = synth.shift_pair((h, w), (0, 0), seed=4)
    c0 = ((w + 1) / 2, (h + 1) / 2)
    H1 = sim_about(c0, math.radians(deg), math.log(scale))
    pa, pb = O.prepare(mv, mode), O.prepare(fx, mode)
    m, info = O.rs_map(pa, pb, H1, symmetric=symmetric)
    assert np.allclose(info["center"], c0)
    assert abs(info["dth"] + math.radians(deg)) <= 2 * math.pi / info["n_th"]
    assert abs(info["dsg"] + math.log(scale)) <= info["dlam"]
    # the correction brings the start pose back to the truth (identity) within a cell's reach
    S = sim_about(info["center"], info["dth"], info["dsg"])
    pts = np.array([[1, 1, 1], [w, 1, 1], [1, h, 1], [w, h, 1]], float).T
    q = S @ H1 @ pts
    assert np.abs(q[:2] / q[2] - pts[:2]).max() < 1.5
