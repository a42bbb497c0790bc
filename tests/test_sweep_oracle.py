"""Pinning of the oracle's rotation × scale sweep (CPU only).

The sweep's pieces are checked against what the mathematics fixes: singular values against
numpy's SVD, the candidate axes against their closed forms, the peak rule against a brute-force
restatement, NMS on hand-built peaks, and the whole sweep against an exact geometric case (a
90° rotation by array transpose, which involves no interpolation) whose pose is known.
"""
import math

import numpy as np
import pytest

from oracle import smi_ref as O
from zcmir_amd import synth


def test_sv2_matches_svd():
    rng = np.random.default_rng(0)
    for _ in range(50):
        L = rng.standard_normal(4)
        s = np.linalg.svd(L.reshape(2, 2), compute_uv=False)
        assert np.allclose(O.sv2(list(L)), s, rtol=1e-12, atol=1e-12)
    # candLinear: σ R(θ) diag(a, 1/a) [1 k; 0 1]
    th, s, k, a = 0.7, 1.3, 0.2, 1.1
    R = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
    M = s * R @ np.diag([a, 1 / a]) @ np.array([[1, k], [0, 1]])
    assert np.allclose(O.cand_linear(th, s, k, a), M.ravel(), atol=1e-14)


def test_sweep_grid_axes():
    g = O.sweep_grid(-180, 180, 72, 0.8, 1.25, 9)
    assert g["wrap"] and len(g["thetas"]) == 72
    assert np.allclose(np.rad2deg(g["thetas"]), -180 + 5 * np.arange(72))
    assert np.allclose(g["sigmas"], 1.25 ** (np.arange(-4, 5) / 4))   # log-spaced 0.8 .. 1.25
    assert g["shears"] == [0.0] and g["anisos"] == [1.0]
    p = O.sweep_grid(-10, 20, 4, 1.0, 1.0, 1, shear_max=0.1, n_shear=3)
    assert not p["wrap"] and np.allclose(np.rad2deg(p["thetas"]), [-10, 0, 10, 20])
    assert np.allclose(p["shears"], [-0.1, 0, 0.1]) and p["sigmas"] == [1.0]
    one = O.sweep_grid(30, 30, 1, 1.0, 1.21, 1)
    assert np.isclose(np.rad2deg(one["thetas"][0]), 30) and np.isclose(one["sigmas"][0], 1.1)


def _peaks_brute(m, n_peak):
    ny, nx = m.shape
    cls = {}
    for t in range(nx * ny):
        y, x = divmod(t, nx)
        v = m[y, x]
        best = cls.get(t % 256, (0.0, 0xFFFFFFFF))
        if not v > best[0]:
            continue
        if all(not m[(y + dy) % ny, (x + dx) % nx] > v for dy in (-1, 0, 1) for dx in (-1, 0, 1) if dx or dy):
            cls[t % 256] = (float(v), t)
    out = [cls.get(c, (0.0, 0xFFFFFFFF)) for c in range(256)]
    out.sort(key=lambda e: (-e[0], e[1]))
    return out[:n_peak]


@pytest.mark.parametrize("shape", [(32, 32), (16, 64), (40, 24)])
def test_sweep_peaks_rule(shape):
    rng = np.random.default_rng(shape[0])
    m = np.maximum(rng.standard_normal(shape), 0)
    m[rng.uniform(size=shape) < 0.3] = 0
    m[3, 5] = m[3, 6] = 9.0          # a plateau: both cells are local maxima (no neighbour exceeds)
    assert O.sweep_peaks(m, 4) == _peaks_brute(m, 4)
    assert O.sweep_peaks(np.zeros(shape), 2) == [(0.0, 0xFFFFFFFF)] * 2


def test_nms4_rules():
    def pk(score, it, is_, sx, sy):
        return dict(score=score, it=it, is_=is_, ik=0, ia=0, sx=sx, sy=sy)
    peaks = [pk(10, 0, 0, 0, 0), pk(9, 71, 1, 2, 2),     # θ wraps: 71 is 1 step from 0 → dropped
             pk(8, 3, 0, 0, 0),                          # 3 θ steps away → kept
             pk(7, 0, 0, 5, 0),                          # 5 cells away → kept
             pk(6, 1, 3, 0, 0)]                          # 3 σ steps away → kept
    kept = O.nms4(peaks, 72, 32, True)
    assert [p["score"] for p in kept] == [10, 8, 7, 6]
    assert [p["score"] for p in O.nms4(peaks, 72, 32, False)] == [10, 9, 8, 7, 6]
    assert len(O.nms4(peaks, 72, 2, True)) == 2


def test_sweep_recovers_exact_quarter_turn():
    """moving = transpose-rotated fixed (exact, no interpolation) with a 1 − x² remap: the best
    candidate of a 30°-step sweep is θ = 90°, σ = 1, at zero shift, i.e. H = about_c(R(90°))."""
    n = 96
    _, fixed, _ = synth.shift_pair((n, n), (0, 0), seed=5)
    moving = np.rot90(fixed).copy()
    moving = 1.0 - moving * moving
    pm, pf = O.prepare(moving), O.prepare(fixed)
    gv = O.sweep_grid(-180, 180, 12, 1.0, 1.0, 1)
    r = O.sweep(pm, pf, np.eye(3), gv, grid=32)
    top = O.nms4(r["peaks"], 12, 8, True)
    assert np.isclose(np.rad2deg(top[0]["th"]), 90.0) and top[0]["s"] == 1.0
    assert (top[0]["sx"], top[0]["sy"]) == (0, 0)
    truth = np.array([[0, -1, n + 1], [1, 0, 0], [0, 0, 1.0]])
    assert np.allclose(top[0]["H"], truth, atol=1e-9)
    # best per candidate: θ-major; the true candidate is the landscape maximum
    assert int(np.argmax(r["best"])) == 9


# ── the sweep's fractional-cell tap average (sweep.wgsl:64-85 tap, 106-145 pack_sweep) ──
def _tap_points(cx, k, cell, o):
    """0-based source coordinates of a cell's k taps along one axis under an identity pose: the
    canvas point o + (c + (t + ½)/k)·cell (1-based pixels), minus 1."""
    return o + (cx + (np.arange(k) + 0.5) / k) * cell - 1.0


@pytest.mark.parametrize("k,cell,ox,oy", [(3, 1.37, -2.25, 3.5), (4, 2.6, 0.0, 0.0), (1, 0.83, 5.1, -1.7)])
def test_sweep_window_constant_planes_count_inside_taps(k, cell, ox, oy):
    """Constant feature planes: every tap whose four bilinear corners lie inside the source adds
    weight 1 and the constant; the others add nothing. So a cell's weight plane is the fraction of
    its k×k taps inside (0 ≤ x < w − 1, 0 ≤ y < h − 1 in 0-based coordinates) and each feature
    plane that fraction times its constant, with the taps at (c + (t + ½)/k)·cell + o − 1."""
    h, w = 19, 23
    consts = np.array([0.5, -1.25, 2.0, 0.125])
    feat = consts[:, None, None] * np.ones((4, h, w))
    grid = 24
    got = O.sweep_window(feat, np.eye(3), 0, 0, grid, grid, grid, k, ox, oy, cell)
    for cy in range(grid):
        ins_y = [(0 <= y < h - 1) for y in _tap_points(cy, k, cell, oy)]
        for cx in range(grid):
            ins_x = [(0 <= x < w - 1) for x in _tap_points(cx, k, cell, ox)]
            frac = sum(ix and iy for iy in ins_y for ix in ins_x) / (k * k)
            assert abs(got[0, cy, cx] - frac) < 1e-12, (cx, cy)
            assert np.abs(got[1:, cy, cx] - consts * frac).max() < 1e-12, (cx, cy)
    assert 0 < got[0].min() < 1 or got[0].min() == 0   # the window reaches past the source


@pytest.mark.parametrize("k,cell", [(3, 1.37), (2, 0.9), (5, 2.2)])
def test_sweep_window_reproduces_ramps_at_cell_centres(k, cell):
    """Bilinear taps reproduce a linear function, and a cell's taps are symmetric about its
    centre, so for a ramp f(x, y) = a + b x + c y (0-based source pixels) every cell whose taps are
    all inside returns f at the image of the cell's centre: under an affine pose Hi (canvas →
    source, 1-based) that is Hi · (o + (c + ½)·cell) − 1. A tap offset of t/k instead of
    (t + ½)/k, a dropped 1-based shift or a transposed pose all move it."""
    h, w = 60, 70
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    coef = [(0.3, 0.02, -0.05), (-1.0, 0.11, 0.07), (2.0, -0.03, 0.01), (0.0, 0.05, 0.05)]
    feat = np.stack([a + b * x + c * y for a, b, c in coef])
    t = np.deg2rad(11.0)
    Hi = np.array([[0.97 * np.cos(t), -0.97 * np.sin(t), 9.0], [0.97 * np.sin(t), 0.97 * np.cos(t), 4.0],
                   [0.0, 0.0, 1.0]])
    ox, oy, grid = 1.5, -0.5, 40
    got = O.sweep_window(feat, Hi, 0, 0, grid, grid, grid, k, ox, oy, cell)
    checked = 0
    for cy in range(grid):
        for cx in range(grid):
            if got[0, cy, cx] != 1.0:
                continue   # a tap outside the source: not a full cell
            px, py = ox + (cx + 0.5) * cell, oy + (cy + 0.5) * cell
            sx, sy = Hi[0, 0] * px + Hi[0, 1] * py + Hi[0, 2] - 1, Hi[1, 0] * px + Hi[1, 1] * py + Hi[1, 2] - 1
            for p, (a, b, c) in enumerate(coef):
                assert abs(got[1 + p, cy, cx] - (a + b * sx + c * sy)) < 1e-9, (k, cell, cx, cy, p)
            checked += 1
    assert checked > 200


def test_cpu_port_sweep_matches_oracle():
    """The float32 CPU port of the sweep (the bench's CPU baseline) computes the oracle's sweep:
    every candidate's best peak within float32 rounding."""
    from oracle import cpu_port

    mv, fx, _ = synth.rotscale_pair((96, 96), theta_deg=20.0, scale=1.05, seed=4)
    pm, pf = O.prepare(mv), O.prepare(fx)
    gv = O.sweep_grid(-40, 40, 3, 0.9, 1.1, 2)
    r = O.sweep(pm, pf, np.eye(3), gv, grid=32)
    done, best = cpu_port.sweep_f32(pm, pf, np.eye(3), gv, grid=32, workers=2)
    assert done == 6
    assert np.abs(best - r["best"]).max() <= 2e-4 * r["best"].max()
