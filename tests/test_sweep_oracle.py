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
