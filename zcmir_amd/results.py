"""Result records returned by ``Registrar`` (field names as the reference package's, so code that
reads them keeps working: python/zcmir/__init__.py:59-160). Each record is built from the C
result struct of its entry point (``include/zcsmi.h``)."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def as_pose(values) -> np.ndarray:
    """Nine numbers (row-major) → a 3×3 float64 pose."""
    return np.array(list(values), dtype=np.float64).reshape(3, 3)


def _nan_none(x: float):
    return None if math.isnan(x) else float(x)


def _translation(dx: float, dy: float) -> np.ndarray:
    return np.array([[1.0, 0.0, dx], [0.0, 1.0, dy], [0.0, 0.0, 1.0]])


def _similarity_about(center, angle: float, log_scale: float) -> np.ndarray:
    """x ↦ c + e^log_scale · R(angle) · (x − c) as a 3×3 pose."""
    k = math.exp(log_scale)
    ca, sa = k * math.cos(angle), k * math.sin(angle)
    cx, cy = center
    lin = np.array([[ca, -sa], [sa, ca]])
    out = np.eye(3)
    out[:2, :2] = lin
    out[:2, 2] = np.array([cx, cy]) - lin @ np.array([cx, cy])
    return out


@dataclass
class Score:
    """SMI of a pose. ``mean`` is what the optimizers maximize: the average of the forward
    (moving → fixed) and inverse directions when the symmetric setting is on, else the forward
    score; ``inv`` is None for a one-directional score."""
    mean: float
    fwd: float
    inv: float | None

    @classmethod
    def from_triple(cls, v):
        return cls(float(v[0]), float(v[1]), _nan_none(float(v[2])))


@dataclass
class Gradient:
    """Score, overlap count ``n``, tangent gradient over the group's coordinates and, on request,
    the Gauss–Newton matrix (``hess``, nk × nk)."""
    score: float
    fwd: float
    inv: float | None
    n: float
    grad: np.ndarray
    hess: np.ndarray | None

    @classmethod
    def from_c(cls, r):
        k = int(r.nk)
        h = np.array(r.hess[: k * k]).reshape(k, k) if r.has_hess else None
        return cls(r.score, r.fwd, _nan_none(r.inv), r.n, np.array(r.grad[:k]), h)


@dataclass
class ShiftMap:
    """The score at every integer translation (lag) of the moving image about a pose.

    ``map`` has ``ny`` rows and ``n`` columns; entry [y, x] is the lag (x, y), read with wrap-around
    (columns past n/2 are negative lags; ``np.fft.fftshift`` puts lag 0 in the middle). A lag step
    is canvas_w / cw fixed pixels across and canvas_h / ch down. (dx, dy) is the best lag in fixed
    pixels; ``corrected(H)`` applies it."""
    peak: float
    dx: float
    dy: float
    zero: float
    n: int
    cw: int
    ch: int
    canvas: tuple
    offset: tuple
    corr_area_px: float
    exact: bool
    calibrated: bool
    map: np.ndarray | None
    ny: int = 0

    @classmethod
    def from_c(cls, r, values=None):
        rows = int(r.n_y or r.n)
        grid = None if values is None else values[: r.n * rows].reshape(rows, r.n).copy()
        return cls(peak=r.peak, dx=r.dx, dy=r.dy, zero=r.zero, n=int(r.n), cw=int(r.cw), ch=int(r.ch),
                   canvas=(int(r.canvas_w), int(r.canvas_h)), offset=(int(r.ox), int(r.oy)),
                   corr_area_px=r.corr_area_px, exact=bool(r.exact), calibrated=bool(r.calibrated), map=grid,
                   ny=rows)

    def corrected(self, H):
        """H moved by the best lag: T(−dx, −dy) · H."""
        return _translation(-self.dx, -self.dy) @ np.asarray(H, np.float64).reshape(3, 3)


@dataclass
class RotoScaleMap:
    """The score over rotations and log-scales of the moving image about ``center`` (fixed pixels,
    1-based). ``map`` is n × n: angle lag across, log-radius lag down, lag 0 at [0, 0] with
    wrap-around; ``fwd`` / ``inv`` are the two directions' maps of a symmetric map. (dth, dsg) is
    the best rotation (radians) and log-scale."""
    peak: float
    dth: float
    dsg: float
    zero: float
    center: tuple
    n: int
    n_th: int
    n_lam: int
    dlam: float
    r0: float
    r1: float
    symmetric: bool
    calibrated: bool
    map: np.ndarray | None
    fwd: np.ndarray | None
    inv: np.ndarray | None

    @classmethod
    def from_c(cls, r, maps=(None, None, None)):
        side = int(r.n)

        def grid(b):
            return None if b is None else b[: side * side].reshape(side, side).copy()

        return cls(peak=r.peak, dth=r.dth, dsg=r.dsg, zero=r.zero, center=(r.cx, r.cy), n=side, n_th=int(r.n_th),
                   n_lam=int(r.n_lam), dlam=r.dlam, r0=r.r0, r1=r.r1, symmetric=bool(r.symmetric),
                   calibrated=bool(r.calibrated), map=grid(maps[0]), fwd=grid(maps[1]),
                   inv=grid(maps[2]) if r.symmetric else None)

    def corrected(self, H):
        """H turned by dth and scaled by e^dsg about the centre."""
        return _similarity_about(self.center, self.dth, self.dsg) @ np.asarray(H, np.float64).reshape(3, 3)


@dataclass
class Climb:
    """Where a climb ended: the pose, its score, iterations run and whether the pose changed."""
    H: np.ndarray
    score: float
    iterations: int
    moved: bool

    @classmethod
    def from_c(cls, r):
        return cls(as_pose(r.H), r.score, int(r.iterations), bool(r.moved))


@dataclass
class Search:
    """Outcome of a global search: the best pose found (the engine's pose when ``improved``), its
    score and the score it had to beat; count_a / count_b are the sweep's peaks after NMS and its
    climbed finalists."""
    H: np.ndarray
    score: float
    previous: float
    improved: bool
    count_a: int
    count_b: int

    @classmethod
    def from_c(cls, r):
        return cls(as_pose(r.H), r.score, r.prev, bool(r.improved), int(r.count_a), int(r.count_b))
