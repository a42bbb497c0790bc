"""Synthetic image pairs of the SMI workloads (seeded, reproducible).

These are inputs for benchmarks, tests and examples; no part of the compute path depends on
them. Workloads:

* ``shift_pair``: fixed = Gaussian-filtered noise (sigma 4 px, unit std) + 20 Gaussian blobs,
  min-max scaled to [0, 1]; moving(x) = 1 - fixed(x - d)^2 + N(0, 0.02), clipped to [0, 1].
  Built on a larger canvas and cropped, so nothing wraps. The true pose (moving → fixed) is
  translate(-dx, -dy) and the shift map's peak lag is (dx, dy). ``quantize`` rounds both to
  k / 255 (8-bit intensities).
* ``rotscale_pair``: moving = fixed rotated by theta and scaled by s about the image centre
  (bilinear), then remapped by x^0.5.
* ``stack_slices``: uneven-size slices, each with a random homography and intensity remap, the
  moving frame cut from the same texture as the fixed one (no empty border).
* ``drift_stack``: square slices whose texture and pose drift slowly from slice to slice, the
  fixed side another modality (1 - x^0.6 plus noise) - the design of the reference's stack test
  (tests/test_api.py:186-202), re-created here.
"""
from __future__ import annotations

import math

import numpy as np


def _texture(h: int, w: int, rng: np.random.Generator, n_blobs: int = 20) -> np.ndarray:
    from scipy.ndimage import gaussian_filter

    f = gaussian_filter(rng.standard_normal((h, w)), 4.0, mode="wrap")
    f /= max(f.std(), 1e-12)
    for _ in range(n_blobs):
        cy, cx = rng.uniform(0, h), rng.uniform(0, w)
        sg = rng.uniform(2.0, max(3.0, min(h, w) / 12.0))
        amp = rng.uniform(-3.0, 3.0)
        # the blob within 9 sigma of its centre (beyond, each term is below e^-40.5 of amp)
        y0, y1 = max(0, int(cy - 9 * sg)), min(h, int(cy + 9 * sg) + 2)
        x0, x1 = max(0, int(cx - 9 * sg)), min(w, int(cx + 9 * sg) + 2)
        if y0 < y1 and x0 < x1:
            gy = np.exp(-((np.arange(y0, y1, dtype=np.float64) - cy) ** 2) / (2 * sg * sg))
            gx = np.exp(-((np.arange(x0, x1, dtype=np.float64) - cx) ** 2) / (2 * sg * sg))
            f[y0:y1, x0:x1] += amp * np.outer(gy, gx)
    f -= f.min()
    f /= max(f.max(), 1e-12)
    return f


def shift_pair(size=(256, 256), shift=(17, -9), seed: int = 0, quantize: bool = False,
               noise: float = 0.02):
    """(moving, fixed, H_true) float32 images of ``size`` = (h, w) (see module docstring)."""
    h, w = size
    dx, dy = shift
    m = max(abs(dx), abs(dy)) + 8
    rng = np.random.default_rng(seed)
    big = _texture(h + 2 * m, w + 2 * m, rng)
    fixed = big[m:m + h, m:m + w]
    src = big[m - dy:m - dy + h, m - dx:m - dx + w]
    moving = 1.0 - src * src + rng.normal(0.0, noise, (h, w))
    moving = np.clip(moving, 0.0, 1.0)
    if quantize:
        fixed = np.round(fixed * 255.0) / 255.0
        moving = np.round(moving * 255.0) / 255.0
    H = np.array([[1.0, 0.0, -dx], [0.0, 1.0, -dy], [0.0, 0.0, 1.0]])
    return moving.astype(np.float32), fixed.astype(np.float32), H


def _bilinear(img: np.ndarray, fx: np.ndarray, fy: np.ndarray) -> np.ndarray:
    h, w = img.shape
    x0 = np.floor(fx).astype(np.int64)
    y0 = np.floor(fy).astype(np.int64)
    ax, ay = fx - x0, fy - y0
    inside = (fx >= 0) & (fy >= 0) & (fx <= w - 1) & (fy <= h - 1)
    xa, xb = np.clip(x0, 0, w - 1), np.clip(x0 + 1, 0, w - 1)
    ya, yb = np.clip(y0, 0, h - 1), np.clip(y0 + 1, 0, h - 1)
    v = (img[ya, xa] * (1 - ax) + img[ya, xb] * ax) * (1 - ay) + (img[yb, xa] * (1 - ax) + img[yb, xb] * ax) * ay
    return np.where(inside, v, 0.0)


def warp(img: np.ndarray, H: np.ndarray, out_shape) -> np.ndarray:
    """img (moving frame) resampled into a frame of out_shape through H (moving → out, 1-based)."""
    ho, wo = out_shape
    y, x = np.mgrid[0:ho, 0:wo].astype(np.float64)
    X = np.stack([x + 1, y + 1, np.ones_like(x)])
    q = np.tensordot(np.linalg.inv(H), X, 1)
    return _bilinear(np.asarray(img, np.float64), q[0] / q[2] - 1, q[1] / q[2] - 1)


def about(c, A):
    """translate(c) · A · translate(-c) for a 2x2 A (3x3 result)."""
    T = np.eye(3)
    T[:2, :2] = A
    cx, cy = c
    return np.array([[1, 0, cx], [0, 1, cy], [0, 0, 1.0]]) @ T @ np.array([[1, 0, -cx], [0, 1, -cy], [0, 0, 1.0]])


def rot(theta_deg: float, s: float = 1.0) -> np.ndarray:
    t = math.radians(theta_deg)
    return s * np.array([[math.cos(t), -math.sin(t)], [math.sin(t), math.cos(t)]])


def rotscale_pair(size=(1024, 1024), theta_deg: float = 37.0, scale: float = 1.12, seed: int = 2,
                  remap=np.sqrt):
    """(moving, fixed, H_true): moving = fixed rotated theta and scaled s about the centre
    (bilinear), then remapped. H_true maps moving → fixed (1-based)."""
    h, w = size
    rng = np.random.default_rng(seed)
    fixed = _texture(h, w, rng)
    c = ((w + 1) / 2.0, (h + 1) / 2.0)
    T = about(c, rot(theta_deg, scale))          # fixed → moving
    moving = remap(np.clip(warp(fixed, T, (h, w)), 0.0, 1.0))
    return moving.astype(np.float32), fixed.astype(np.float32), np.linalg.inv(T)


REMAPS = {
    "1-x^2": lambda x: 1.0 - x * x,
    "x^0.5": np.sqrt,
    "sin(pi x)": lambda x: np.sin(np.pi * x),
}


def random_homography(rng: np.random.Generator, w: int, h: int, rot_deg=10.0, scale=(0.9, 1.1),
                      persp=1e-4, shift_frac=0.03):
    """A random homography about the image centre: rotation ±rot_deg, isotropic scale, a small
    translation and perspective terms of magnitude <= persp."""
    c = ((w + 1) / 2.0, (h + 1) / 2.0)
    A = rot(rng.uniform(-rot_deg, rot_deg), rng.uniform(*scale))
    H = about(c, A)
    H[0, 2] += rng.uniform(-shift_frac, shift_frac) * w
    H[1, 2] += rng.uniform(-shift_frac, shift_frac) * h
    P = np.eye(3)
    P[2, 0] = rng.uniform(-persp, persp)
    P[2, 1] = rng.uniform(-persp, persp)
    return about(c, np.eye(2)) @ (np.array([[1, 0, c[0]], [0, 1, c[1]], [0, 0, 1]]) @ P @ np.array([[1, 0, -c[0]], [0, 1, -c[1]], [0, 0, 1.0]])) @ H


def stack_slices(n: int = 64, side=(512, 1536), max_aspect: float = 1.5, seed: int = 3):
    """[(moving, fixed, H_true, remap name)] uneven-size slices (see module docstring).

    Both frames are cut from one texture that extends 0.35 · max(h, w) + 8 px past the fixed
    frame on every side, so every moving pixel holds real content: moving(X) = remap(tex(H_true X))
    with the fixed frame at the texture's offset. (Filling the part of the moving frame that falls
    outside the fixed image with zeros instead leaves an artificial edge there whose gradient and
    Laplacian features pull the SMI optimum a median 0.8 px off H_true on these slices.)"""
    rng = np.random.default_rng(seed)
    out = []
    names = list(REMAPS)
    for _ in range(n):
        a = int(rng.integers(side[0], side[1] + 1))
        asp = rng.uniform(1.0, max_aspect)
        b = int(max(side[0] // 2, min(side[1], round(a / asp))))
        h, w = (a, b) if rng.uniform() < 0.5 else (b, a)
        m = int(0.35 * max(h, w)) + 8
        tex = _texture(h + 2 * m, w + 2 * m, rng)
        fixed = tex[m:m + h, m:m + w]
        Ht = random_homography(rng, w, h)         # moving → fixed
        to_tex = np.array([[1.0, 0.0, m], [0.0, 1.0, m], [0.0, 0.0, 1.0]]) @ Ht
        mv = warp(tex, np.linalg.inv(to_tex), (h, w))
        name = names[int(rng.integers(len(names)))]
        moving = np.clip(REMAPS[name](np.clip(mv, 0.0, 1.0)), 0.0, 1.0)
        out.append((moving.astype(np.float32), fixed.astype(np.float32), Ht, name))
    return out


def rect_texture(n: int, seed: int) -> np.ndarray:
    """An n × n texture with strong edges: smooth power-law noise under 80 randomly rotated
    rectangles of random flat intensity, plus N(0, 0.02) noise, in [0, 1]."""
    rng = np.random.default_rng(seed)
    fy, fx = np.meshgrid(np.fft.fftfreq(n), np.fft.fftfreq(n), indexing="ij")
    r = np.hypot(fy, fx)
    r[0, 0] = 1.0
    spec = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * r ** -1.6
    spec[0, 0] = 0.0
    base = np.fft.ifft2(spec).real
    img = 0.6 * (base - base.min()) / max(np.ptp(base), 1e-12)
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float64)
    for _ in range(80):
        cx, cy = rng.uniform(0, n, 2)
        half_w, half_h = rng.uniform(6, 40, 2)
        t = rng.uniform(0, math.pi)
        u = (xx - cx) * math.cos(t) + (yy - cy) * math.sin(t)
        v = (yy - cy) * math.cos(t) - (xx - cx) * math.sin(t)
        img[(np.abs(u) < half_w) & (np.abs(v) < half_h)] = rng.uniform(0, 1)
    return np.clip(img + 0.02 * rng.standard_normal((n, n)), 0.0, 1.0)


def drift_stack(n_slices: int = 5, n: int = 384, seed: int = 11):
    """[(moving, fixed, H_true)]: slice k's moving image mixes a base texture with its own
    (85 / 15), its pose drifts (rotation -6 + 0.4 k degrees, scale 1.04 + 0.004 k, translation
    (10 + 1.5 k, -7 + k), perspective (2e-5, -1e-5) about the centre), and the fixed image is
    1 - moving^0.6 warped by the pose, plus N(0, 0.03), zero outside the moving image."""
    base = rect_texture(n, seed)
    rng = np.random.default_rng(seed + 1)
    c = ((n + 1) / 2.0, (n + 1) / 2.0)
    out = []
    for k in range(n_slices):
        img = np.clip(0.85 * base + 0.15 * rect_texture(n, 100 + k) + 0.01 * rng.standard_normal((n, n)), 0.0, 1.0)
        R = np.eye(3)
        R[:2, :2] = rot(-6.0 + 0.4 * k, 1.04 + 0.004 * k)
        R[2, :2] = (2e-5, -1e-5)
        C = np.array([[1, 0, c[0]], [0, 1, c[1]], [0, 0, 1.0]])
        H = np.array([[1, 0, 10.0 + 1.5 * k], [0, 1, -7.0 + k], [0, 0, 1.0]]) @ C @ R @ np.linalg.inv(C)
        src = warp(img, H, (n, n))
        inside = warp(np.ones_like(img), H, (n, n)) > 0.999
        fixed = np.where(inside, 1.0 - np.clip(src, 0, 1) ** 0.6, 0.0)
        fixed = np.clip(fixed + 0.03 * rng.standard_normal((n, n)), 0.0, 1.0)
        out.append((img.astype(np.float32), fixed.astype(np.float32), H))
    return out
