"""Float32 restatement of zcmir's work-image preprocessing (CLAHE, band pass, invert, RGBA to gray,
f16 texels).

TEST INFRASTRUCTURE ONLY. This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it, and
never as the thing measured or shipped. The product path (``zcmir_amd``) never imports it.

Every operation is IEEE float32 in the order the reference's shaders and host code write it
(numpy float32 scalars and arrays round every operation; nothing is fused), so the HIP kernels,
which compile these steps without contraction, match bit for bit. WGSL ``mix(a, b, t)`` is taken
as its specified ``a * (1 - t) + b * t``; the 7-tap Gaussian weights are the float32 roundings of
exp(-k^2 / 2). Whether the reference's compiled shaders round the same way (FMA contraction, the
accuracy of its exp) is parity unpinned: no reference fixture holds a preprocessed image.

Pins (tests/test_pre_oracle.py): the histogram against np.histogram per tile; with an unlimited
clip the tile CDFs against the cumulative histogram; a constant image's clipped CDF by hand; the
Gaussian pass against scipy.ndimage.convolve1d (mode nearest); RGBA weights on pure channels.
"""
from __future__ import annotations

import math

import numpy as np

F = np.float32


def round_half(x: np.ndarray) -> np.ndarray:
    """store_t with f16 texels (gls.zig:433-442): round to binary16, nearest even."""
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def store(x: np.ndarray, half: bool) -> np.ndarray:
    return round_half(x) if half else np.asarray(x, np.float32)


# ── RGBA8 → gray (gls_mift.wgsl:283-291 rgba_to_gray_buf; gls.zig:409-427 uploadRgba) ──────────
UNORM8 = (np.arange(256, dtype=np.float64) / 255.0).astype(np.float32)


def rgba_to_gray(rgba: np.ndarray) -> np.ndarray:
    """(h, w, 4) uint8 → f32 0.299 R + 0.587 G + 0.114 B, left to right."""
    r, g, b = UNORM8[rgba[..., 0]], UNORM8[rgba[..., 1]], UNORM8[rgba[..., 2]]
    return (F(0.299) * r + F(0.587) * g) + F(0.114) * b


# ── CLAHE (gls.zig:462-489 clahe; gls_mift.wgsl:1973-2066) ─────────────────────────────────────
def clahe_params(w: int, h: int, n_bins0: int = 64, grid: int = 4) -> dict:
    hg = min(8, max(2, grid))
    nb = min(256, max(8, n_bins0))
    h_pad = (hg - h % hg) % hg
    w_pad = (hg - w % hg) % hg
    hp, wp = h + h_pad, w + w_pad
    return dict(hg=hg, nb=nb, h_pad=h_pad, w_pad=w_pad, h_tile=max(1, hp // hg), w_tile=max(1, wp // hg),
                hp=hp, wp=wp)


def clahe_reflect(i: np.ndarray, n: int, pad: int) -> np.ndarray:
    """Padded index → image index: the pad sits before the image and mirrors it
    (gls_mift.wgsl:1973-1979 clahe_reflect)."""
    if n <= 1:
        return np.zeros_like(i)
    x = i - pad
    x = np.where(x < 0, -x, x)
    x = np.where(x >= n, 2 * n - 2 - x, x)
    return np.clip(x, 0, n - 1)


def clahe_hist(src: np.ndarray, p: dict) -> np.ndarray:
    """Per-tile histograms (hg * hg, nb) of the padded image (gls_mift.wgsl:1981-2000)."""
    h, w = src.shape
    gy, gx = np.mgrid[0:p["hp"], 0:p["wp"]]
    oy = clahe_reflect(gy, h, p["h_pad"])
    ox = clahe_reflect(gx, w, p["w_pad"])
    v = np.clip(src[oy, ox].astype(np.float32), F(0), F(1))
    nb = p["nb"]
    b = np.clip(v * F(nb - 1), F(0), F(nb - 1)).astype(np.int64)
    tile = (gy // p["h_tile"]) * p["hg"] + (gx // p["w_tile"])
    return np.bincount((tile * nb + b).ravel(), minlength=p["hg"] * p["hg"] * nb).reshape(-1, nb).astype(np.uint32)


def clahe_cdf(hist: np.ndarray, p: dict, clip_limit: float = 40.0) -> np.ndarray:
    """Clipped, redistributed, normalized cumulative histograms (gls_mift.wgsl:2002-2026)."""
    nb = p["nb"]
    voxels = max(p["h_tile"] * p["w_tile"], 1)
    limit = max(F(clip_limit) * F(voxels) / F(nb), F(1))
    out = np.zeros(hist.shape, np.float32)
    for t in range(hist.shape[0]):
        hc = np.minimum(hist[t].astype(np.float32), limit)
        sumc = F(0)
        for i in range(nb):
            sumc = F(sumc + hc[i])
        clipped = F(F(voxels) - sumc)
        nbf = F(nb)
        residual = int(F(clipped - F(nbf * F(math.floor(F(clipped / nbf))))))
        redist = F(F(clipped - F(residual)) / nbf)
        den = max(F(voxels), F(1))
        acc = F(0)
        for i in range(nb):
            c = F(hc[i] + redist)
            if i < residual:
                c = F(c + F(1))
            acc = F(acc + c)
            out[t, i] = min(max(F(acc / den), F(0)), F(1))
    return out


def _mix(a, b, t):
    return a * (F(1) - t) + b * t


def clahe_apply(src: np.ndarray, cdf: np.ndarray, p: dict) -> np.ndarray:
    """Bilinear in the bins, then across the tile grid (gls_mift.wgsl:2028-2066)."""
    h, w = src.shape
    hg, nb = p["hg"], p["nb"]
    v = np.clip(src.astype(np.float32), F(0), F(1))
    y = np.arange(h, dtype=np.float32)[:, None]
    x = np.arange(w, dtype=np.float32)[None, :]
    gy = np.broadcast_to((y + F(0.5)) / F(h) * F(hg - 1), (h, w))
    gx = np.broadcast_to((x + F(0.5)) / F(w) * F(hg - 1), (h, w))
    y0 = np.floor(gy).astype(np.int64)
    x0 = np.floor(gx).astype(np.int64)
    y1 = np.minimum(y0 + 1, hg - 1)
    x1 = np.minimum(x0 + 1, hg - 1)
    fy = gy - y0.astype(np.float32)
    fx = gx - x0.astype(np.float32)
    bf = v * F(nb - 1)
    b0 = np.clip(np.floor(bf), F(0), F(nb - 2)).astype(np.int64)
    fb = bf - b0.astype(np.float32)
    b1 = b0 + 1

    def at(ty, tx):
        t = ty * hg + tx
        return _mix(cdf[t, b0], cdf[t, b1], fb)

    c0 = _mix(at(y0, x0), at(y0, x1), fx)
    c1 = _mix(at(y1, x0), at(y1, x1), fx)
    return _mix(c0, c1, fy).astype(np.float32)


def clahe(src: np.ndarray, clip_limit: float = 40.0, n_bins: int = 64, grid: int = 4, half: bool = False):
    h, w = src.shape
    p = clahe_params(w, h, n_bins, grid)
    cdf = clahe_cdf(clahe_hist(src, p), p, clip_limit)
    return store(clahe_apply(src, cdf, p), half)


# ── band pass (gls.zig:441-448 gauss, 502-545 blurPasses + bandpass; gls_mift.wgsl:293-318) ────
GAUSS7 = np.exp(-0.5 * np.arange(-3, 4, dtype=np.float64) ** 2).astype(np.float32)


def gauss_pass(src: np.ndarray, axis: int, half: bool) -> np.ndarray:
    """sigma 1, radius 3, clamped borders: sum_k at(x + k) w_k / sum_k w_k, k = -3..3 in order."""
    h, w = src.shape
    acc = np.zeros((h, w), np.float32)
    wsum = F(0)
    for i, k in enumerate(range(-3, 4)):
        if axis == 1:
            idx = np.clip(np.arange(w) + k, 0, w - 1)
            tap = src[:, idx]
        else:
            idx = np.clip(np.arange(h) + k, 0, h - 1)
            tap = src[idx, :]
        acc = acc + tap.astype(np.float32) * GAUSS7[i]
        wsum = F(wsum + GAUSS7[i])
    return store(acc / wsum, half)


def blur_passes(src: np.ndarray, passes: int, half: bool) -> np.ndarray:
    out = np.asarray(src, np.float32)
    for _ in range(passes):
        out = gauss_pass(gauss_pass(out, 1, half), 0, half)
    return out


def band_counts(fine_px: float = 1.0, coarse_px: float = 6.0):
    """Gaussian passes of the fine and coarse blurs (gls.zig:521-522; Zig @round: half away from 0)."""
    rnd = lambda v: math.floor(v + 0.5)
    fine_n = int(max(1, min(16, rnd((fine_px if fine_px > 0 else 1) ** 2))))
    coarse_n = int(max(fine_n + 1, min(64, rnd((coarse_px if coarse_px > 0 else 6) ** 2))))
    return fine_n, coarse_n


def bandpass(work: np.ndarray, fine_px: float = 1.0, coarse_px: float = 6.0, half: bool = False):
    """Difference of the fine and coarse blurs (each clamped to [0, 1]), rescaled to [0, 1]."""
    fine_n, coarse_n = band_counts(fine_px, coarse_px)
    fine = blur_passes(work, fine_n, half)
    coarse = blur_passes(fine, coarse_n - fine_n, half)
    v = np.clip(fine, F(0), F(1)) - np.clip(coarse, F(0), F(1))
    lo, hi = v.min(), v.max()
    sc = F(F(1) / F(hi - lo)) if hi > lo else F(0)
    return store((v - lo) * sc, half)


def invert(work: np.ndarray, half: bool = False) -> np.ndarray:
    """1 - clamp(v, 0, 1) (gls.zig:547-554)."""
    return store(F(1) - np.clip(np.asarray(work, np.float32), F(0), F(1)), half)


def work_image(gray: np.ndarray, preprocess: bool = True, half: bool = True, clahe_on: bool = True,
               clahe_bins: int = 64, clahe_grid: int = 4, band: bool = False, bp_fine: float = 1.0,
               bp_coarse: float = 6.0, invert_on: bool = False) -> np.ndarray:
    """The work image of a side (engine.zig:321-340 bake): CLAHE of the f32 gray (clip 40) or the
    gray's texels, then band pass, then invert, each stored as texels."""
    g = np.asarray(gray, np.float32)
    if preprocess and clahe_on:
        w = clahe(g, 40.0, clahe_bins, clahe_grid, half)
    else:
        w = store(g, half)
    if preprocess and band:
        w = bandpass(w, bp_fine, bp_coarse, half)
    if preprocess and invert_on:
        w = invert(w, half)
    return w
