"""Float64 NumPy oracle of zcmir's SMI hot path.

TEST INFRASTRUCTURE ONLY. This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it, and
never as the thing measured or shipped. The product path (``zcmir_amd``) never imports it.

Every function restates one piece of the reference algorithm and cites the reference file and
lines it follows (paths relative to the zcmir repository). Arithmetic is float64 except where
the reference's integer/bit-exact steps are defined (the rank transform), which are restated in
float32 / integer arithmetic exactly as the HIP kernels perform them.

Pinning (what checks each function beyond retyping it) is listed in ``tests/test_oracle.py``;
functions with no independent check say "parity unpinned" in their docstring.
"""
from __future__ import annotations

import math

import numpy as np

MIN_N = 12            # smi.zig:17
RIDGE = 1e-4          # smi.zig:18, moments.zig:22
N_BINS = 256          # smi.zig:19
MAP_CAP = 1024        # smi.zig:22
MAP_AUTO_MIN = 128    # smi.zig:29
MAP_AUTO_MAX = 512    # smi.zig:30
TAPER = 2.0           # smi.wgsl:49

#: upper-triangle index pairs of a 4x4 (smi.wgsl:60-66 tri_i / tri_j)
TRI = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]

FEATURE_MODES = ("laplacian", "three", "cubic")   # moments.zig:13 SmiFeatures


def tri_k(i: int, j: int) -> int:
    """Row-major upper-triangle index of (i, j) (smi.wgsl:67-72, moments.zig:31-35)."""
    a, b = min(i, j), max(i, j)
    return a * 4 - (a * (a + 1)) // 2 + b


# ── exact image statistics (the inputs of the bit-exact rank) ───────────────────────────────
# The reference finishes the image mean and spread with f32 GPU reductions (smi.zig:366-375,
# smi.wgsl:521-529, 532-551, 1541-1570), whose rounding depends on the reduction order. This
# build defines both from exact integer sums so the GPU and this oracle produce the same f32
# mean and spread, and therefore the same histogram bins, bit for bit:
#   K      = 62 - e, where max|x| < 2^e (frexp), so |trunc(x 2^K)| < 2^62
#   S1     = sum_i trunc(x_i 2^K)                           (exact integer)
#   mean   = f32( ((double)hi 2^64 + (double)lo) / n  2^-K ),  S1 = hi 2^64 + lo
#   d_i    = f32(x_i - mean)
#   K2     = 60 - 2e  (d^2 < 2^(2e+2))
#   S2     = sum_i trunc(d_i^2 2^K2)                         (d_i^2 exact in f64)
#   spread = f32( sqrt(max(S2 / n 2^-K2, 1e-12)) )           (smi.wgsl:1567-1570 finish_std)

def _exponent(maxabs: float) -> int:
    return math.frexp(maxabs)[1] if maxabs > 0 else 0


def _exact_trunc_sum(values: np.ndarray, K: int) -> int:
    t = np.trunc(np.ldexp(values.astype(np.float64), K)).astype(np.int64)
    hi = t >> 32
    lo = t & 0xFFFFFFFF
    return int(hi.sum(dtype=np.int64)) * (1 << 32) + int(lo.sum(dtype=np.int64))


def _fixed_to_f64(S: int, n: int, K: int) -> float:
    hi = S >> 64
    lo = S & ((1 << 64) - 1)
    v = (float(hi) * 18446744073709551616.0 + float(lo)) / float(n)
    return math.ldexp(v, -K)


def image_stats(x: np.ndarray) -> tuple[np.float32, np.float32]:
    """The f32 mean and spread of a gray image, from exact integer sums (see above).

    Restates smi.zig:366-375 (reduce1 → finish_sum, zscore_p, feat_moments → finish_std) with
    an order-independent definition of the two sums.
    """
    x = np.ascontiguousarray(x, np.float32).ravel()
    n = x.size
    maxabs = float(np.max(np.abs(x))) if n else 0.0
    e = _exponent(maxabs)
    K = 62 - e
    S1 = _exact_trunc_sum(x, K)
    mean = np.float32(_fixed_to_f64(S1, n, K))
    d = (x - mean).astype(np.float32)
    K2 = 60 - 2 * e
    d64 = d.astype(np.float64)
    S2 = _exact_trunc_sum(d64 * d64, K2)
    var = _fixed_to_f64(S2, n, K2)
    spread = np.float32(math.sqrt(max(var, 1e-12)))
    return mean, spread


def round_half(x: np.ndarray) -> np.ndarray:
    """The f16 work image (settings ``half``): gray rounded to binary16, round-to-nearest-even
    (gls.zig:433-442 uploadGray + packTexels/promoteTexels)."""
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def rank_transform(gray: np.ndarray):
    """Histogram-equalized rank of a gray image (smi.wgsl:394-420 hist_bins, cdf_scan, rank_eq).

    z = (x - mean) / max(spread, 1e-8) in f32; bin = trunc(clamp((z + 4) 32, 0, 255));
    exclusive CDF; r = (cdf[b] + 0.5 hist[b]) / n * 2 - 1 in f32.
    Returns (rank f32 (h, w), hist u32[257], cdf u32[258], bins u32 (h, w)).
    """
    g = np.ascontiguousarray(gray, np.float32)
    mean, spread = image_stats(g)
    d = (g - mean).astype(np.float32)
    z = (d / np.maximum(spread, np.float32(1e-8))).astype(np.float32)
    t = ((z + np.float32(4.0)) * np.float32(32.0)).astype(np.float32)
    bins = np.clip(t, np.float32(0.0), np.float32(255.0)).astype(np.uint32)
    hist = np.zeros(N_BINS + 1, np.uint32)
    hist[:N_BINS] = np.bincount(bins.ravel(), minlength=N_BINS).astype(np.uint32)
    cdf = np.zeros(N_BINS + 2, np.uint32)
    csum = np.cumsum(hist[:N_BINS], dtype=np.uint64)
    cdf[1:N_BINS] = csum[:N_BINS - 1].astype(np.uint32)   # exclusive scan
    cdf[N_BINS] = np.uint32(csum[-1])                      # the total
    tot = np.float32(max(int(cdf[N_BINS]), 1))
    c = cdf[bins].astype(np.float32)
    h = hist[bins].astype(np.float32)
    r = ((((c + np.float32(0.5) * h).astype(np.float32) / tot).astype(np.float32) * np.float32(2.0))
         .astype(np.float32) - np.float32(1.0)).astype(np.float32)
    return r, hist, cdf, bins


def raw_features(rank: np.ndarray, mode: str = "laplacian"):
    """[r, r^2, |grad r|, f3] and the Sobel magnitude (smi.wgsl:423-462 feats), in float64.

    Sobel with clamped borders; f3: 4-neighbour Laplacian (laplacian), 0 (three), r^3 (cubic).
    """
    r = np.asarray(rank, np.float64)
    p = np.pad(r, 1, mode="edge")
    h, w = r.shape
    at = lambda dx, dy: p[1 + dy:1 + dy + h, 1 + dx:1 + dx + w]
    gx = -at(-1, -1) + at(1, -1) - 2.0 * at(-1, 0) + 2.0 * at(1, 0) - at(-1, 1) + at(1, 1)
    gy = -at(-1, -1) - 2.0 * at(0, -1) - at(1, -1) + at(-1, 1) + 2.0 * at(0, 1) + at(1, 1)
    g = np.sqrt(gx * gx + gy * gy)
    if mode == "laplacian":
        f3 = at(-1, 0) + at(1, 0) + at(0, -1) + at(0, 1) - 4.0 * r
    elif mode == "three":
        f3 = np.zeros_like(r)
    elif mode == "cubic":
        f3 = r * r * r
    else:
        raise ValueError(f"smi_features must be one of {FEATURE_MODES}")
    return np.stack([r, r * r, g, f3]), g


def whiten(feat: np.ndarray, ridge: float = RIDGE):
    """Global ridge whitening f <- L^-1 (f - mu) (smi.wgsl:532-599 feat_moments, chol, whiten).

    G = cov + ridge max(tr/4, 1e-12) I; L its Cholesky with diagonal sqrt(max(s, 1e-12)).
    Returns (whitened planes, L^-1, mu).
    """
    f = feat.reshape(4, -1).astype(np.float64)
    n = max(f.shape[1], 1)
    mu = f.sum(1) / n
    G = (f @ f.T) / n - np.outer(mu, mu)
    tr = np.trace(G)
    G = G + ridge * max(tr * 0.25, 1e-12) * np.eye(4)
    L = np.zeros((4, 4))
    for i in range(4):
        for j in range(i + 1):
            s = G[i, j] - np.dot(L[i, :j], L[j, :j])
            if i == j:
                L[i, i] = math.sqrt(max(s, 1e-12))
            else:
                L[i, j] = s / L[j, j]
    X = np.linalg.inv(L)
    X = np.tril(X)
    out = X @ (f - mu[:, None])
    return out.reshape(feat.shape), X, mu


def prepare(gray: np.ndarray, mode: str = "laplacian", half: bool = False):
    """Smi.prepare (smi.zig:348-400) for SMI features: rank → features → whitening → stack.

    Returns dict with rank (f32), hist, cdf, bins, feat (4, h, w) f64 whitened, mag, stack
    (5, h, w) = [w = 1, f0..f3] (smi.wgsl:609-624 fill + stack_planes).
    """
    g = np.asarray(gray, np.float32)
    if half:
        g = round_half(g)
    r, hist, cdf, bins = rank_transform(g)
    raw, mag = raw_features(r, mode)
    feat, _, _ = whiten(raw)
    stack = np.concatenate([np.ones((1,) + r.shape), feat])
    return dict(rank=r, hist=hist, cdf=cdf, bins=bins, feat=feat, mag=mag, stack=stack, mode=mode)


# ── autocorrelation and the pair's correlation area (calibrated maps) ─────────────────────
def feature_acf(feat: np.ndarray) -> np.ndarray:
    """Normalized autocorrelation at the 17x17 lags |dx|, |dy| <= 8 (smi.wgsl:936-975 feat_acf,
    1574-1580 acf_norm). Returns (289, 4), lag index (dy + 8) 17 + (dx + 8)."""
    _, h, w = feat.shape
    out = np.zeros((289, 4))
    for k in range(289):
        dx, dy = k % 17 - 8, k // 17 - 8
        ys, ye = max(0, -dy), min(h, h - dy)
        xs, xe = max(0, -dx), min(w, w - dx)
        if ye <= ys or xe <= xs:
            continue
        a = feat[:, ys:ye, xs:xe]
        b = feat[:, ys + dy:ye + dy, xs + dx:xe + dx]
        cnt = max((ye - ys) * (xe - xs), 1)
        out[k] = (a * b).sum(axis=(1, 2)) / cnt
    zero = out[144].copy()
    return out / np.maximum(zero, 1e-12)


def corr_area(acf_a: np.ndarray, acf_b: np.ndarray) -> float:
    """Bartlett correlation area (px^2) of a pair, >= 1 (smi.wgsl:1582-1595 corr_area)."""
    s = float((acf_a.sum(1) * acf_b.sum(1)).sum())
    return max(1.0, s / 16.0)


# ── canvas geometry (lie.zig:537-579) ───────────────────────────────────────────────────────
def translate(tx: float, ty: float) -> np.ndarray:
    return np.array([[1.0, 0.0, tx], [0.0, 1.0, ty], [0.0, 0.0, 1.0]])


def reg_frame(H, lw, lh, rw, rh):
    """Canvas of a pose: the fixed image plus the moving image's warped corners, clipped to
    ±2 max(size) (lie.zig:542-574 regFrame). Returns (ox, oy, w, h)."""
    H = np.asarray(H, np.float64).reshape(9)
    corners = [(1.0, 1.0), (float(lw), 1.0), (float(lw), float(lh)), (1.0, float(lh))]
    xs_min, ys_min, xs_max, ys_max = 0.0, 0.0, float(rw), float(rh)
    lim = max(max(float(lw), float(rw)), max(max(float(lh), float(rh)), 8.0)) * 2.0
    for px, py in corners:
        x = H[0] * px + H[1] * py + H[2]
        y = H[3] * px + H[4] * py + H[5]
        z = H[6] * px + H[7] * py + H[8]
        zz = 1e-8 if abs(z) < 1e-8 else z
        xs_min = min(xs_min, x / zz - 1)
        xs_max = max(xs_max, x / zz - 1)
        ys_min = min(ys_min, y / zz - 1)
        ys_max = max(ys_max, y / zz - 1)
    clip = lambda v, l: min(l, max(-l, v))
    ox = min(math.floor(clip(xs_min, lim)), 0)
    oy = min(math.floor(clip(ys_min, lim)), 0)
    x1 = max(math.ceil(clip(xs_max, lim)), float(rw))
    y1 = max(math.ceil(clip(ys_max, lim)), float(rh))
    return int(ox), int(oy), int(max(8, x1 - ox)), int(max(8, y1 - oy))


def canvas_h(H, ox, oy):
    """H in canvas pixels: translate(-ox, -oy) H (lie.zig:577-579 canvasH)."""
    return translate(-float(ox), -float(oy)) @ np.asarray(H, np.float64).reshape(3, 3)


def warp_stack(stack: np.ndarray, Hc: np.ndarray, cw: int, ch: int) -> np.ndarray:
    """Moving stack warped onto a cw x ch canvas by Hc (moving → canvas, 1-based): bilinear with
    clamped taps, 0 outside [0, w-1] x [0, h-1] (smi.wgsl:1659-1690 warp_stack). The inverse is
    taken in float64 here (the reference: f32 adjugate, inverse3_plain)."""
    P, h, w = stack.shape
    out = np.zeros((P, ch, cw))
    if abs(np.linalg.det(Hc)) < 1e-12:
        return out
    Hi = np.linalg.inv(Hc)
    gy, gx = np.mgrid[0:ch, 0:cw].astype(np.float64)
    X = np.stack([gx + 1, gy + 1, np.ones_like(gx)])
    q = np.tensordot(Hi, X, 1)
    okz = np.abs(q[2]) >= 1e-8
    qz = np.where(okz, q[2], 1.0)
    fx = q[0] / qz - 1.0
    fy = q[1] / qz - 1.0
    ok = okz & ~((fx < 0) | (fy < 0) | (fx > w - 1) | (fy > h - 1))
    fxs = np.where(ok, fx, 0.0)
    fys = np.where(ok, fy, 0.0)
    x0 = np.floor(fxs).astype(np.int64)
    y0 = np.floor(fys).astype(np.int64)
    tx = fxs - x0
    ty = fys - y0
    xa, xb = np.clip(x0, 0, w - 1), np.clip(x0 + 1, 0, w - 1)
    ya, yb = np.clip(y0, 0, h - 1), np.clip(y0 + 1, 0, h - 1)
    for p in range(P):
        s = stack[p]
        v = (s[ya, xa] * (1 - tx) + s[ya, xb] * tx) * (1 - ty) + (s[yb, xa] * (1 - tx) + s[yb, xb] * tx) * ty
        out[p] = np.where(ok, v, 0.0)
    return out


def paste_stack(stack: np.ndarray, ox: int, oy: int, cw: int, ch: int) -> np.ndarray:
    """Fixed stack pasted into a cw x ch canvas at (-ox, -oy), 0 elsewhere (smi.wgsl:1692-1706)."""
    P, h, w = stack.shape
    out = np.zeros((P, ch, cw))
    out[:, -oy:-oy + h, -ox:-ox + w] = stack
    return out


# ── grid plans (fft.zig:20-43, 96-109; smi.zig:35-47) ───────────────────────────────────────
def _smooth(n: int, sizes: str) -> bool:
    primes = (2, 3) if sizes == "ladder" else (2, 3, 5, 7, 11, 13)
    for p in primes:
        while n % p == 0:
            n //= p
    return n == 1


def plan_size(target: int, div: int = 1, sizes: str = "ladder") -> int:
    """Smallest length >= target in `sizes`, divisible by div (fft.zig:20-27 planSize)."""
    n = max(target, 1)
    while True:
        if n % div == 0 and _smooth(n, sizes):
            return n
        n += 1


def largest_smooth(max_n: int, sizes: str = "ladder") -> int:
    """Largest length <= max_n (>= 32) in `sizes` (fft.zig:37-43 largestSmooth)."""
    n = max(32, max_n)
    while n >= 32:
        if _smooth(n, sizes):
            return n
        n -= 1
    return 32


def js_round(x: float) -> int:
    return int(math.floor(x + 0.5))


def linear_corr_plan_rect(w: int, h: int, max_n0: int, sizes: str = "ladder"):
    """(nx, ny, cw, ch) for a linear correlation of a w x h canvas with itself
    (fft.zig:96-109 linearCorrPlanRect)."""
    max_n = largest_smooth(max_n0, sizes)
    s = min(1.0, max(8, max_n >> 1) / max(max(w, h), 1))
    for _ in range(16):
        cw = max(8, js_round(w * s))
        ch = max(8, js_round(h * s))
        nx = min(plan_size(max(cw + cw - 1, 32), 1, sizes), max_n)
        ny = min(plan_size(max(ch + ch - 1, 32), 1, sizes), max_n)
        if nx >= cw + cw - 1 and ny >= ch + ch - 1:
            return nx, ny, cw, ch
        s *= 0.85
    c = max(8, (max_n + 1) >> 1)
    return max_n, max_n, c, c


def auto_shift_plan(w: int, h: int):
    """(nx, ny, cw, ch) of the automatic shift map (smi.zig:35-47 autoShiftPlan)."""
    d = max(w, h, 1)
    g = 1
    while 2 * ((d + g - 1) // g) - 1 > MAP_AUTO_MAX:
        g *= 2
    cw = max(1, (w + g - 1) // g)
    ch = max(1, (h + g - 1) // g)
    pow2 = lambda need: min(max(1 << (max(need, 1) - 1).bit_length(), MAP_AUTO_MIN), MAP_AUTO_MAX)
    return pow2(2 * cw - 1), pow2(2 * ch - 1), cw, ch


# ── correlation plan (smi.zig:77-159 CorrPlan.make) ─────────────────────────────────────────
FOURTH_PLANES = (4, 8, 11, 13, 14)


def corr_list(exact: bool = True):
    """The correlation list by number: (real plane of A, real plane of B) (smi.zig:91-116)."""
    lst = [(0, 0)] + [(1 + i, 0) for i in range(4)] + [(0, 1 + j) for j in range(4)]
    lst += [(1 + i, 1 + j) for i in range(4) for j in range(4)]
    if exact:
        lst += [(5 + k, 0) for k in range(10)] + [(0, 5 + k) for k in range(10)]
    return lst


def kept_corrs(exact: bool = True, three: bool = False):
    """Correlation numbers a plan keeps: 45/25 with four features, 28/16 with three."""
    lst = corr_list(exact)
    return [c for c, (a, b) in enumerate(lst) if not (three and (a in FOURTH_PLANES or b in FOURTH_PLANES))]


def real_planes(s: np.ndarray, exact: bool = True) -> np.ndarray:
    """Real correlation planes from premultiplied stack samples s = [w, w f0..w f3]:
    0 → w, 1..4 → w f_i, 5..14 → s_i s_j / s_0 (0 where s_0 <= 1e-6) (smi.wgsl:643-649)."""
    n_planes = 15 if exact else 5
    out = np.zeros((n_planes,) + s.shape[1:])
    out[:5] = s[:5]
    if exact:
        w = s[0]
        ok = w > 1e-6
        ws = np.where(ok, w, 1.0)
        for k, (i, j) in enumerate(TRI):
            out[5 + k] = np.where(ok, s[1 + i] * s[1 + j] / ws, 0.0)
    return out


# ── resample onto the map grid (smi.wgsl:678-705 pack_resample) ──────────────────────────
def pack_resample(canvas: np.ndarray, cw: int, ch: int) -> np.ndarray:
    """Canvas stack (5, H, W) → (5, ch, cw) cells, each the mean of k x k bilinear taps,
    k = clamp(ceil(W / cw - 1e-4), 1, 8) (per axis), sample_plane's pixel-centre convention."""
    P, H, W = canvas.shape
    fx = W / max(cw, 1)
    fy = H / max(ch, 1)
    kx = int(min(max(math.ceil(fx - 1e-4), 1), 8))
    ky = int(min(max(math.ceil(fy - 1e-4), 1), 8))
    if kx == 1 and ky == 1 and cw == W and ch == H:
        return canvas.copy()
    gy, gx = np.mgrid[0:ch, 0:cw].astype(np.float64)
    acc = np.zeros((P, ch, cw))
    for j in range(ky):
        py = (gy + (j + 0.5) / ky) * fy
        for i in range(kx):
            px = (gx + (i + 0.5) / kx) * fx
            x = px - 0.5
            y = py - 0.5
            x0 = np.floor(x).astype(np.int64)
            y0 = np.floor(y).astype(np.int64)
            ax = x - x0
            ay = y - y0
            xa, xb = np.clip(x0, 0, W - 1), np.clip(x0 + 1, 0, W - 1)
            ya, yb = np.clip(y0, 0, H - 1), np.clip(y0 + 1, 0, H - 1)
            for p in range(P):
                s = canvas[p]
                acc[p] += (s[ya, xa] * (1 - ax) + s[ya, xb] * ax) * (1 - ay) + (s[yb, xa] * (1 - ax) + s[yb, xb] * ax) * ay
    return acc / (kx * ky)


# ── correlations ─────────────────────────────────────────────────────────────────────────
def correlations_fft(ra: np.ndarray, rb: np.ndarray, nx: int, ny: int, pairs) -> np.ndarray:
    """Sum_x a(x + t) b(x) for every lag t on an ny x nx grid (lag 0 at [0, 0], negative lags
    wrapped): IFFT(FFT(a) conj(FFT(b))) in float64 (smi.wgsl:730-796 cmul_pair + extract_pair,
    the A conj(B) convention at 751-752). ra, rb: (R, ch, cw) real planes."""
    FA = {}
    FB = {}
    out = np.zeros((len(pairs), ny, nx))
    for k, (a, b) in enumerate(pairs):
        if a not in FA:
            FA[a] = np.fft.rfft2(ra[a], (ny, nx))
        if b not in FB:
            FB[b] = np.fft.rfft2(rb[b], (ny, nx))
        out[k] = np.fft.irfft2(FA[a] * np.conj(FB[b]), (ny, nx))
    return out


def correlations_direct(ra: np.ndarray, rb: np.ndarray, nx: int, ny: int, pairs) -> np.ndarray:
    """Brute-force Sum_x a(x + t) b(x) (tiny images; the FFT path's independent check)."""
    _, h, w = ra.shape
    out = np.zeros((len(pairs), ny, nx))
    for ty in range(-(h - 1), h):
        for tx in range(-(w - 1), w):
            ys, ye = max(0, -ty), min(h, h - ty)
            xs, xe = max(0, -tx), min(w, w - tx)
            for k, (a, b) in enumerate(pairs):
                out[k, ty % ny, tx % nx] = np.sum(ra[a, ys + ty:ye + ty, xs + tx:xe + tx] * rb[b, ys:ye, xs:xe])
    return out


# ── per-lag score (smi.wgsl:799-931 combine_cell) ───────────────────────────────────────────
def f_to_z(f, d1, d2):
    """Wilson–Hilferty normal score of an F(d1, d2) value (smi.wgsl:816-822)."""
    a = 2.0 / (9.0 * d1)
    b = 2.0 / (9.0 * d2)
    c = np.cbrt(np.maximum(f, 0.0))
    return ((1.0 - b) * c - (1.0 - a)) / np.sqrt(b * c * c + a)


def combine(corr: np.ndarray, exact: bool = True, calibrated: bool = False, nf: int = 4,
            min_n: float = MIN_N, ridge: float = RIDGE, area: float = 1.0, zero_n: float | None = None) -> np.ndarray:
    """Per-lag SMI from the correlation planes, float64 (smi.wgsl:825-931 combine_cell).

    corr: (45, ...) (or 25 for global) indexed by correlation number; entries of the fourth
    feature unused when nf == 3. Non-calibrated floor: n < max(min_n, 0.08 n0) → 0, n0 = n at
    lag 0 (or `zero_n`: the sweep's zero-displacement lag, sweep.wgsl:247);
    calibrated: n < min_n or n / area <= 10 → 0.
    """
    corr = np.asarray(corr, np.float64)
    shp = corr.shape[1:]
    c = corr.reshape(corr.shape[0], -1)
    n = c[0]
    calib = exact and calibrated
    neff = n / max(area, 1.0)
    n0 = float(c[0, 0]) if zero_n is None else float(zero_n)
    floor_n = max(float(min_n), 0.08 * max(n0, 0.0))
    if calib:
        floor_n = float(min_n)
    ok = ~(n < floor_n)
    if calib:
        ok &= neff > 10.0
    den = np.maximum(n, 1e-8)
    out = np.zeros(n.shape)
    ma = np.stack([c[1 + i] / den for i in range(nf)])
    mb = np.stack([c[5 + i] / den for i in range(nf)])
    C = np.zeros((nf, nf) + n.shape)
    for i in range(nf):
        for j in range(nf):
            C[i, j] = c[9 + i * 4 + j] / den - ma[i] * mb[j]
    if not exact:
        out = n * np.sum(np.clip(C, -1.0, 1.0) ** 2, axis=(0, 1))
        return np.where(ok, out, 0.0).reshape(shp)
    GA = np.zeros((nf, nf) + n.shape)
    GB = np.zeros((nf, nf) + n.shape)
    for i in range(nf):
        for j in range(nf):
            k = tri_k(i, j)
            GA[i, j] = c[25 + k] / den - ma[i] * ma[j]
            GB[i, j] = c[35 + k] / den - mb[i] * mb[j]
    tra = np.trace(GA)
    trb = np.trace(GB)
    for i in range(nf):
        GA[i, i] += ridge * np.maximum(tra * 0.25, 1e-12)
        GB[i, i] += ridge * np.maximum(trb * 0.25, 1e-12)
    GA = np.moveaxis(GA, (0, 1), (-2, -1))
    GB = np.moveaxis(GB, (0, 1), (-2, -1))
    Cm = np.moveaxis(C, (0, 1), (-2, -1))
    # Cholesky per lag; lags that are not positive definite score 0 (chol4 false)
    pd = np.ones(n.shape, bool)
    LA = np.zeros_like(GA)
    LB = np.zeros_like(GB)
    for G, L in ((GA, LA), (GB, LB)):
        for i in range(nf):
            for j in range(i + 1):
                s = G[..., i, j] - np.sum(L[..., i, :j] * L[..., j, :j], axis=-1)
                if i == j:
                    pd &= s > 0
                    L[..., i, i] = np.sqrt(np.where(s > 0, s, 1.0))
                else:
                    L[..., i, j] = s / L[..., j, j]
    X = np.zeros_like(Cm)
    for j in range(nf):
        for i in range(nf):
            s = Cm[..., i, j] - np.sum(LA[..., i, :i] * X[..., :i, j], axis=-1)
            X[..., i, j] = s / LA[..., i, i]
    acc = np.zeros(n.shape)
    for r in range(nf):
        y = np.zeros(n.shape + (nf,))
        for i in range(nf):
            s = X[..., r, i] - np.sum(LB[..., i, :i] * y[..., :i], axis=-1)
            y[..., i] = s / LB[..., i, i]
            acc += y[..., i] ** 2
    score = n * acc
    if calib:
        q = float(nf)
        v = np.minimum(acc, q - 1e-4)
        with np.errstate(divide="ignore", invalid="ignore"):
            f = (neff - q - 1.0) / q * v / (q - v)
            score = np.maximum(f_to_z(f, q * q, q * (neff - q - 1.0)), 0.0)
    score = np.where(np.isfinite(score), score, 0.0)
    out = np.where(ok & pd, score, 0.0)
    return out.reshape(shp)


def argmax_first(m: np.ndarray) -> int:
    """Flat argmax, first index on ties (smi.wgsl:1600-1644): np.argmax already takes the first."""
    return int(np.argmax(m.ravel()))


def peak_lag(idx: int, nx: int, ny: int):
    """Lag of flat index idx in an nx-wide, ny-high map, negative lags wrapped
    (engine.zig:789-793 peakLag)."""
    py, px = divmod(idx, nx)
    return (px if px <= nx // 2 else px - nx, py if py <= ny // 2 else py - ny)


def shift_map(prep_a: dict, prep_b: dict, H, exact: bool = True, calibrated: bool = False,
              cap: int = MAP_CAP, min_overlap: float = 0.0, sizes: str = "ladder",
              plan: tuple | None = None, direct: bool = False, frame: tuple | None = None):
    """Dense shift map of pose H (smi.zig:540-572 shiftMap/shiftMapFrom; engine.zig:795-819).

    cap 0: automatic plan; else linearCorrPlanRect at largest_smooth(cap). `plan` overrides
    (nx, ny, cw, ch) (full-resolution maps: cw = canvas width). frame: (ox, oy, w, h) instead of
    H's own (smi.zig:548 shiftMapAt: the climb's batch frame). Returns (map, info dict).
    """
    sa, sb = prep_a["stack"], prep_b["stack"]
    _, lh, lw = sa.shape
    _, rh, rw = sb.shape
    ox, oy, W, Hh = reg_frame(H, lw, lh, rw, rh) if frame is None else frame
    Hc = canvas_h(H, ox, oy)
    ca = warp_stack(sa, Hc, W, Hh)
    cb = paste_stack(sb, ox, oy, W, Hh)
    three = prep_a["mode"] == "three"
    if plan is None:
        plan = auto_shift_plan(W, Hh) if cap == 0 else linear_corr_plan_rect(W, Hh, cap, sizes)
    nx, ny, cw, ch = plan
    ga = pack_resample(ca, cw, ch)
    gb = pack_resample(cb, cw, ch)
    pa = real_planes(ga, exact)
    pb = real_planes(gb, exact)
    lst = corr_list(exact)
    keep = kept_corrs(exact, three)
    pairs = [lst[c] for c in keep]
    corr_k = (correlations_direct if direct else correlations_fft)(pa, pb, nx, ny, pairs)
    corr = np.zeros((len(lst), ny, nx))
    corr[keep] = corr_k
    floor_n = math.floor(min_overlap * cw * ch) if min_overlap > 0 else 0
    area = 1.0
    area_px = 1.0
    if exact and calibrated:
        area_px = corr_area(feature_acf(prep_a["feat"]), feature_acf(prep_b["feat"]))
        gcell = W / cw
        area = max(1.0, area_px / (gcell * gcell))
    m = combine(corr, exact=exact, calibrated=calibrated, nf=3 if three else 4,
                min_n=max(MIN_N, floor_n), area=area)
    idx = argmax_first(m)
    lag = peak_lag(idx, nx, ny)
    info = dict(n=nx, ny=ny, cw=cw, ch=ch, canvas=(W, Hh), offset=(ox, oy), peak=float(m.ravel()[idx]),
                peak_index=idx, dx=lag[0] * (W / cw), dy=lag[1] * (Hh / ch), zero=float(m.ravel()[0]),
                corr=corr, floor=max(MIN_N, floor_n), area=area, area_px=area_px,
                thresholds=([float(max(MIN_N, floor_n)), 10.0 * max(area, 1.0)] if (exact and calibrated)
                            else [max(float(max(MIN_N, floor_n)), 0.08 * max(float(corr[0, 0, 0]), 0.0))]))
    return m, info


# ── roto-scale (log-polar) map (smi.zig:577-651; smi.wgsl:1434-1452, 1708-1731, 778-796) ──
MAX_FFT = 4096        # the longest FFT line of the HIP build (maps at one cell per pixel)


def moving_centroid(H, lw: int, lh: int):
    """Mean of the moving image's four corners mapped by H (lie.zig:319-329 movingCentroid)."""
    pts = [map_pt(H, x, y) for x, y in ((1.0, 1.0), (float(lw), 1.0), (float(lw), float(lh)), (1.0, float(lh)))]
    return (sum(p[0] for p in pts) * 0.25, sum(p[1] for p in pts) * 0.25)


def rs_plan(W: int, Hh: int, ox: int, oy: int, c, r0=None, r1=None, cap: int = MAP_CAP,
            sizes: str = "ladder") -> dict:
    """Grid of a roto-scale map on a W x Hh canvas about c (fixed pixels, 1-based)
    (smi.zig:590-620): radii, n = n_theta (FFT width), n_lambda = max(16, n / 2), log spacing."""
    cx = c[0] - 1 - ox
    cy = c[1] - 1 - oy
    if not r1:
        r1 = max(math.hypot(px - cx, py - cy)
                 for px, py in ((0, 0), (W - 1, 0), (W - 1, Hh - 1), (0, Hh - 1)))
    if r0 is None:
        r0 = max(2.0, min(r1 / 16.0, 8.0))
    r_out = max(r1, 2.0)
    r_in = min(max(r0, 1e-3), r_out / 8.0)
    n_ring = max(64, int(math.ceil(2.0 * math.pi * r_out)))
    if cap == 0:
        n = min(max(1 << (n_ring - 1).bit_length(), MAP_AUTO_MIN), MAP_AUTO_MAX)
    else:
        n = min(plan_size(n_ring, 1, sizes), largest_smooth(min(cap, MAX_FFT), sizes))
    n_lam = max(16, n >> 1)
    log0 = math.log(r_in)
    span = math.log(r_out) - log0
    return dict(n=n, n_th=n, n_lam=n_lam, log0=log0, span=span, dlam=span / n_lam, dth=2.0 * math.pi / n,
                r0=r_in, r1=r_out, cx=cx, cy=cy)


def sample_plane_z(plane: np.ndarray, xs: np.ndarray, ys: np.ndarray) -> np.ndarray:
    """Bilinear sample at pixel coordinates (pixel i at i), 0 unless all four taps are inside
    (smi.wgsl:1418-1432 sample_plane_z)."""
    h, w = plane.shape
    x0 = np.floor(xs).astype(np.int64)
    y0 = np.floor(ys).astype(np.int64)
    ok = (x0 >= 0) & (y0 >= 0) & (x0 + 1 < w) & (y0 + 1 < h)
    xa = np.clip(x0, 0, w - 1)
    ya = np.clip(y0, 0, h - 1)
    xb = np.clip(x0 + 1, 0, w - 1)
    yb = np.clip(y0 + 1, 0, h - 1)
    fx = xs - x0
    fy = ys - y0
    v = (plane[ya, xa] * (1 - fx) + plane[ya, xb] * fx) * (1 - fy) + (plane[yb, xa] * (1 - fx) + plane[yb, xb] * fx) * fy
    return np.where(ok, v, 0.0)


def spatial_logpolar(canvas: np.ndarray, p: dict) -> np.ndarray:
    """Canvas stack (5, H, W) on the log-polar grid: rows lambda (radius r_in e^((y+1/2) dlam)),
    columns theta ((x+1/2) 2 pi / n_th) (smi.wgsl:1437-1452 spatial_logpolar)."""
    y, x = np.mgrid[0:p["n_lam"], 0:p["n_th"]].astype(np.float64)
    rho = np.exp(p["log0"] + (y + 0.5) / p["n_lam"] * p["span"])
    th = (x + 0.5) / p["n_th"] * 2.0 * math.pi
    xs = p["cx"] + rho * np.cos(th)
    ys = p["cy"] + rho * np.sin(th)
    return np.stack([sample_plane_z(canvas[k], xs, ys) for k in range(canvas.shape[0])])


def rs_map_dir(prep_a: dict, prep_b: dict, H, c, r0=None, r1=None, exact: bool = True,
               calibrated: bool = False, cap: int = MAP_CAP, direct: bool = False, frame: tuple | None = None):
    """One direction's roto-scale map (smi.zig:589-634 rsMapFrom): canvas stacks at H, both on the
    log-polar grid about c, every real plane x rho (pack_pair's area element), circular in theta and
    linear in log radius on the n x n grid, each correlation x e^(k dlam) dlam dtheta for lag row k
    (extract_pair), then the per-lag score with area scale 1. Returns (map n x n, plan dict)."""
    sa, sb = prep_a["stack"], prep_b["stack"]
    _, lh, lw = sa.shape
    _, rh, rw = sb.shape
    ox, oy, W, Hh = reg_frame(H, lw, lh, rw, rh) if frame is None else frame
    Hc = canvas_h(H, ox, oy)
    ca = warp_stack(sa, Hc, W, Hh)
    cb = paste_stack(sb, ox, oy, W, Hh)
    p = rs_plan(W, Hh, ox, oy, c, r0, r1, cap)
    n, n_lam = p["n"], p["n_lam"]
    om = np.exp(p["log0"] + (np.arange(n_lam) + 0.5) / n_lam * p["span"])[:, None]
    pa = real_planes(spatial_logpolar(ca, p), exact) * om
    pb = real_planes(spatial_logpolar(cb, p), exact) * om
    three = prep_a["mode"] == "three"
    lst = corr_list(exact)
    keep = kept_corrs(exact, three)
    pairs = [lst[k] for k in keep]
    corr_k = (correlations_direct_circ_x if direct else correlations_fft)(pa, pb, n, n, pairs)
    k = np.arange(n)
    k = np.where(k > n // 2, k - n, k)
    scale = np.exp(k * p["dlam"]) * p["dlam"] * p["dth"]
    corr = np.zeros((len(lst), n, n))
    corr[keep] = corr_k * scale[None, :, None]
    area = 1.0
    area_px = 1.0
    if exact and calibrated:
        area_px = corr_area(feature_acf(prep_a["feat"]), feature_acf(prep_b["feat"]))
        area = max(1.0, area_px)
    m = combine(corr, exact=exact, calibrated=calibrated, nf=3 if three else 4, min_n=MIN_N, area=area)
    p.update(canvas=(W, Hh), offset=(ox, oy), corr=corr, area=area)
    return m, p


def correlations_direct_circ_x(ra: np.ndarray, rb: np.ndarray, nx: int, ny: int, pairs) -> np.ndarray:
    """Brute-force sum_x a(x + t) b(x), circular in x over nx columns (ra, rb: (R, h, nx)) and
    linear in y (the log-polar map's correlation; the FFT path's independent check)."""
    _, h, w = ra.shape
    assert w == nx
    out = np.zeros((len(pairs), ny, nx))
    for ty in range(-(h - 1), h):
        ys, ye = max(0, -ty), min(h, h - ty)
        for tx in range(nx):
            for k, (a, b) in enumerate(pairs):
                out[k, ty % ny, tx] = np.sum(np.roll(ra[a], -tx, axis=1)[ys + ty:ye + ty] * rb[b, ys:ye])
    return out


def rs_finish(fwd: np.ndarray, inv, mirror_theta: bool, dlam: float, smin: float = 0.2, smax: float = 5.0):
    """Average with the inverse-frame map at the mirrored lag (-y, -x), or (-y, x) for a
    reflecting pose, then zero the lambda rows whose applied scale exp(-k dlam) lies outside
    [smin, smax] (smi.zig:636-650 rsFinish; smi.wgsl:1708-1731 rs_finish)."""
    n = fwd.shape[0]
    v = fwd.astype(np.float64).copy()
    if inv is not None:
        ym = (n - np.arange(n)) % n
        xm = (n - np.arange(n)) % n if mirror_theta else np.arange(n)
        v = 0.5 * (v + inv[ym][:, xm])
    lo = math.log(max(smin if smin > 0 else 0.2, 1e-6))
    hi = math.log(max(smax if smax > 0 else 5.0, 1.000001))
    k = np.arange(n)
    k = np.where(k > n // 2, k - n, k)
    sg = -k * dlam
    v[(sg < lo) | (sg > hi), :] = 0.0
    return v


def rs_map(prep_a: dict, prep_b: dict, H, c=None, symmetric: bool = True, exact: bool = True,
           calibrated: bool = False, cap: int = MAP_CAP, smin: float = 0.2, smax: float = 5.0):
    """The roto-scale map at H about c (default the moving centroid) (engine.zig:836-879 rsMap):
    forward map, optional inverse-frame map with the forward radii, finish, argmax. Returns
    (finished map, info dict with fwd, inv, peak, dth, dsg and the plan)."""
    H = np.asarray(H, np.float64).reshape(3, 3)
    _, lh, lw = prep_a["stack"].shape
    if c is None:
        c = moving_centroid(H, lw, lh)
    fwd, p = rs_map_dir(prep_a, prep_b, H, c, None, None, exact, calibrated, cap)
    inv = None
    pi = None
    if symmetric:
        Hi = np.linalg.inv(H)
        ci = map_pt(Hi, c[0], c[1])
        inv, pi = rs_map_dir(prep_b, prep_a, Hi, ci, p["r0"], p["r1"], exact, calibrated, cap)
        assert pi["n"] == p["n"], "GridMismatch"
    mirror = H[0, 0] * H[1, 1] - H[0, 1] * H[1, 0] > 0
    m = rs_finish(fwd, inv, mirror, p["dlam"], smin, smax)
    n = p["n"]
    idx = argmax_first(m)
    lag = peak_lag(idx, n, n)
    info = dict(p, fwd=fwd, inv=inv, inv_plan=pi, peak=float(m.ravel()[idx]), peak_index=idx,
                zero=float(m.ravel()[0]), dth=-(lag[0] / max(p["n_th"], 1)) * 2.0 * math.pi, dsg=-lag[1] * p["dlam"],
                center=tuple(c))
    return m, info


# ── pose moments and score (smi.wgsl:980-1057 overlap_moments; moments.zig:91-172) ──────────
def ramp(v, lo, hi):
    """Edge taper (value, slope) over TAPER px inside [lo, hi] (smi.wgsl:338-346)."""
    a, b = (v - lo) / TAPER, (hi - v) / TAPER
    ca, cb = np.clip(a, 0, 1), np.clip(b, 0, 1)
    da = np.where((a > 0) & (a < 1), 1.0 / TAPER, 0.0)
    db = np.where((b > 0) & (b < 1), -1.0 / TAPER, 0.0)
    return ca * cb, da * cb + ca * db


def sample_pose(fa: np.ndarray, hb: int, wb: int, H):
    """Moving features at every fixed pixel under H (bilinear, clamped taps), the taper weight
    and the feature gradients (smi.wgsl:252-275, 989-1027; examples/03_smi.py:62-86)."""
    _, hm, wm = fa.shape
    y, x = np.mgrid[0:hb, 0:wb].astype(np.float64)
    X = np.stack([x + 1, y + 1, np.ones_like(x)])
    Hi = np.linalg.inv(np.asarray(H, np.float64).reshape(3, 3))
    ph = np.tensordot(Hi, X, 1)
    okz = np.abs(ph[2]) >= 1e-8
    pz = np.where(okz, ph[2], 1.0)
    fx, fy = ph[0] / pz - 1, ph[1] / pz - 1
    ok = okz & (fx >= 2) & (fy >= 2) & (fx <= wm - 3) & (fy <= hm - 3)
    tx, dtx = ramp(fx, 2.0, wm - 3.0)
    ty, dty = ramp(fy, 2.0, hm - 3.0)
    fxs, fys = np.where(ok, fx, 2.0), np.where(ok, fy, 2.0)
    x0 = np.floor(fxs).astype(np.int64)
    y0 = np.floor(fys).astype(np.int64)
    ax, ay = fxs - x0, fys - y0
    x0c, x1c = np.clip(x0, 0, wm - 1), np.clip(x0 + 1, 0, wm - 1)
    y0c, y1c = np.clip(y0, 0, hm - 1), np.clip(y0 + 1, 0, hm - 1)
    v00, v10, v01, v11 = fa[:, y0c, x0c], fa[:, y0c, x1c], fa[:, y1c, x0c], fa[:, y1c, x1c]
    va = (v00 * (1 - ax) + v10 * ax) * (1 - ay) + (v01 * (1 - ax) + v11 * ax) * ay
    gx = (v10 - v00) * (1 - ay) + (v11 - v01) * ay
    gy = (v01 - v00) * (1 - ax) + (v11 - v10) * ax
    w = np.where(ok, tx * ty, 0.0)
    dw = np.where(ok, np.stack([dtx * ty, tx * dty]), 0.0)
    return dict(va=va, gx=gx, gy=gy, w=w, dw=dw, ph=ph, X=X)


def moments(va, vb, w):
    """The 45 overlap sums (layout smi.wgsl:12-17; examples/03_smi.py:109-113)."""
    s = [w.sum()] + [(w * va[i]).sum() for i in range(4)] + [(w * vb[j]).sum() for j in range(4)]
    s += [(w * va[i] * vb[j]).sum() for i in range(4) for j in range(4)]
    s += [(w * va[i] * va[j]).sum() for i, j in TRI] + [(w * vb[i] * vb[j]).sum() for i, j in TRI]
    return np.array(s)


def pose_moments(feat_a, feat_b, H):
    """45 moment sums of the moving features onto the fixed at H (zc_moments)."""
    _, hb, wb = feat_b.shape
    S = sample_pose(feat_a, hb, wb, H)
    return moments(S["va"], feat_b, S["w"])


def score_from_moments(st, exact: bool = True, ridge: float = RIDGE) -> float:
    """SMI n tr(G_A^-1 C G_B^-1 C^T) (exact) or n ||C||^2 (global) from 45 sums
    (moments.zig:91-134 scoreFromMoments; examples/03_smi.py:129-147). 0 below MIN_N."""
    st = np.asarray(st, np.float64)
    n = st[0]
    if not (n >= MIN_N):
        return 0.0
    ma, mb = st[1:5] / n, st[5:9] / n
    C = st[9:25].reshape(4, 4) / n - np.outer(ma, mb)
    if not exact:
        return float(n * np.sum(C * C))
    GA = np.zeros((4, 4))
    GB = np.zeros((4, 4))
    for i in range(4):
        for j in range(4):
            k = tri_k(i, j)
            GA[i, j] = st[25 + k] / n - ma[i] * ma[j]
            GB[i, j] = st[35 + k] / n - mb[i] * mb[j]
    GA += ridge * max(np.trace(GA) / 4, 1e-12) * np.eye(4)
    GB += ridge * max(np.trace(GB) / 4, 1e-12) * np.eye(4)
    try:
        np.linalg.cholesky(GA)
        np.linalg.cholesky(GB)
    except np.linalg.LinAlgError:
        return 0.0
    W = np.linalg.solve(GA, C) @ np.linalg.inv(GB)
    s = float(n * np.sum(W * C))
    return s if math.isfinite(s) else 0.0


def pose_score(feat_a, feat_b, H, exact: bool = True, symmetric: bool = False):
    """(mean, fwd, inv) pose score (pair.zig:151-180 parts): inverse direction = fixed onto
    moving by H^-1; mean of the two when symmetric."""
    fwd = score_from_moments(pose_moments(feat_a, feat_b, H), exact)
    if not symmetric:
        return fwd, fwd, None
    inv = score_from_moments(pose_moments(feat_b, feat_a, np.linalg.inv(np.asarray(H, np.float64).reshape(3, 3))), exact)
    return 0.5 * (fwd + inv), fwd, inv


# ── rotation × scale sweep (search.zig:78-143, 300-529; sweep.wgsl:64-371) ──────────────────
def sweep_grid(rot_lo, rot_hi, n_theta, s_min, s_max, n_sigma, shear_max=0.0, n_shear=1, aniso_max=0.0,
               n_aniso=1) -> dict:
    """Candidate axes (search.zig:89-116 sweepGrid): the full circle without repeating ±180°, a
    partial range with both ends; scales and stretches log-spaced, shears linear about 0."""
    def linspace(lo, hi, n):
        return [0.5 * (lo + hi)] if n == 1 else [lo + (i * (hi - lo)) / (n - 1) for i in range(n)]
    nT = max(1, n_theta)
    lo_d, span = min(rot_lo, rot_hi), abs(rot_hi - rot_lo)
    wrap = span >= 360 - 1e-9
    th = [(-180 + 360 * i / nT if wrap else (lo_d + span / 2 if nT == 1 else lo_d + span * i / (nT - 1)))
          * math.pi / 180 for i in range(nT)]
    llo = math.log(max(1e-3, min(s_min, s_max)))
    lhi = math.log(max(1e-3, max(s_min, s_max)))
    sg = [math.exp(v) for v in linspace(llo, lhi, n_sigma)]
    kmax = abs(shear_max)
    sh = linspace(-kmax, kmax, n_shear if kmax > 0 else 1)
    amax = math.log(1 + abs(aniso_max))
    an = [math.exp(v) for v in linspace(-amax, amax, n_aniso if amax > 0 else 1)]
    return dict(thetas=th, sigmas=sg, shears=sh, anisos=an, wrap=wrap)


def cand_linear(th, s, k=0.0, a=1.0):
    """σ R(θ) diag(a, 1/a) [1 k; 0 1] row-major (search.zig:119-127)."""
    c, sn = s * math.cos(th), s * math.sin(th)
    return [c * a, c * a * k - sn / a, sn * a, sn * a * k + c / a]


def about_c(c, L):
    """X ↦ c + L (X − c) (search.zig:130-132)."""
    return np.array([[L[0], L[1], c[0] - L[0] * c[0] - L[1] * c[1]],
                     [L[2], L[3], c[1] - L[2] * c[0] - L[3] * c[1]], [0.0, 0.0, 1.0]])


def sv2(L):
    """Singular values of a 2×2, largest first (search.zig:135-143); checked against np.linalg.svd."""
    a = L[0] ** 2 + L[2] ** 2
    b = L[0] * L[1] + L[2] * L[3]
    d = L[1] ** 2 + L[3] ** 2
    t = 0.5 * (a + d)
    r = math.sqrt(max(0.0, 0.25 * (a - d) ** 2 + b * b))
    return math.sqrt(t + r), math.sqrt(max(0.0, t - r))


def map_pt(H, x, y):
    """lie.mapPt (lie.zig:311-317): |Z| < 1e-8 divides by +1e-8."""
    H = np.asarray(H, np.float64).reshape(3, 3)
    X, Y, Z = H @ np.array([x, y, 1.0])
    z = 1e-8 if abs(Z) < 1e-8 else Z
    return X / z, Y / z


def _cells(a, b, o0, c, g):
    """Canvas cells [lo, hi) over dest-pixel extent [a, b], one cell of margin (search.zig:329-334)."""
    lo = min(max(math.floor((a - o0) / c) - 1, 0), g)
    hi = min(max(math.ceil((b - o0) / c) + 1, lo + 1), g)
    return lo, hi


def sweep_window(feat, Hi, gx0, gy0, win_x, win_y, grid, k, ox, oy, cell):
    """(5, win_y, win_x) stack values [w, w f0..w f3] (w = 1) of a candidate's window: cell
    (gx0 + x, gy0 + y) is the mean of k×k bilinear taps at Hi · tap point, each tap zero unless
    all four corners lie inside the source (sweep.wgsl:64-85 tap, 106-145 pack_sweep); cells
    outside the canvas are zero."""
    _, sh, sw = feat.shape
    Hi = np.asarray(Hi, np.float64).reshape(3, 3)
    cy, cx = np.mgrid[0:win_y, 0:win_x]
    cx = cx + gx0
    cy = cy + gy0
    inside = (cx < grid) & (cy < grid)
    acc = np.zeros((5, win_y, win_x))
    for ty in range(k):
        py = oy + (cy + (ty + 0.5) / k) * cell
        for tx in range(k):
            px = ox + (cx + (tx + 0.5) / k) * cell
            qx = Hi[0, 0] * px + Hi[0, 1] * py + Hi[0, 2]
            qy = Hi[1, 0] * px + Hi[1, 1] * py + Hi[1, 2]
            qz = Hi[2, 0] * px + Hi[2, 1] * py + Hi[2, 2]
            okz = np.abs(qz) >= 1e-8
            qz = np.where(okz, qz, 1.0)
            x = qx / qz - 1.0
            y = qy / qz - 1.0
            ok = inside & okz & (x >= 0) & (y >= 0) & (x < sw - 1) & (y < sh - 1)
            xs, ys = np.where(ok, x, 0.0), np.where(ok, y, 0.0)
            x0, y0 = np.floor(xs).astype(np.int64), np.floor(ys).astype(np.int64)
            fx, fy = xs - x0, ys - y0
            w00, w10, w01, w11 = (1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy
            acc[0] += np.where(ok, w00 + w10 + w01 + w11, 0.0)
            for p in range(4):
                f = feat[p]
                v = w00 * f[y0, x0] + w10 * f[y0, x0 + 1] + w01 * f[y0 + 1, x0] + w11 * f[y0 + 1, x0 + 1]
                acc[1 + p] += np.where(ok, v, 0.0)
    return acc / (k * k)


def sweep_peaks(m: np.ndarray, n_peak: int = 4):
    """Top local maxima of a score map (sweep.wgsl:326-371 peaks_sweep): lag class c = t mod 256
    keeps its largest value above 0 that no 3×3 periodic neighbour exceeds (first lag on ties);
    the n_peak best classes, as (value, flat lag) sorted by value (lower lag on ties); empty
    classes are (0, 0xffffffff)."""
    ny, nx = m.shape
    ok = m > 0
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dx or dy:
                ok &= ~(np.roll(m, (-dy, -dx), axis=(0, 1)) > m)
    flat = m.ravel()
    okf = ok.ravel()
    best = []
    for c in range(256):
        idx = np.arange(c, flat.size, 256)
        idx = idx[okf[idx]]
        if idx.size == 0:
            best.append((0.0, 0xFFFFFFFF))
            continue
        v = flat[idx]
        j = int(np.argmax(v))   # first maximum = lowest lag of this class
        best.append((float(v[j]), int(idx[j])))
    best.sort(key=lambda e: (-e[0], e[1]))
    return best[:n_peak]


def sweep(prep_m: dict, prep_f: dict, H0, gv: dict, grid: int = 128, exact: bool = True, n_peak: int = 4,
          min_overlap: float = 0.05, full: bool = False, keep_maps: bool = False):
    """Every candidate's shift map on the sweep canvas, reduced to its top local maxima
    (search.zig:300-494 Sweep.sweep; sweep.wgsl pack_sweep / cmul_sweep / combine_sweep /
    peaks_sweep). Plane lengths are the powers of two >= max(32, M + F − 1) (the layout the
    peak classes are defined on); the values per lag do not depend on the length.

    Returns dict(peaks=[...], best, maps (if keep_maps: per candidate the map, its n plane and
    overlap floor), cell, grid, D, ox, oy, fc, cands, min_n)."""
    feat_m, feat_f = prep_m["feat"], prep_f["feat"]
    three = prep_m["mode"] == "three"
    _, mh, mw = feat_m.shape
    _, fh, fw = feat_f.shape
    H0 = np.asarray(H0, np.float64).reshape(3, 3)
    cm = map_pt(H0, (mw + 1) / 2, (mh + 1) / 2)
    s_max = max(sv2(cand_linear(0.0, s, k, a))[0] for s in gv["sigmas"] for k in gv["shears"] for a in gv["anisos"])
    D = min(2 * max(fw, fh), max(max(fw, fh), s_max * math.hypot(mw, mh)))
    if full:
        grid = int(math.ceil(D - 1e-9))
    cell = D / grid
    ox, oy = (fw + 1) / 2 - D / 2, (fh + 1) / 2 - D / 2
    f32 = lambda v: float(np.float32(v))
    oxf, oyf, cellf = f32(ox), f32(oy), f32(cell)
    fcx, fcy = _cells(0.5, fw + 0.5, ox, cell, grid), _cells(0.5, fh + 0.5, oy, cell, grid)
    Fx, Fy = fcx[1] - fcx[0], fcy[1] - fcy[0]
    s0r = math.sqrt(abs(H0[0, 0] * H0[1, 1] - H0[0, 1] * H0[1, 0]))
    s0 = 1.0 if (s0r == 0 or math.isnan(s0r)) else s0r
    plen = lambda need: 1 << max(5, (max(need, 1) - 1).bit_length())
    lst = corr_list(exact)
    keep = kept_corrs(exact, three)
    pairs = [lst[c] for c in keep]
    min_n = max(MIN_N, js_round(min_overlap * (fw * fh) / (cell * cell)))
    kf = int(min(6, max(1, math.ceil(cell - 1e-3))))
    fixed_cache = {}
    peaks, best, maps, cands = [], [], [], []
    ci = 0
    for it, th in enumerate(gv["thetas"]):
        for is_, s in enumerate(gv["sigmas"]):
            for ik, kk in enumerate(gv["shears"]):
                for ia, a in enumerate(gv["anisos"]):
                    L = cand_linear(th, s, kk, a)
                    H = about_c(cm, L) @ H0
                    pts = [map_pt(H, qx, qy) for qx, qy in ((0.5, 0.5), (mw + 0.5, 0.5), (mw + 0.5, mh + 0.5),
                                                            (0.5, mh + 0.5))]
                    xs, ys = [p[0] for p in pts], [p[1] for p in pts]
                    if all(math.isfinite(v) for v in xs + ys):
                        mcx = _cells(min(xs), max(xs), ox, cell, grid)
                        mcy = _cells(min(ys), max(ys), oy, cell, grid)
                    else:
                        mcx = mcy = (0, grid)
                    nx, ny = plen(mcx[1] - mcx[0] + Fx - 1), plen(mcy[1] - mcy[0] + Fy - 1)
                    taps = int(min(6, max(1, math.ceil(cell / (sv2(L)[1] * s0) - 1e-3))))
                    if (nx, ny) not in fixed_cache:
                        sf = sweep_window(feat_f, np.eye(3), fcx[0], fcy[0], Fx, Fy, grid, kf, oxf, oyf, cellf)
                        fixed_cache[(nx, ny)] = real_planes(sf, exact)
                    pb = fixed_cache[(nx, ny)]
                    Hi = np.linalg.inv(H).astype(np.float32).astype(np.float64)
                    wx, wy = nx - Fx + 1, ny - Fy + 1
                    sm = sweep_window(feat_m, Hi, mcx[0], mcy[0], wx, wy, grid, taps, oxf, oyf, cellf)
                    pa = real_planes(sm, exact)
                    ck = correlations_fft(pa, pb, nx, ny, pairs)
                    corr = np.zeros((len(lst), ny, nx))
                    corr[keep] = ck
                    zl = ((fcy[0] - mcy[0]) % ny) * nx + ((fcx[0] - mcx[0]) % nx)
                    n0 = corr[0].ravel()[zl]
                    m = combine(corr, exact=exact, nf=3 if three else 4, min_n=min_n, zero_n=n0)
                    if keep_maps:
                        maps.append(dict(map=m, n=corr[0], floor=max(float(min_n), 0.08 * max(n0, 0.0))))
                    cands.append(dict(it=it, is_=is_, ik=ik, ia=ia, th=th, s=s, H=H, gx0=mcx[0], gy0=mcy[0],
                                      nx=nx, ny=ny, taps=taps))
                    b = 0.0
                    for k, (v, t) in enumerate(sweep_peaks(m, n_peak)):
                        v *= cell * cell
                        if not (v > 0) or t >= nx * ny:
                            continue
                        if k == 0:
                            b = v
                        py, px = divmod(t, nx)
                        lx = px if px < wx else px - nx
                        ly = py if py < wy else py - ny
                        sx, sy = mcx[0] + lx - fcx[0], mcy[0] + ly - fcy[0]
                        Hc = translate(-sx * cell, -sy * cell) @ H
                        peaks.append(dict(score=v, cand=ci, it=it, is_=is_, ik=ik, ia=ia, th=th, s=s, sx=sx, sy=sy,
                                          H=Hc, lag=t))
                    best.append(b)
                    ci += 1
    return dict(peaks=peaks, best=np.array(best), maps=maps, cell=cell, grid=grid, D=D, ox=ox, oy=oy,
                fc=(fcx, fcy), cands=cands, min_n=min_n)


def nms4(peaks, n_theta: int, keep: int, wrap: bool):
    """Non-maximum suppression over (θ, σ, shear, stretch, shift) (search.zig:500-529): strongest
    first (stable), dropped within ±2 θ steps (circular when wrap), ±2 σ, ±1 shear and stretch
    steps and 4 cells of a kept peak."""
    kept = []
    for p in sorted(peaks, key=lambda q: -q["score"]):
        ok = True
        for q in kept:
            d = {k: abs(int(p[k]) - int(q[k])) for k in ("it", "is_", "ik", "ia", "sx", "sy")}
            dt = min(d["it"], n_theta - d["it"]) if wrap else d["it"]
            if (dt <= 2 and d["is_"] <= 2 and d["ik"] <= 1 and d["ia"] <= 1
                    and math.hypot(d["sx"], d["sy"]) <= 4):
                ok = False
                break
        if ok:
            kept.append(p)
            if len(kept) >= keep:
                break
    return kept
