"""Pins the float64 oracle (CPU only) against independent statements of the same mathematics,
the reference's own known answers, and fixtures the reference holds.

What pins what:
* exact image statistics — Python big-integer / Fraction arithmetic on small inputs, and agreement
  with float64 mean/std;
* rank transform — an independent mid-rank formula from np.histogram / np.searchsorted;
* features — scipy.ndimage.correlate with the Sobel / Laplacian stencils (mode "nearest");
* whitening — the whitened covariance is (Σ + ridge I)^-1/2 Σ (Σ + ridge I)^-1/2 (closed form);
* correlations — brute-force direct sums; FFT-length independence;
* combine — the 03_smi.py smi() formulation (np.linalg.solve / inv) on random moment vectors;
  special cases: affine-related sides score n · nf (ridge-limited), independent noise scores
  about nf^2 (Pillai's E[V] = nf^2 / n), constant images score 0;
* shift map — the integer shift of the synthetic workload is found exactly; the reference's
  landsat-a preset (pure translation (80, 15)) and the 03_smi.py landsat-b known answer
  (peak exactly (14, -9), examples/03_smi.py:364) are reproduced;
* grid plans — the values the reference's comments state (smi.zig:21-47, fft.zig:96-109);
* lie inv3 — the reference's unit test matrix (lie.zig:581-586);
* pose moments — brute-force per-pixel sums on a pure translation, where bilinear sampling is
  exact.
"""
import json
import math
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest
from scipy import ndimage

from oracle import smi_ref as O
from zcmir_amd import synth

GOLD = Path(__file__).resolve().parent / "golden"


# ── exact statistics ─────────────────────────────────────────────────────────────────────
@pytest.mark.parametrize("seed", range(4))
def test_image_stats_exact_integer_sums(seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.3, 1.7, (7, 11)).astype(np.float32)
    x[0, 0] = 0.0
    mean, spread = O.image_stats(x)
    e = math.frexp(float(np.abs(x).max()))[1]
    K = 62 - e
    S1 = sum(int(Fraction(float(v)) * 2 ** K) for v in x.ravel())          # trunc toward zero
    hi, lo = S1 >> 64, S1 & ((1 << 64) - 1)
    m64 = math.ldexp((float(hi) * 2.0 ** 64 + float(lo)) / x.size, -K)
    assert np.float32(m64) == mean
    d = (x - mean).astype(np.float32)
    K2 = 60 - 2 * e
    S2 = sum(int(Fraction(float(v)) ** 2 * 2 ** K2) for v in d.ravel())
    hi, lo = S2 >> 64, S2 & ((1 << 64) - 1)
    v64 = math.ldexp((float(hi) * 2.0 ** 64 + float(lo)) / x.size, -K2)
    assert np.float32(math.sqrt(max(v64, 1e-12))) == spread
    assert abs(float(mean) - x.astype(np.float64).mean()) < 1e-6
    assert abs(float(spread) - x.astype(np.float64).std()) < 1e-6


def test_rank_matches_midrank_formula():
    rng = np.random.default_rng(5)
    g = rng.uniform(0, 1, (90, 70)).astype(np.float32)
    g[:10] = np.round(g[:10] * 20) / 20
    r, hist, cdf, bins = O.rank_transform(g)
    h, _ = np.histogram(bins, bins=np.arange(257))
    assert np.array_equal(hist[:256], h)
    assert cdf[256] == g.size
    below = np.searchsorted(np.sort(bins.ravel()), bins, side="left")
    same = h[bins]
    ref = ((below + 0.5 * same) / g.size) * 2 - 1
    assert np.abs(r - ref).max() < 1e-6
    # monotone in the image value
    o = np.argsort(g.ravel(), kind="stable")
    assert np.all(np.diff(r.ravel()[o]) >= 0)


def test_features_match_ndimage():
    rng = np.random.default_rng(6)
    r = rng.uniform(-1, 1, (33, 41))
    feat, g = O.raw_features(r, "laplacian")
    sx = ndimage.correlate(r, np.array([[-1, 0, 1], [-2, 0, 2], [-1, 0, 1.0]]), mode="nearest")
    sy = ndimage.correlate(r, np.array([[-1, -2, -1], [0, 0, 0], [1, 2, 1.0]]), mode="nearest")
    lap = ndimage.correlate(r, np.array([[0, 1, 0], [1, -4, 1], [0, 1, 0.0]]), mode="nearest")
    assert np.allclose(feat[2], np.hypot(sx, sy))
    assert np.allclose(feat[3], lap)
    assert np.allclose(feat[1], r * r)
    assert np.all(O.raw_features(r, "three")[0][3] == 0)
    assert np.allclose(O.raw_features(r, "cubic")[0][3], r ** 3)


def test_whitening_closed_form():
    rng = np.random.default_rng(7)
    f = rng.standard_normal((4, 50, 60)) @ np.eye(60) + 0.5
    f[1] += 0.7 * f[0]
    w, X, mu = O.whiten(f)
    flat = f.reshape(4, -1)
    S = np.cov(flat, bias=True)
    G = S + 1e-4 * np.trace(S) / 4 * np.eye(4)
    L = np.linalg.cholesky(G)
    assert np.allclose(X, np.linalg.inv(L), atol=1e-12)
    cw = np.cov(w.reshape(4, -1), bias=True)
    expect = np.linalg.inv(L) @ S @ np.linalg.inv(L).T
    assert np.allclose(cw, expect, atol=1e-10)
    assert np.abs(cw - np.eye(4)).max() < 1e-3


# ── correlations and the combine ─────────────────────────────────────────────────────────
def test_correlations_fft_equal_direct():
    rng = np.random.default_rng(8)
    ra = rng.standard_normal((3, 9, 13))
    rb = rng.standard_normal((3, 9, 13))
    pairs = [(0, 0), (1, 2), (2, 1)]
    a = O.correlations_fft(ra, rb, 32, 32, pairs)
    b = O.correlations_direct(ra, rb, 32, 32, pairs)
    assert np.abs(a - b).max() < 1e-10
    c = O.correlations_fft(ra, rb, 27, 18, pairs)       # any length >= 2c - 1 gives the same lags
    assert np.allclose(c[:, 0, 0], a[:, 0, 0]) and np.allclose(c[:, -1, -1], a[:, -1, -1])


def _smi_0303(s, ridge=1e-4):
    """examples/03_smi.py:129-147 smi() without the overlap floor (its own formulation)."""
    s = np.asarray(s, np.float64)
    n = s[0]
    ma, mb = s[1:5] / n, s[5:9] / n
    C = s[9:25].reshape(4, 4) / n - np.outer(ma, mb)
    Ga, Gb = np.zeros((4, 4)), np.zeros((4, 4))
    for k, (i, j) in enumerate(O.TRI):
        Ga[i, j] = Ga[j, i] = s[25 + k] / n - ma[i] * ma[j]
        Gb[i, j] = Gb[j, i] = s[35 + k] / n - mb[i] * mb[j]
    for G in (Ga, Gb):
        G += ridge * np.trace(G) / 4 * np.eye(4)
    W = np.linalg.solve(Ga, C) @ np.linalg.inv(Gb)
    return n * np.sum(W * C)


def _random_moments(rng, n=500):
    a = rng.standard_normal((4, n))
    b = 0.4 * a[[2, 0, 3, 1]] + rng.standard_normal((4, n))
    w = rng.uniform(0.5, 1.0, n)
    return O.moments(a, b, w)


@pytest.mark.parametrize("seed", range(3))
def test_combine_matches_0303_formula(seed):
    rng = np.random.default_rng(seed)
    st = np.stack([_random_moments(rng) for _ in range(6)], axis=1)        # (45, 6) "lags"
    got = O.combine(st.reshape(45, 2, 3), exact=True, min_n=0)
    ref = np.array([_smi_0303(st[:, k]) for k in range(6)]).reshape(2, 3)
    assert np.allclose(got, ref, rtol=1e-10)
    assert np.allclose(O.score_from_moments(st[:, 0]), ref.ravel()[0], rtol=1e-10)


def test_combine_special_cases():
    rng = np.random.default_rng(3)
    n = 4000
    a = rng.standard_normal((4, n))
    A = rng.standard_normal((4, 4)) + 3 * np.eye(4)
    b = A @ a + 2.0                                  # an invertible affine map of a
    w = np.ones(n)
    s = O.combine(O.moments(a, b, w)[:, None], exact=True, min_n=0)[0]
    assert abs(s / (n * 4) - 1) < 1e-3              # n nf, short of it only by the ridge
    # independent sides: Pillai E[V] = nf^2 / n, so E[S] = nf^2 = 16
    vals = [O.combine(O.moments(rng.standard_normal((4, n)), rng.standard_normal((4, n)), w)[:, None],
                      exact=True, min_n=0)[0] for _ in range(40)]
    assert 12 < np.mean(vals) < 20
    # a constant image: no covariance, score 0
    c = np.ones((4, n))
    assert O.combine(O.moments(c, b, w)[:, None], exact=True, min_n=0)[0] == 0.0
    assert O.score_from_moments(O.moments(a, b, w)) > 0


def test_f_to_z_matches_scipy():
    from scipy import stats
    for f, d1, d2 in ((1.0, 16, 1000), (2.5, 16, 400), (0.5, 9, 90)):
        z_wh = O.f_to_z(np.array(f), d1, d2)
        z_ex = stats.norm.ppf(stats.f.cdf(f, d1, d2))
        assert abs(z_wh - z_ex) < 0.05


# ── grid plans and geometry ──────────────────────────────────────────────────────────────
def test_grid_plans():
    # smi.zig:21-22: the map cap 1024 on a 1024² canvas: FFT 1024², a 512² cell grid
    assert O.linear_corr_plan_rect(1024, 1024, 1024) == (1024, 1024, 512, 512)
    # full resolution: 1 cell = 1 px
    assert O.linear_corr_plan_rect(1024, 1024, 2048) == (2048, 2048, 1024, 1024)
    assert O.linear_corr_plan_rect(256, 256, 1024) == (512, 512, 256, 256)
    nx, ny, cw, ch = O.linear_corr_plan_rect(640, 400, 1024)
    assert nx >= 2 * cw - 1 and ny >= 2 * ch - 1 and O._smooth(nx, "ladder")
    # autoShiftPlan: FFT 128-512 per side, power of two, one grid factor
    assert O.auto_shift_plan(1024, 1024) == (512, 512, 256, 256)
    assert O.auto_shift_plan(100, 40) == (256, 128, 100, 40)
    assert O.plan_size(2047) == 2048 and O.plan_size(1399) == 1458
    assert O.largest_smooth(1000) == 972


def test_reg_frame():
    assert O.reg_frame(np.eye(3), 100, 80, 120, 90) == (0, 0, 120, 90)
    ox, oy, w, h = O.reg_frame(O.translate(-10, 5), 100, 80, 120, 90)
    assert (ox, oy) == (-10, 0) and w == 130 and h == 90


def test_inv3_reference_matrix():
    # lie.zig:581-586 "inv3 · H = I"
    H = np.array([1.01, 0.012, -3.06, -0.006, 1.008, 4.86, 1e-5, 2e-5, 1]).reshape(3, 3)
    assert np.allclose(np.linalg.inv(H) @ H, np.eye(3), atol=1e-12)


# ── shift maps: known answers ────────────────────────────────────────────────────────────
def test_shift_map_finds_integer_shift():
    mv, fx, H = synth.shift_pair((128, 128), (11, -6), seed=3)
    pa, pb = O.prepare(mv), O.prepare(fx)
    m, info = O.shift_map(pa, pb, np.eye(3), cap=256)
    assert (info["dx"], info["dy"]) == (11.0, -6.0)
    m2, info2 = O.shift_map(pa, pb, np.eye(3), cap=256, min_overlap=0.25, exact=False)
    assert (info2["dx"], info2["dy"]) == (11.0, -6.0)
    # the map about the true pose peaks at lag 0
    m3, info3 = O.shift_map(pa, pb, H, cap=256)
    assert (info3["dx"], info3["dy"]) == (0.0, 0.0)


def test_shift_map_direct_equals_fft_small():
    mv, fx, _ = synth.shift_pair((14, 12), (2, -1), seed=1)
    pa, pb = O.prepare(mv), O.prepare(fx)
    a, ia = O.shift_map(pa, pb, np.eye(3), cap=64)
    b, ib = O.shift_map(pa, pb, np.eye(3), cap=64, direct=True)
    n = ib["corr"][0]
    clear = np.abs(n - ib["thresholds"][0]) > 1e-3
    assert np.abs(a - b)[clear].max() < 1e-8 * np.abs(b).max()


def _preset(name):
    from PIL import Image
    d = GOLD / "presets"
    m = np.asarray(Image.open(d / f"{name}_moving.jpg").convert("L")).astype(np.float32) / 255
    f = np.asarray(Image.open(d / f"{name}_fixed.jpg").convert("L")).astype(np.float32) / 255
    truth = np.array(json.loads((d / "truth.json").read_text())[name])
    return m, f, truth


def test_preset_landsat_a_translation():
    m, f, truth = _preset("landsat-a")
    pa, pb = O.prepare(m), O.prepare(f)
    _, info = O.shift_map(pa, pb, np.eye(3), cap=1024)
    T = np.array([[1, 0, -info["dx"]], [0, 1, -info["dy"]], [0, 0, 1.0]])
    assert np.allclose(T, truth)


def test_preset_landsat_b_0303_known_answer():
    """examples/03_smi.py:150-169, 364: about truth moved by (14, -9), the peak is (14, -9)."""
    m, f, truth = _preset("landsat-b")
    pa, pb = O.prepare(m), O.prepare(f)
    H0 = O.translate(14, -9) @ truth
    _, info = O.shift_map(pa, pb, H0, cap=2048)
    assert (info["dx"], info["dy"]) == (14.0, -9.0)


# ── pose moments ─────────────────────────────────────────────────────────────────────────
def test_pose_moments_brute_force_translation():
    rng = np.random.default_rng(9)
    fa = rng.standard_normal((4, 30, 34))
    fb = rng.standard_normal((4, 28, 31))
    H = O.translate(-3.0, 2.0)          # fixed x = moving x - 3 → sample moving at x + 3
    st = O.pose_moments(fa, fb, H)
    s = np.zeros(45)
    for y in range(28):
        for x in range(31):
            mx, my = x + 3, y - 2                       # 0-based moving pixel (exact, no interpolation)
            if not (2 <= mx <= 34 - 3 and 2 <= my <= 30 - 3):
                continue
            wt = min(max((mx - 2) / 2, 0), 1) * min(max((34 - 3 - mx) / 2, 0), 1)
            wt *= min(max((my - 2) / 2, 0), 1) * min(max((30 - 3 - my) / 2, 0), 1)
            va, vb = fa[:, my, mx], fb[:, y, x]
            s += O.moments(va[:, None], vb[:, None], np.array([wt]))
    assert np.allclose(st, s, rtol=1e-12, atol=1e-9)


# ── pack_resample (smi.wgsl:626-700): closed forms of the k×k bilinear box filter ───────────────
def test_pack_resample_integer_factors_are_block_means():
    """At an integer factor f (k = f taps per axis) every tap lands on a pixel centre
    (x = f·gx + i exactly), so each cell is the plain f×f block mean."""
    rng = np.random.default_rng(5)
    for f in (2, 3, 4):
        canvas = rng.uniform(-1, 1, (5, 6 * f, 7 * f))
        got = O.pack_resample(canvas, 7, 6)
        want = canvas.reshape(5, 6, f, 7, f).mean(axis=(2, 4))
        assert np.abs(got - want).max() < 1e-12, f


def test_pack_resample_reproduces_linear_ramps():
    """Bilinear taps reproduce a linear function, and the taps of a cell are symmetric about its
    centre, so a ramp a·x + b·y + c comes back at x = (gx + ½)·fx − ½ (and y alike) wherever no
    tap is clamped at the border — for non-integer factors and for k clamped at 8."""
    for (H, W, ch, cw) in ((90, 100, 37, 41), (64, 100, 5, 5), (33, 47, 33, 20)):
        y, x = np.mgrid[0:H, 0:W].astype(np.float64)
        a, b, c = 0.37, -1.3, 2.0
        canvas = np.stack([a * x + b * y + c] * 5)
        got = O.pack_resample(canvas, cw, ch)
        fx, fy = W / cw, H / ch
        kx, ky = min(max(math.ceil(fx - 1e-4), 1), 8), min(max(math.ceil(fy - 1e-4), 1), 8)
        gy, gx = np.mgrid[0:ch, 0:cw].astype(np.float64)
        want = a * ((gx + 0.5) * fx - 0.5) + b * ((gy + 0.5) * fy - 0.5) + c
        inside = (((gx + 0.5 / kx) * fx - 0.5 >= 0) & ((gx + 1 - 0.5 / kx) * fx - 0.5 <= W - 1)
                  & ((gy + 0.5 / ky) * fy - 0.5 >= 0) & ((gy + 1 - 0.5 / ky) * fy - 0.5 <= H - 1))
        assert inside.sum() > 0.5 * inside.size
        assert np.abs(got[:, inside] - want[inside]).max() < 1e-9, (H, W, ch, cw)


def test_pack_resample_constant_and_identity():
    rng = np.random.default_rng(6)
    canvas = np.full((5, 50, 70), 0.625)
    assert np.abs(O.pack_resample(canvas, 23, 17) - 0.625).max() < 1e-12
    canvas = rng.uniform(0, 1, (5, 20, 30))
    assert np.array_equal(O.pack_resample(canvas, 30, 20), canvas)


# ── the calibrated map's autocorrelation and correlation area (smi.wgsl:936-975, 1574-1595) ──
def _acf_index(dx, dy):
    return (dy + 8) * 17 + (dx + 8)


def test_feature_acf_of_two_impulses_is_exact():
    """A plane with two unit impulses d = (3, 2) apart has Σ a(x) a(x + t) = 2 at t = 0, 1 at
    t = ±d and 0 at every other lag; divided by the overlap count (w − |dx|)(h − |dy|) and then
    by the zero lag, lag ±d is w·h / (2 (w − 3)(h − 2)). A transposed lag index, a count of the
    whole image or a lag read the wrong way round all fail this."""
    h, w = 23, 31
    feat = np.zeros((4, h, w))
    for p in range(4):
        y0, x0 = 4 + p, 6 + 2 * p
        feat[p, y0, x0] = 1.0
        feat[p, y0 + 2, x0 + 3] = 1.0
    acf = O.feature_acf(feat)
    want = np.zeros(289)
    want[_acf_index(0, 0)] = 1.0
    want[_acf_index(3, 2)] = want[_acf_index(-3, -2)] = w * h / (2.0 * (w - 3) * (h - 2))
    for p in range(4):
        assert np.abs(acf[:, p] - want).max() < 1e-12, p


def test_feature_acf_of_constants_is_one_everywhere():
    """A constant plane's lag products average to the constant squared over any overlap, so
    every normalized lag is 1 (a sum divided by the whole image instead of the overlap would
    fall to (w − 8)(h − 8)/(w h) at the corner lags)."""
    feat = np.stack([np.full((12, 14), v) for v in (1.0, -2.5, 0.25, 3.0)])
    assert np.abs(O.feature_acf(feat) - 1.0).max() < 1e-12


def test_feature_acf_white_noise_and_box_blur():
    """Statistics of known processes: white noise has ρ(t) = δ(t); a k×k box average of white
    noise has the separable triangle ρ(dx, dy) = (1 − |dx|/k)(1 − |dy|/k) for |dx|, |dy| < k and
    0 beyond (sampling error ~ 1/√(pixels/k²))."""
    rng = np.random.default_rng(17)
    noise = rng.standard_normal((4, 384, 384))
    acf = O.feature_acf(noise)
    off = np.delete(acf, _acf_index(0, 0), axis=0)
    assert np.abs(acf[_acf_index(0, 0)] - 1).max() < 1e-12 and np.abs(off).max() < 0.02
    k = 5
    big = rng.standard_normal((4, 384 + k - 1, 384 + k - 1))
    box = np.stack([np.lib.stride_tricks.sliding_window_view(b, (k, k)).mean(axis=(2, 3)) for b in big])
    acf = O.feature_acf(box)
    tri = lambda d: max(0.0, 1.0 - abs(d) / k)   # noqa: E731
    want = np.array([tri(kk % 17 - 8) * tri(kk // 17 - 8) for kk in range(289)])
    assert np.abs(acf - want[:, None]).max() < 0.05


def test_corr_area_closed_forms():
    """Bartlett area (1/16) Σ_t (Σ_i ρ_a,i(t)) (Σ_j ρ_b,j(t)), at least 1: white noise on both
    sides (ρ = δ) gives exactly 1; two k×k-box-blurred sides (triangle ρ on all four planes) give
    (Σ_{|d|<k} (1 − |d|/k)²)² — 3.4² = 11.56 for k = 5; one white side and one blurred side give
    ρ_b(0) = 1, i.e. 1; and the area is symmetric in its sides."""
    delta = np.zeros((289, 4))
    delta[_acf_index(0, 0)] = 1.0
    assert O.corr_area(delta, delta) == 1.0
    k = 5
    tri = lambda d: max(0.0, 1.0 - abs(d) / k)   # noqa: E731
    t = np.array([tri(kk % 17 - 8) * tri(kk // 17 - 8) for kk in range(289)])
    T = np.repeat(t[:, None], 4, axis=1)
    line = sum(tri(d) ** 2 for d in range(-k + 1, k))
    assert abs(line - 3.4) < 1e-12
    assert abs(O.corr_area(T, T) - line ** 2) < 1e-9
    assert O.corr_area(delta, T) == 1.0
    rng = np.random.default_rng(3)
    A, B = rng.uniform(0, 1, (289, 4)), rng.uniform(0, 1, (289, 4))
    assert abs(O.corr_area(A, B) - O.corr_area(B, A)) < 1e-12
    # measured end to end: box-blurred noise through feature_acf → corr_area near the closed form
    big = rng.standard_normal((4, 384 + k - 1, 384 + k - 1))
    box = np.stack([np.lib.stride_tricks.sliding_window_view(b, (k, k)).mean(axis=(2, 3)) for b in big])
    acf = O.feature_acf(box)
    assert abs(O.corr_area(acf, acf) / line ** 2 - 1) < 0.1
