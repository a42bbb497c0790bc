"""Pins of the preprocessing oracle (oracle/pre_ref.py), CPU only.

* CLAHE histograms: np.histogram of each tile of the reflect-padded image;
* CLAHE CDFs: with a clip limit above every count, the plain normalized cumulative histogram;
  a constant image's clipped histogram by hand (the excess spread evenly, the remainder to the
  first bins);
* CLAHE output: at a pixel centre on a tile node with a value on a bin edge, that tile's share of
  values up to the bin, counted directly; monotone in the input;
* the Gaussian pass: scipy.ndimage.convolve1d with normalized exp(-k^2/2) weights, mode nearest;
* band-pass counts (1 and 36 passes by default), its [0, 1] range, invert, RGBA weights.
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle import pre_ref as R


def test_clahe_hist_matches_np_histogram_per_tile():
    rng = np.random.default_rng(0)
    img = rng.uniform(0, 1, (37, 50)).astype(np.float32)
    p = R.clahe_params(50, 37, 64, 4)
    assert (p["h_pad"], p["w_pad"], p["h_tile"], p["w_tile"]) == (3, 2, 10, 13)
    hist = R.clahe_hist(img, p)
    # the padded image: reflect the first rows / columns in front of the image
    ys = R.clahe_reflect(np.arange(p["hp"]), 37, 3)
    xs = R.clahe_reflect(np.arange(p["wp"]), 50, 2)
    assert list(ys[:4]) == [3, 2, 1, 0] and list(xs[:3]) == [2, 1, 0]
    pad = img[np.ix_(ys, xs)]
    for ty in range(4):
        for tx in range(4):
            tile = pad[ty * 10:(ty + 1) * 10, tx * 13:(tx + 1) * 13]
            b = np.floor(tile.astype(np.float64) * 63).astype(int)   # truncation of v * (nb - 1)
            ref = np.bincount(b.ravel(), minlength=64)
            assert np.array_equal(hist[ty * 4 + tx], ref)


def test_clahe_cdf_unclipped_is_the_cumulative_histogram():
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 1, (64, 64)).astype(np.float32)
    p = R.clahe_params(64, 64, 32, 4)
    hist = R.clahe_hist(img, p)
    cdf = R.clahe_cdf(hist, p, clip_limit=1e6)
    ref = np.cumsum(hist, axis=1) / (16 * 16)
    assert np.abs(cdf - ref).max() < 1e-6


def test_clahe_cdf_constant_image_by_hand():
    img = np.full((32, 32), 0.5, np.float32)
    p = R.clahe_params(32, 32, 8, 2)          # 2 x 2 tiles of 16 x 16 = 256 px, 8 bins
    hist = R.clahe_hist(img, p)
    assert hist[:, 3].tolist() == [256] * 4 and hist.sum() == 1024     # 0.5 * 7 = 3.5 → bin 3
    cdf = R.clahe_cdf(hist, p, clip_limit=2.0)
    # limit = 2 * 256 / 8 = 64: bin 3 keeps 64, 192 clipped, 24 per bin, remainder 0
    ref = np.cumsum([24, 24, 24, 88, 24, 24, 24, 24]) / 256.0
    assert np.allclose(cdf, ref[None, :], atol=1e-7)


def test_clahe_output_at_a_tile_node_and_monotone():
    """9 x 9 image, 3 x 3 tiles of 3 x 3, 9 bins, no clipping: at pixel (4, 4) the tile coordinate
    (4 + 1/2) / 9 * 2 is exactly 1 and a value 3/8 sits exactly on bin 3, so the output is tile 4's
    cumulative share of values in bins 0..3."""
    rng = np.random.default_rng(2)
    img = rng.uniform(0, 1, (9, 9)).astype(np.float32)
    img[4, 4] = np.float32(0.375)
    p = R.clahe_params(9, 9, 9, 3)
    assert (p["hg"], p["nb"], p["h_tile"], p["w_tile"], p["h_pad"]) == (3, 9, 3, 3, 0)
    cdf = R.clahe_cdf(R.clahe_hist(img, p), p, clip_limit=1e6)
    out = R.clahe_apply(img, cdf, p)
    tile = img[3:6, 3:6].astype(np.float64)
    assert out[4, 4] == np.float32(np.sum(np.floor(tile * 8) <= 3) / 9.0)
    # monotone in the input at every pixel, within [0, 1]
    cdf = R.clahe_cdf(R.clahe_hist(img, p), p, 40.0)
    out = R.clahe_apply(img, cdf, p)
    out2 = R.clahe_apply(np.clip(img + np.float32(0.05), 0, 1).astype(np.float32), cdf, p)
    assert np.all(out2 >= out - 1e-7) and out.min() >= 0 and out.max() <= 1


def test_gauss_pass_matches_convolve1d():
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 1, (40, 33)).astype(np.float32)
    wts = np.exp(-0.5 * np.arange(-3, 4) ** 2)
    wts = wts / wts.sum()
    for axis in (0, 1):
        ref = ndimage.convolve1d(img.astype(np.float64), wts, axis=axis, mode="nearest")
        got = R.gauss_pass(img, axis, False)
        assert np.abs(got - ref).max() < 1e-6


def test_band_and_invert():
    assert R.band_counts() == (1, 36) and R.band_counts(2.0, 3.0) == (4, 9) and R.band_counts(0.5, 1.0) == (1, 2)
    rng = np.random.default_rng(4)
    img = rng.uniform(0, 1, (48, 40)).astype(np.float32)
    b = R.bandpass(img, 1.0, 2.0, False)
    assert b.min() == 0.0 and abs(float(b.max()) - 1.0) < 1e-6
    assert np.array_equal(R.invert(np.array([[-0.5, 0.25, 2.0]], np.float32)), np.array([[1.0, 0.75, 0.0]], np.float32))


def test_rgba_to_gray_weights():
    px = np.array([[[255, 0, 0, 255], [0, 255, 0, 255], [0, 0, 255, 0], [255, 255, 255, 255]]], np.uint8)
    g = R.rgba_to_gray(px)
    assert g.dtype == np.float32
    assert np.allclose(g[0], [0.299, 0.587, 0.114, 1.0], atol=1e-7)


@pytest.mark.parametrize("half", [False, True])
def test_work_image_order(half):
    rng = np.random.default_rng(5)
    img = rng.uniform(0, 1, (40, 40)).astype(np.float32)
    w0 = R.work_image(img, preprocess=False, half=half)
    assert np.array_equal(w0, R.store(img, half))
    w1 = R.work_image(img, preprocess=True, half=half, invert_on=True)
    assert np.array_equal(w1, R.invert(R.clahe(img, 40.0, 64, 4, half), half))
