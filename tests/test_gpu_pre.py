"""GPU parity of the work-image preprocessing (CLAHE, band pass, invert, RGBA to gray, f16 texels)
against the float32 oracle (oracle/pre_ref.py), through the C ABI: bit-exact work images, then the
rank transform of the preprocessed image bit-exact as well (the features follow from it).
"""
import numpy as np
import pytest

from oracle import pre_ref as R
from oracle import smi_ref as O

pytestmark = pytest.mark.gpu


def _img(h, w, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    base = 0.5 + 0.3 * np.sin(x / 7.0) * np.cos(y / 11.0) + 0.15 * rng.standard_normal((h, w))
    return np.clip(base, 0, 1).astype(np.float32)


def _work(reg, side=0):
    return reg.work_image(side)


@pytest.mark.parametrize("size,grid,bins,half", [((96, 128), 4, 64, False), ((97, 131), 4, 64, True),
                                                  ((120, 90), 2, 8, False), ((64, 200), 8, 256, True),
                                                  ((257, 300), 3, 100, True)])
def test_clahe_bit_exact(reg_factory, size, grid, bins, half):
    img = _img(*size, seed=sum(size))
    reg = reg_factory(half=half, clahe=True, clahe_grid=grid, clahe_bins=bins, band=False, invert=False)
    reg.set_images(img, img[::-1].copy(), preprocess=True)
    ref = R.work_image(img, preprocess=True, half=half, clahe_on=True, clahe_bins=bins, clahe_grid=grid)
    got = _work(reg)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), np.abs(got - ref).max()
    r, hist, cdf, _ = O.rank_transform(ref)
    gr, ghist, gcdf, _ = reg.rank(0)
    assert np.array_equal(ghist, hist) and np.array_equal(gcdf, cdf)
    assert np.array_equal(gr.view(np.uint32), r.view(np.uint32))


@pytest.mark.parametrize("half,fine,coarse,inv", [(False, 1.0, 2.0, False), (True, 1.0, 2.0, True),
                                                   (True, 1.0, 6.0, False)])
def test_band_pass_and_invert_bit_exact(reg_factory, half, fine, coarse, inv):
    img = _img(80, 72, seed=3)
    reg = reg_factory(half=half, clahe=True, band=True, bp_fine=fine, bp_coarse=coarse, invert=inv)
    reg.set_images(img, img.T.copy(), preprocess=True)
    ref = R.work_image(img, preprocess=True, half=half, band=True, bp_fine=fine, bp_coarse=coarse, invert_on=inv)
    assert np.array_equal(_work(reg).view(np.uint32), ref.view(np.uint32))


def test_rgba_and_settings_rebake(reg_factory):
    rng = np.random.default_rng(9)
    rgba = rng.integers(0, 256, (70, 90, 4), dtype=np.uint8)
    reg = reg_factory(half=True)
    reg.set_images(rgba, rgba[:, ::-1].copy())          # uint8: always preprocessed (CLAHE on by default)
    gray = R.rgba_to_gray(rgba)
    assert np.array_equal(_work(reg).view(np.uint32), R.work_image(gray, half=True).view(np.uint32))
    # a settings change re-bakes the work image and the features from the kept input
    reg.configure(clahe=False, invert=True)
    assert np.array_equal(_work(reg).view(np.uint32),
                          R.work_image(gray, half=True, clahe_on=False, invert_on=True).view(np.uint32))
    reg.configure(half=False)
    ref = R.work_image(gray, half=False, clahe_on=False, invert_on=True)
    r, hist, _, _ = O.rank_transform(ref)
    gr, ghist, _, _ = reg.rank(0)
    assert np.array_equal(_work(reg).view(np.uint32), ref.view(np.uint32)) and np.array_equal(ghist, hist)


def test_preprocess_off_is_the_gray(reg_factory):
    img = _img(50, 60, seed=5)
    reg = reg_factory(half=True)
    reg.set_images(img, img, preprocess=False)
    assert np.array_equal(_work(reg).view(np.uint32), R.round_half(img).view(np.uint32))
