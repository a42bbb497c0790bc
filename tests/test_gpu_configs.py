"""The large workloads at their stated sizes, on the GPU, through the C ABI: properties that hold
whatever the size (the planted pose is found; the search lands on it; the stack registers).

* the rotation × scale sweep of a 1024² pair (moving = fixed rotated 37°, scaled 1.12, remapped
  x^0.5): 72 rotations × 9 scales, 648 candidate maps, on the reference's 128² canvas and at full
  resolution (planes of 2048, 2304 and 3072 points);
* the uneven stack: 64 slices of 512–1536 px, random homographies and remaps, 30 Gauss–Newton
  iterations on the device after a small sweep per slice.
The large SMI maps (1024² and 2048² pairs) are covered in test_gpu_map.py.
"""
import numpy as np
import pytest

from zcmir_amd import synth

pytestmark = pytest.mark.gpu

SWEEP_GRID = dict(sw_nth=72, sw_smin=0.8, sw_smax=1.25, sw_ns=9)


def _corner_err(Ha, Hb, w, h):
    pts = np.array([[0.5, 0.5, 1], [w + 0.5, 0.5, 1], [0.5, h + 0.5, 1], [w + 0.5, h + 0.5, 1]]).T
    p, q = Ha @ pts, Hb @ pts
    return float(np.max(np.hypot(*(p[:2] / p[2] - q[:2] / q[2]))))


@pytest.fixture(scope="module")
def rotscale_1024():
    return synth.rotscale_pair((1024, 1024), theta_deg=37.0, scale=1.12, seed=2)


@pytest.mark.parametrize("canvas", [dict(sw_grid=128), dict(sw_full=True)])
def test_sweep_648_candidates_1024_finds_the_planted_cell(reg_factory, rotscale_1024, canvas):
    """Every one of the 648 candidates is scored; the best is the grid cell nearest the planted
    pose (moving → fixed is a turn of −37° and a scale of 1/1.12: the cell −35°, 1.25^−0.5)."""
    mv, fx, _ = rotscale_1024
    reg = reg_factory(half=False, symmetric=True, **SWEEP_GRID, **canvas)
    reg.set_images(mv, fx, preprocess=False)
    info, peaks, best = reg.sweep()
    assert info["nc"] == 648 and best.shape == (648,) and np.all(np.isfinite(best))
    k = int(np.argmax(best))
    assert -180 + 5 * (k // 9) == -35 and k % 9 == 2, k
    if canvas.get("sw_full"):
        assert info["n"] == 3072 and info["cell"] <= 1.0


def test_search_648_candidates_1024_lands_on_the_pose(reg_factory, rotscale_1024):
    """The search (sweep on the 128² canvas, NMS, Gauss–Newton screen and finals) ends within
    2 px of the planted pose at the image corners (the SMI optimum of a bilinear-resampled,
    remapped pair sits within about a pixel of the planted pose)."""
    mv, fx, Ht = rotscale_1024
    reg = reg_factory(half=False, symmetric=True, **SWEEP_GRID, sw_grid=128)
    reg.set_images(mv, fx, preprocess=False)
    reg.search()
    assert _corner_err(reg.pose, Ht, 1024, 1024) < 2.0


def test_stack_64_uneven_slices_register(reg_factory):
    """64 slices of 512–1536 px (aspect up to 1.5) with random homographies (±10°, 0.9–1.1,
    perspective <= 1e-4) and remaps (1 − x², √x, sin πx): after a small sweep per slice and 30
    Gauss–Newton iterations on the device, at least 95 % of the slices are within 1 px of their
    planted pose at the corners."""
    sl = synth.stack_slices(n=64, side=(512, 1536), seed=3)
    reg = reg_factory(half=False)
    reg.set_stack([q[0] for q in sl], [q[1] for q in sl])
    Hs, _, _, it = reg.register_stack(iters=30, init="sweep")
    errs = np.array([_corner_err(Hs[k], q[2], q[0].shape[1], q[0].shape[0]) for k, q in enumerate(sl)])
    assert it >= 1
    assert np.mean(errs < 1.0) >= 0.95, (np.mean(errs < 1.0), np.sort(errs)[-5:])
