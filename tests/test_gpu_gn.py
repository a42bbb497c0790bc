"""GPU parity of the pose gradient, the Gauss–Newton climb and the sweep search against the
float64 oracle (oracle/gn_ref.py), through the C ABI.

Bars: score within 1e-4 relative; gradient and Gauss–Newton matrix within 2e-4 of their largest
entry (FP32 per-pixel sums, FP64 across blocks); the climb and the search reach the oracle's /
the true pose within a fraction of a pixel at the image corners.
"""
import numpy as np
import pytest

from oracle import gn_ref as G
from oracle import smi_ref as O
from zcmir_amd import synth

pytestmark = pytest.mark.gpu


def _corner_err(Ha, Hb, w, h):
    pts = np.array([[1, 1, 1], [w, 1, 1], [1, h, 1], [w, h, 1]], float).T
    a, b = Ha @ pts, Hb @ pts
    return float(np.abs(a[:2] / a[2] - b[:2] / b[2]).max())


@pytest.mark.parametrize("symmetric,exact,group", [(False, True, "homography"), (True, True, "homography"),
                                                   (True, False, "homography"), (True, True, "affine")])
def test_gradient_matches_oracle(reg_factory, symmetric, exact, group):
    mv, fx, _ = synth.shift_pair((120, 136), (4, -3), seed=8)
    H = np.array([[1.01, 0.02, -2.5], [-0.01, 0.99, 1.7], [2e-5, -1e-5, 1.0]])
    if group == "affine":
        H[2] = [0, 0, 1]
    reg = reg_factory(half=False, symmetric=symmetric, exact=exact, group=group)
    reg.set_images(mv, fx, preprocess=False)
    g = reg.gradient(H, hess=True)
    fa, fb = O.prepare(mv)["feat"], O.prepare(fx)["feat"]
    o = G.pair_grad(fa, fb, H, exact, group, symmetric, True)
    nk = G.n_keys(group)
    assert g.grad.shape == (nk,) and g.hess.shape == (nk, nk)
    assert abs(g.score - o["score"]) <= 1e-4 * abs(o["score"])
    assert np.abs(g.grad - o["grad"][:nk]).max() <= 2e-4 * np.abs(o["grad"][:nk]).max()
    assert np.abs(g.hess - o["hess"]).max() <= 2e-4 * np.abs(o["hess"]).max()


@pytest.mark.parametrize("offset,hop,hop_fm", [((1.5, -1.0, 0.0, 1.0), False, False),
                                               ((6.5, -4.0, 0.0, 1.0), True, False),
                                               ((2.0, 1.0, 7.0, 1.06), True, True)])
def test_climb_matches_oracle(reg_factory, offset, hop, hop_fm):
    """The climb (shift-map hop, Gauss–Newton blocks, roto-scale hop) from a start pose off the
    truth by a translation and a turn / rescale about the centre follows the oracle's climb: the
    same number of iterations and the same end pose within 0.1 px at the corners."""
    import math

    mv, fx, Ht = synth.shift_pair((96, 96), (5, -3), seed=2)
    tx, ty, deg, sc = offset
    a = math.radians(deg)
    C = np.array([[1, 0, 48.5], [0, 1, 48.5], [0, 0, 1.0]])
    R = np.array([[sc * math.cos(a), -sc * math.sin(a), 0], [sc * math.sin(a), sc * math.cos(a), 0], [0, 0, 1]])
    H0 = C @ R @ np.linalg.inv(C) @ Ht @ np.array([[1, 0, tx], [0, 1, ty], [0, 0, 1]])
    reg = reg_factory(half=False, g_iters=25, hop=hop, hop_fm=hop_fm)
    reg.set_images(mv, fx, preprocess=False)
    reg.pose = H0
    c = reg.climb()
    pa, pb = O.prepare(mv), O.prepare(fx)
    Ho, so, it, moved = G.climb(pa, pb, H0, 25, hop=hop, hop_fm=hop_fm)
    assert c.moved and moved and c.iterations == it, (c.iterations, it)
    assert _corner_err(c.H, Ho, 96, 96) < 0.1
    assert _corner_err(c.H, Ht, 96, 96) < 1.5
    assert abs(c.score - so) <= 1e-3 * abs(so)
    assert np.allclose(reg.pose, c.H)


@pytest.mark.parametrize("offset,hop,hop_fm,symmetric", [((1.5, -1.0, 0.0, 1.0), False, False, True),
                                                         ((6.5, -4.0, 0.0, 1.0), True, False, True),
                                                         ((2.0, 1.0, 7.0, 1.06), True, True, True),
                                                         ((2.0, 1.0, 7.0, 1.06), True, True, False)])
def test_climb_on_device_matches_host_loop(reg_factory, offset, hop, hop_fm, symmetric):
    """The climb with its decisions on the device (climb_graph, the default: one captured graph
    per iteration) takes the host-driven loop's path: the same iteration count and end pose; a
    second climb from the same start replays the cached graph from its first iteration and ends
    at the same pose bit for bit."""
    import math

    mv, fx, Ht = synth.shift_pair((96, 96), (5, -3), seed=2)
    tx, ty, deg, sc = offset
    a = math.radians(deg)
    C = np.array([[1, 0, 48.5], [0, 1, 48.5], [0, 0, 1.0]])
    R = np.array([[sc * math.cos(a), -sc * math.sin(a), 0], [sc * math.sin(a), sc * math.cos(a), 0], [0, 0, 1]])
    H0 = C @ R @ np.linalg.inv(C) @ Ht @ np.array([[1, 0, tx], [0, 1, ty], [0, 0, 1]])
    out = {}
    for graph in (False, True):
        reg = reg_factory(half=False, g_iters=25, hop=hop, hop_fm=hop_fm, symmetric=symmetric, climb_graph=graph)
        reg.set_images(mv, fx, preprocess=False)
        reg.pose = H0
        out[graph] = reg.climb()
        if graph:
            reg.pose = H0
            again = reg.climb()
            assert again.iterations == out[graph].iterations
            assert np.array_equal(again.H, out[graph].H) and again.score == out[graph].score
    d, h = out[True], out[False]
    assert d.iterations == h.iterations and d.moved == h.moved, (d.iterations, h.iterations)
    assert _corner_err(d.H, h.H, 96, 96) < 1e-3
    assert abs(d.score - h.score) <= 1e-6 * abs(h.score)


def test_search_exact_quarter_turn(reg_factory):
    n = 96
    _, fixed, _ = synth.shift_pair((n, n), (0, 0), seed=5)
    moving = 1.0 - np.rot90(fixed) ** 2
    reg = reg_factory(half=False, sw_grid=32, sw_nth=12, sw_smin=1.0, sw_smax=1.0, sw_ns=1)
    reg.set_images(moving, fixed, preprocess=False)
    s = reg.search()
    truth = np.array([[0, -1, n + 1], [1, 0, 0], [0, 0, 1.0]])
    assert s.improved and s.count_a >= 1 and s.count_b >= 1
    assert _corner_err(s.H, truth, n, n) < 0.6   # 96² with 8 homography parameters
    assert np.allclose(reg.pose, s.H)


def test_search_rotation_scale_256(reg_factory):
    """Moving = fixed rotated 37°, scaled 1.12, then x^0.5 at 256²: the search (72 θ × 9 σ sweep,
    Gauss–Newton on the peaks) reaches the true pose's basin."""
    mv, fx, Ht = synth.rotscale_pair((256, 256), theta_deg=37.0, scale=1.12, seed=2)
    reg = reg_factory(half=False, sw_grid=128, sw_nth=72, sw_smin=0.8, sw_smax=1.25, sw_ns=9)
    reg.set_images(mv, fx, preprocess=False)
    s = reg.search()
    assert s.improved
    at_truth = reg.score(Ht).mean
    err = _corner_err(s.H, Ht, 256, 256)
    # the search's pose scores at least as high as the true pose, and lies within a few pixels of it
    assert s.score >= at_truth * (1 - 1e-3) and err < 3.0, (s.score, at_truth, err)
