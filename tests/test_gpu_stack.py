"""GPU stack registration: uneven slice pairs with random homographies and intensity remaps,
registered together (shift-map start, batched Gauss–Newton), against the known poses and the
float64 oracle's Gauss–Newton from the same start.
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


def test_stack_drift_within_one_pixel(reg_factory):
    """Five 384² slices whose texture and pose drift (rotation about -6°, scale about 1.04,
    translation (10..16, -7..-3) px, perspective (2e-5, -1e-5)), the fixed side another modality:
    every slice within 1 px of its homography (the bar of tests/test_api.py:186-202)."""
    sl = synth.drift_stack(5, 384)
    reg = reg_factory(half=False)
    reg.set_stack([s[0] for s in sl], [s[1] for s in sl])
    Hs, sc, Hi, it = reg.register_stack(iters=30, init="map")
    errs = [_corner_err(Hs[k], Ht, 384, 384) for k, (_, _, Ht) in enumerate(sl)]
    assert max(errs) < 1.0, errs


def test_stack_random_homographies(reg_factory):
    """Five slices of 320–448 px (aspect up to 1.5), rotation ±10°, scale 0.9–1.1, perspective
    <= 1e-4, remaps 1 − x², x^0.5, sin(πx), started from a small rotation × scale sweep: every
    slice within 1 px of its homography, and no slice scoring below its true pose."""
    sl = synth.stack_slices(n=5, side=(320, 448), seed=4)
    reg = reg_factory(half=False)
    reg.set_stack([s[0] for s in sl], [s[1] for s in sl])
    Hs, sc, Hi, it = reg.register_stack(iters=30, init="sweep")
    assert it >= 1 and np.all(sc > 0)
    errs = [_corner_err(Hs[k], Ht, mv.shape[1], mv.shape[0]) for k, (mv, fx, Ht, _) in enumerate(sl)]
    assert max(errs) < 1.0, errs
    truth = _scores_at(reg_factory, sl)
    assert all(sc[k] >= truth[k] * (1 - 1e-3) for k in range(len(sl))), (list(sc), truth)


def _scores_at(reg_factory, sl):
    """The symmetric score of each slice at its true pose."""
    out = []
    for mv, fx, Ht, _ in sl:
        r = reg_factory(half=False)
        r.set_images(mv, fx, preprocess=False)
        out.append(r.score(Ht).mean)
        r.close()
    return out


def test_stack_gn_matches_oracle(reg_factory):
    """Ten Gauss–Newton iterations of one slice from the same start: the GPU's pose and score as
    the oracle's."""
    mv, fx, Ht, _ = synth.stack_slices(n=1, side=(160, 200), seed=9)[0]
    reg = reg_factory(half=False)
    reg.set_stack([mv], [fx])
    _, _, Hi, _ = reg.register_stack(iters=0, init="map")
    Hs, sc, _, _ = reg.register_stack(H0=Hi, iters=10, init=None)
    fa, fb = O.prepare(mv)["feat"], O.prepare(fx)["feat"]
    (Ho, so), = G.gn_batch(fa, fb, [Hi[0]], 10, amps=G.geo_amps(0.08, 0.5, 7))
    h, w = mv.shape
    assert _corner_err(Hs[0], Ho, w, h) < 0.1
    assert abs(sc[0] - so) <= 1e-3 * abs(so)


def test_stack_device_batch_matches_host_loop(reg_factory):
    """The Gauss–Newton batch run on the device (one captured graph per iteration) reaches the
    poses and scores of the host-driven loop from the same starts: same arithmetic, FP64 on both."""
    sl = synth.stack_slices(n=3, side=(200, 260), seed=12)
    out = {}
    for graph in (True, False):
        reg = reg_factory(half=False, gn_graph=graph)
        reg.set_stack([s[0] for s in sl], [s[1] for s in sl])
        _, _, Hi, _ = reg.register_stack(iters=0, init="map")
        out[graph] = reg.register_stack(H0=Hi, iters=12, init=None)
    (Hd, sd, _, itd), (Hh, sh, _, ith) = out[True], out[False]
    assert itd == ith
    for k, (mv, _, _, _) in enumerate(sl):
        h, w = mv.shape
        assert _corner_err(Hd[k], Hh[k], w, h) < 1e-3
        assert abs(sd[k] - sh[k]) <= 1e-6 * abs(sh[k])
