"""Pinning of the oracle's pose gradient and Gauss–Newton pieces (CPU only).

Each analytic quantity is checked against a central difference of the float64 function it
differentiates (the method of examples/03_smi.py:318-359), the pose algebra against scipy and
numpy, and the Gauss–Newton step at the optimum gnBatch reaches against zero.
"""
import math

import numpy as np
import pytest
import scipy.linalg

from oracle import gn_ref as G
from oracle import smi_ref as O
from zcmir_amd import synth


def test_lie_basics():
    rng = np.random.default_rng(0)
    for scale in (0.1, 3.0):
        A = rng.standard_normal((3, 3)) * scale
        assert np.allclose(G.expm3(A), scipy.linalg.expm(A), rtol=1e-10, atol=1e-12)
    M = rng.standard_normal((3, 3)) + 3 * np.eye(3)
    assert np.allclose(G.inv3(M), np.linalg.inv(M), rtol=1e-13)
    for group in ("homography", "affine"):
        x = rng.standard_normal(8)
        if group == "affine":
            x[6:] = 0
        assert np.allclose(G.vee(G.hat(x, group), group), x, atol=1e-14)
    # composeN with ξ = translation tx moves the destination by tx · w / 2 pixels
    H = G.compose_n(np.eye(3), [0.1, 0, 0, 0, 0, 0, 0, 0], "homography", 200, 100)
    assert np.allclose(H, [[1, 0, 10], [0, 1, 0], [0, 0, 1]], atol=1e-12)
    A = rng.standard_normal((5, 5)) + 5 * np.eye(5)
    b = rng.standard_normal(5)
    assert np.allclose(G.solve_linear(A, b), np.linalg.solve(A, b), rtol=1e-12)
    assert G.solve_linear(np.zeros((2, 2)), np.ones(2)) is None
    assert G.geo_amps(0.08, 0.5, 7) == pytest.approx([0.08 * 0.5 ** k for k in range(7)])
    assert G.beats(1.0 + 2e-6, 1.0) and not G.beats(1.0 + 5e-7, 1.0)


def _pair(size=(72, 80), seed=1):
    mv, fx, _ = synth.shift_pair(size, (3, -2), seed=seed)
    return O.prepare(mv)["feat"], O.prepare(fx)["feat"]


@pytest.mark.parametrize("exact", [True, False])
def test_score_coefs_gradient_over_sums(exact):
    fa, fb = _pair()
    H = np.array([[1.01, 0.02, -2.5], [-0.01, 0.99, 1.7], [1e-5, -1e-5, 1.0]])
    st = O.pose_moments(fa, fb, H)
    co = G.score_coefs(st, exact)
    assert co["score"] == pytest.approx(O.score_from_moments(st, exact), rel=1e-12)
    for k in range(45):
        h = 1e-6 * max(abs(st[k]), 1.0)
        sp, sm = st.copy(), st.copy()
        sp[k] += h
        sm[k] -= h
        fd = (G.score_coefs(sp, exact)["score"] - G.score_coefs(sm, exact)["score"]) / (2 * h)
        assert co["g"][k] == pytest.approx(fd, rel=2e-4, abs=1e-6 * co["score"]), k


@pytest.mark.parametrize("group,exact", [("homography", True), ("affine", True), ("homography", False)])
def test_tangent_gradient_matches_differences(group, exact):
    fa, fb = _pair()
    H = np.array([[1.01, 0.02, -2.5], [-0.01, 0.99, 1.7], [2e-5, -1e-5, 1.0]])
    if group == "affine":
        H[2] = [0, 0, 1]
    r = G.grad_dir(fa, fb, H, exact, group, True)
    nk = G.n_keys(group)
    _, hb, wb = fb.shape
    S = lambda Hx: G.score_coefs(O.pose_moments(fa, fb, Hx), exact)["score"]
    for k in range(nk):
        e = np.zeros(8)
        e[k] = 1e-5
        fd = (S(G.compose_n(H, e, group, wb, hb)) - S(G.compose_n(H, -e, group, wb, hb))) / 2e-5
        assert r["grad"][k] == pytest.approx(fd, rel=5e-3, abs=2e-3 * np.abs(r["grad"]).max()), k
    # the Gauss–Newton matrix is symmetric and positive semi-definite
    A = r["hess"]
    assert np.allclose(A, A.T, rtol=1e-10, atol=1e-10 * np.abs(A).max())
    assert np.linalg.eigvalsh(A).min() > -1e-8 * np.abs(A).max()


def test_symmetric_gradient_matches_differences():
    fa, fb = _pair()
    H = np.array([[1.01, 0.02, -2.5], [-0.01, 0.99, 1.7], [2e-5, -1e-5, 1.0]])
    r = G.pair_grad(fa, fb, H, True, "homography", True, True)
    _, hb, wb = fb.shape
    for k in range(8):
        e = np.zeros(8)
        e[k] = 1e-5
        fd = (G.pair_score(fa, fb, G.compose_n(H, e, "homography", wb, hb)) -
              G.pair_score(fa, fb, G.compose_n(H, -e, "homography", wb, hb))) / 2e-5
        assert r["grad"][k] == pytest.approx(fd, rel=5e-3, abs=2e-3 * np.abs(r["grad"]).max()), k


def test_gn_step_vanishes_at_optimum():
    """An integer shift pair: gnBatch from a pose 1.5 px off climbs back to the true pose, and the
    Gauss–Newton step at the pose it reaches is small (the criterion of tests/test_api.py:219-221,
    |step| < 5e-3 tangent units)."""
    mv, fx, Ht = synth.shift_pair((96, 96), (5, -3), seed=2)
    fa, fb = O.prepare(mv)["feat"], O.prepare(fx)["feat"]
    H0 = Ht @ np.array([[1, 0, 1.5], [0, 1, -1.0], [0, 0, 1]])
    (H1, s1), = G.gn_batch(fa, fb, [H0], 25, amps=G.geo_amps(0.08, 0.5, 7))
    pts = np.array([[1, 1, 1], [96, 1, 1], [1, 96, 1], [96, 96, 1]], float).T
    a, b = H1 @ pts, Ht @ pts
    assert np.abs(a[:2] / a[2] - b[:2] / b[2]).max() < 0.5
    assert s1 > G.pair_score(fa, fb, H0)
    r = G.pair_grad(fa, fb, H1, True, "homography", True, True)
    A = r["hess"] + 1e-4 * np.diag(np.maximum(np.diag(r["hess"]), 1e-12))
    step = G.solve_linear(A, 0.5 * r["grad"])
    assert np.linalg.norm(step) < 5e-3


def test_climb_hop_reaches_the_shift_gn_alone_cannot():
    """The oracle climb (gclimb.zig run): 6.5 px off the truth, beyond the Gauss–Newton basin's
    reach in 25 iterations of the hop-less climb, the shift-map hop lands on the true shift's cell
    and the Gauss–Newton block refines it to the same end pose as the climb from 1.5 px off; the
    Gauss–Newton step there is small (tests/test_api.py:219-221's criterion)."""
    mv, fx, Ht = synth.shift_pair((96, 96), (5, -3), seed=2)
    pa, pb = O.prepare(mv), O.prepare(fx)
    pts = np.array([[1, 1, 1], [96, 1, 1], [1, 96, 1], [96, 96, 1]], float).T

    def err(Ha, Hb):
        a, b = Ha @ pts, Hb @ pts
        return np.abs(a[:2] / a[2] - b[:2] / b[2]).max()

    far = Ht @ np.array([[1, 0, 6.5], [0, 1, -4.0], [0, 0, 1]])
    near = Ht @ np.array([[1, 0, 1.5], [0, 1, -1.0], [0, 0, 1]])
    Hh, sh, it_h, moved_h = G.climb(pa, pb, far, 25, hop=True)
    Hn, sn, _, _ = G.climb(pa, pb, near, 25, hop=False)
    Hg, sg, _, _ = G.climb(pa, pb, far, 25, hop=False)
    assert moved_h and it_h >= 2 and err(Hh, Hn) < 0.05 and abs(sh - sn) <= 1e-6 * abs(sn)
    assert err(Hg, Ht) > 1.0 and sg < sh
    r = G.pair_grad(pa["feat"], pb["feat"], Hh, True, "homography", True, True)
    A = r["hess"] + 1e-4 * np.diag(np.maximum(np.diag(r["hess"]), 1e-12))
    assert np.linalg.norm(G.solve_linear(A, 0.5 * r["grad"])) < 5e-3
