"""GPU parity of the rotation × scale sweep against the float64 oracle, through the C ABI.

Bars: each candidate's score map within max |Δ| / max |map| <= 1e-4 on lags clear of the
overlap floor; each candidate's best local maximum at the same lag with the same score to 1e-4;
the sweep's landscape (best score per candidate) within 1e-4 of the oracle's maximum; the
recovered pose equal to the oracle's (an exact quarter turn, and the nearest grid cell of an
off-grid rotation / scale).
"""
import math

import numpy as np
import pytest

from oracle import smi_ref as O
from zcmir_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _single(theta_deg, sigma):
    """Settings of a one-candidate sweep at (θ, σ)."""
    lo, hi = (sigma * sigma, 1.0) if sigma < 1 else (1.0, sigma * sigma)
    return dict(sw_rot_range=True, sw_rot_lo=theta_deg, sw_rot_hi=theta_deg, sw_nth=1, sw_smin=lo, sw_smax=hi,
                sw_ns=1)


def _compare_map(gm, om):
    m, n, floor = om["map"], om["n"], om["floor"]
    assert gm.shape == m.shape
    clear = np.abs(n - floor) > 1e-3 * max(floor, 1.0) + 1.0
    scale = max(np.abs(m).max(), 1e-30)
    return float(np.abs(gm - m)[clear].max() / scale), scale


@pytest.mark.parametrize("theta,sigma,grid,exact,mode", [
    (37.0, 1.12, 64, True, "laplacian"),
    (-120.0, 0.9, 32, True, "laplacian"),
    (10.0, 1.0, 128, True, "three"),
    (37.0, 1.12, 64, False, "laplacian"),
])
def test_sweep_candidate_map(reg_factory, theta, sigma, grid, exact, mode):
    mv, fx, _ = synth.rotscale_pair((112, 136), theta_deg=-theta, scale=1 / sigma, seed=3)
    reg = reg_factory(half=False, smi_features=mode, exact=exact, sw_grid=grid, **_single(theta, sigma))
    reg.set_images(mv, fx, preprocess=False)
    info, peaks, best = reg.sweep()
    gm = reg.sweep_map()
    pm, pf = O.prepare(mv, mode), O.prepare(fx, mode)
    gv = O.sweep_grid(theta, theta, 1, *((sigma * sigma, 1.0) if sigma < 1 else (1.0, sigma * sigma)), 1)
    r = O.sweep(pm, pf, np.array([[1, 0, (136 - 136) / 2], [0, 1, (112 - 112) / 2], [0, 0, 1.0]]), gv, grid=grid,
                exact=exact, keep_maps=True)
    assert info["nc"] == 1 and info["grid"] == grid and abs(info["cell"] - r["cell"]) < 1e-12
    err, scale = _compare_map(gm, r["maps"][0])
    assert err <= TOL, (err, scale)
    assert len(peaks) > 0 and len(r["peaks"]) > 0
    assert (peaks[0]["sx"], peaks[0]["sy"]) == (r["peaks"][0]["sx"], r["peaks"][0]["sy"])
    assert abs(peaks[0]["score"] - r["peaks"][0]["score"]) <= TOL * r["peaks"][0]["score"]
    assert np.allclose(peaks[0]["H"], r["peaks"][0]["H"], rtol=1e-12, atol=1e-9)


def test_sweep_full_resolution_candidate(reg_factory):
    """sw_full: one canvas cell per fixed pixel (grid = ceil(D))."""
    mv, fx, _ = synth.rotscale_pair((72, 80), theta_deg=-20.0, scale=1.0, seed=4)
    reg = reg_factory(half=False, sw_full=True, **_single(20.0, 1.0))
    reg.set_images(mv, fx, preprocess=False)
    info, peaks, _ = reg.sweep()
    gm = reg.sweep_map()
    gv = O.sweep_grid(20.0, 20.0, 1, 1.0, 1.0, 1)
    r = O.sweep(O.prepare(mv), O.prepare(fx), np.eye(3), gv, full=True, keep_maps=True)
    assert info["grid"] == r["grid"] == math.ceil(r["D"] - 1e-9) and info["cell"] <= 1.0
    err, _ = _compare_map(gm, r["maps"][0])
    assert err <= TOL
    assert (peaks[0]["sx"], peaks[0]["sy"]) == (r["peaks"][0]["sx"], r["peaks"][0]["sy"])


@pytest.mark.parametrize("size", [512, 1024])
def test_sweep_smooth_planes_match_powers_of_two(reg_factory, size):
    """Full-resolution candidates whose windows need 1258-1356 (512² pair) or 2516-2715 (1024² pair)
    points run on planes of 1536 or 3072 points (radix-3 stages) instead of 2048 or 4096:
    a linear correlation does not depend on the plane length, so every candidate's best score and
    best lag match the power-of-two sweep (itself checked against the oracle above)."""
    mv, fx, _ = synth.rotscale_pair((size, size), theta_deg=37.0, scale=1.12, seed=9)
    runs = {}
    for smooth in (False, True):
        reg = reg_factory(half=False, sw_full=True, sweep_smooth=smooth, sw_rot_range=True, sw_rot_lo=33.0,
                          sw_rot_hi=41.0, sw_nth=3, sw_smin=1.1, sw_smax=1.4, sw_ns=2)
        reg.set_images(mv, fx, preprocess=False)
        info, peaks, best = reg.sweep()
        runs[smooth] = (info, peaks, best, reg.sweep_map().shape)
    (ip, pp, bp, mp), (is_, ps, bs, ms) = runs[False], runs[True]
    want = 1536 if size == 512 else 3072
    assert ip["n"] == 4 * want // 3 and mp == (4 * want // 3,) * 2, (ip["n"], mp)
    assert is_["n"] == want and ms[0] in (1536, 2304, 3072), (is_["n"], ms)
    assert np.abs(bs - bp).max() <= TOL * bp.max()
    def firsts(lst):
        out = {}
        for p in lst:
            out.setdefault(int(p["cand"]), p)
        return out
    fp, fs = firsts(pp), firsts(ps)
    assert set(fp) == set(fs)
    for c in fp:
        assert (fp[c]["sx"], fp[c]["sy"]) == (fs[c]["sx"], fs[c]["sy"]), c


def test_sweep_landscape_matches_oracle(reg_factory):
    """A 12 θ × 3 σ sweep: every candidate's best score and best lag as the oracle's."""
    mv, fx, _ = synth.rotscale_pair((128, 128), theta_deg=37.0, scale=1.12, seed=6)
    reg = reg_factory(half=False, sw_grid=64, sw_nth=12, sw_smin=0.85, sw_smax=1.15, sw_ns=3)
    reg.set_images(mv, fx, preprocess=False)
    info, peaks, best = reg.sweep()
    gv = O.sweep_grid(-180, 180, 12, 0.85, 1.15, 3)
    r = O.sweep(O.prepare(mv), O.prepare(fx), np.eye(3), gv, grid=64)
    assert info["nc"] == 36
    assert np.abs(best - r["best"]).max() <= TOL * r["best"].max()
    def firsts(lst):   # each candidate's first listed peak is its best
        out = {}
        for p in lst:
            out.setdefault(p["cand"], p)
        return out
    top_g, top_o = firsts(peaks), firsts(r["peaks"])
    assert set(top_g) == set(top_o)
    for c, po in top_o.items():
        pg = top_g[c]
        assert (pg["sx"], pg["sy"]) == (po["sx"], po["sy"]), c


def test_sweep_exact_quarter_turn(reg_factory):
    n = 96
    _, fixed, _ = synth.shift_pair((n, n), (0, 0), seed=5)
    moving = np.rot90(fixed).copy()
    moving = 1.0 - moving * moving
    reg = reg_factory(half=False, sw_grid=32, sw_nth=12, sw_smin=1.0, sw_smax=1.0, sw_ns=1)
    reg.set_images(moving, fixed, preprocess=False)
    info, peaks, best = reg.sweep()
    top = O.nms4(peaks, 12, 8, True)
    assert np.isclose(np.rad2deg(top[0]["th"]), 90.0) and (top[0]["sx"], top[0]["sy"]) == (0, 0)
    assert np.allclose(top[0]["H"], [[0, -1, n + 1], [1, 0, 0], [0, 0, 1]], atol=1e-9)
    assert int(np.argmax(best)) == 9


def test_sweep_nearest_grid_cell(reg_factory):
    """Moving = fixed rotated 37° and scaled 1.12, then x^0.5 (the rotation × scale workload at
    256²): the best NMS peak of 72 θ × 9 σ (5° steps, 0.8 – 1.25) is the nearest grid cell of the
    inverse pose, θ = −35°, σ = 1.25^(−2/4)."""
    mv, fx, Ht = synth.rotscale_pair((256, 256), theta_deg=37.0, scale=1.12, seed=2)
    reg = reg_factory(half=False, sw_grid=128, sw_nth=72, sw_smin=0.8, sw_smax=1.25, sw_ns=9)
    reg.set_images(mv, fx, preprocess=False)
    info, peaks, best = reg.sweep()
    assert info["nc"] == 648
    top = O.nms4(peaks, 72, 32, True)
    assert np.isclose(np.rad2deg(top[0]["th"]), -35.0)
    assert np.isclose(top[0]["s"], 1.25 ** -0.5)
    # the peak pose is within a few pixels of the truth over the image
    pts = np.array([[1, 1, 1], [256, 1, 1], [1, 256, 1], [256, 256, 1], [128, 128, 1]], float).T
    a = top[0]["H"] @ pts
    b = Ht @ pts
    err = np.abs(a[:2] / a[2] - b[:2] / b[2]).max()
    assert err < 12.0, err


def test_sweep_graph_replays_match_eager(reg_factory):
    """The sweep's launches replayed from the captured HIP graph give the eager launches' peaks bit
    for bit on every replay, with other GPU work and synchronisations between replays, and again
    after the key changes (another batch size: a new capture); the batch size (4, 7, the default)
    changes no result."""
    import zcmir_amd

    lib = zcmir_amd._load()
    mv, fx, _ = synth.rotscale_pair((128, 128), theta_deg=37.0, scale=1.12, seed=6)
    kw = dict(half=False, sw_grid=64, sw_nth=12, sw_smin=0.85, sw_smax=1.15, sw_ns=3)
    first = None
    try:
        for batch in (4, 7, 0):
            lib.zc_sweep_tune(batch)        # several batches per plane size: many launches per graph
            eager = reg_factory(sweep_graph=False, **kw)
            eager.set_images(mv, fx, preprocess=False)
            _, pe, be = eager.sweep()
            if first is None:
                first = (pe.tobytes(), be.tobytes())
            assert (pe.tobytes(), be.tobytes()) == first, batch
            graph = reg_factory(sweep_graph=True, **kw)
            graph.set_images(mv, fx, preprocess=False)
            for r in range(4):
                _, pg, bg = graph.sweep()
                assert pg.tobytes() == pe.tobytes() and bg.tobytes() == be.tobytes(), (batch, r)
                assert len(pg) > 0 and not pg["_pad"].any()
                assert np.array_equal(graph.sweep_map(), eager.sweep_map()), (batch, r)
                graph.shift_map(np.eye(3), exact=True, calibrated=False)   # eager launches between replays
    finally:
        lib.zc_sweep_tune(0)


@pytest.mark.parametrize("rows", [1, 3, 4])
def test_sweep_row_tilings_agree_on_lags_without_overlap(reg_factory, rows):
    """A full-resolution candidate of a 640² pair (2048-point planes, about 400 lag rows and columns
    where its footprint and the fixed image cannot overlap) gives the default tiling's map bit for
    bit under the inverse row tilings that run the same radix stages at 2048 points (32 or 16
    points per lane): tiling 4 holds two lines per workgroup, so lag rows without overlap next to
    rows with it take the per-line path, while the others store n = 0 per whole workgroup. The
    default tiling's candidate maps are checked against the oracle above."""
    import zcmir_amd

    lib = zcmir_amd._load()
    mv, fx, _ = synth.rotscale_pair((640, 640), theta_deg=-37.0, scale=1 / 1.12, seed=9)
    kw = dict(half=False, sw_full=True, sweep_smooth=False, **_single(37.0, 1.12))
    out = {}
    try:
        for r in (0, rows):
            lib.zc_fft_tune(r, 0)
            reg = reg_factory(**kw)
            reg.set_images(mv, fx, preprocess=False)
            info, peaks, best = reg.sweep()
            out[r] = (info["n"], reg.sweep_map(), peaks[0])
    finally:
        lib.zc_fft_tune(0, 0)
    (n0, m0, p0), (n1, m1, p1) = out[0], out[rows]
    assert n0 == n1 == 2048 and m0.shape == m1.shape == (2048, 2048)
    assert (m0 == 0.0).mean() > 0.3, (m0 == 0.0).mean()   # lags below the floor, those without overlap among them
    assert np.array_equal(m0, m1)
    assert (p0["sx"], p0["sy"], p0["score"]) == (p1["sx"], p1["sy"], p1["score"])
