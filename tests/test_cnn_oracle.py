"""Pinning of the prior network's oracle (CPU only): every gradient of oracle/cnn_ref.py against
central differences of its float64 forward pass, and im2col / col2im as adjoints."""
import numpy as np
import pytest

from oracle import cnn_ref as C

LAYERS = [(2, 4, 1), (4, 4, 2), (4, 8, 2)]


def _net(rng):
    params = [(rng.standard_normal((9 * ci, co)) * np.sqrt(2 / (9 * ci)), rng.standard_normal(co) * 0.1)
              for ci, co, _ in LAYERS]
    params.append((rng.standard_normal((2 * 2 * 8, 5)) * 0.1, rng.standard_normal(5) * 0.1))
    return params


def test_im2col_adjoint():
    rng = np.random.default_rng(0)
    for s in (1, 2):
        x = rng.standard_normal((2, 8, 6, 3))
        c = rng.standard_normal((2 * (8 // s) * (6 // s), 27))
        assert np.isclose(np.sum(C.im2col(x, s) * c), np.sum(x * C.col2im(c, x.shape, s)))


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159], np.float32)
    assert np.array_equal(C.bf16(x), np.array([1.0, 1.0, 1.0 + 2 ** -7, -2.5, 3.140625], np.float32))


def test_network_gradients_match_differences():
    rng = np.random.default_rng(1)
    params = _net(rng)
    x = rng.standard_normal((2, 8, 8, 2))
    labels = np.array([1, 3])

    def loss_of(ps, xx):
        acts, z = C.net_forward(xx, ps, LAYERS)
        return C.xent(z, labels)[0].mean()

    acts, z = C.net_forward(x, params, LAYERS)
    _, dz = C.xent(z, labels)
    grads, dx = C.net_backward(acts, params, LAYERS, dz)
    h = 1e-6
    for li, (W, b) in enumerate(params):
        for arr, g in ((W, grads[li][0]), (b, grads[li][1])):
            for idx in list(np.ndindex(arr.shape))[:: max(1, arr.size // 12)]:
                old = arr[idx]
                arr[idx] = old + h
                lp = loss_of(params, x)
                arr[idx] = old - h
                lm = loss_of(params, x)
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                assert g[idx] == pytest.approx(fd, rel=1e-4, abs=1e-8), (li, idx)
    for idx in list(np.ndindex(x.shape))[::37]:
        old = x[idx]
        x[idx] = old + h
        lp = loss_of(params, x)
        x[idx] = old - h
        lm = loss_of(params, x)
        x[idx] = old
        assert dx[idx] == pytest.approx((lp - lm) / (2 * h), rel=1e-4, abs=1e-8), idx


def test_lp_features_rotation_is_a_shift():
    """A 90° turn of the image moves its log-polar magnitude map by T/2 angle bins (exact on the
    pixel grid); a 2× zoom out by log(2)/log(r1/r0)·R radius bins."""
    from zcmir_amd import synth
    img = synth.shift_pair((256, 256), (0, 0), seed=3)[1]
    a = C.lp_features(img, S=256, T=48, R=64)
    b = C.lp_features(np.rot90(img).copy(), S=256, T=48, R=64)
    err = [np.abs(np.roll(a, k, axis=0) - b).mean() for k in range(48)]
    assert int(np.argmin(err)) == 24


def test_xcorr_channel_is_the_direct_correlation():
    """The input's third channel against the direct double sum c[t, d + R/2] = Σ_{t', r}
    a[(t' + t) mod T, r + d] b[t', r] / (T (R − |d|)) (then standardized), and a map shifted by
    k angle bins and j radius bins against itself peaks at lag (k, j)."""
    import torch

    from zcmir_amd import prior

    rng = np.random.default_rng(11)
    T, R = 12, 10
    a, b = rng.standard_normal((T, R)), rng.standard_normal((T, R))
    got = prior.xcorr_channel(torch.tensor(a[None]), torch.tensor(b[None]))[0].numpy()
    ref = np.zeros((T, R))
    for t in range(T):
        for d in range(-(R // 2), R - R // 2):
            s = 0.0
            for t2 in range(T):
                for r in range(R):
                    if 0 <= r + d < R:
                        s += a[(t2 + t) % T, r + d] * b[t2, r]
            ref[t, d + R // 2] = s / (T * (R - abs(d)))
    ref = (ref - ref.mean()) / ref.std(ddof=1)
    assert np.abs(got - ref).max() < 1e-9
    big = rng.standard_normal((48, 64))
    moved = np.roll(big, 5, axis=0)
    moved[:, 3:] = moved[:, :-3].copy()
    c = prior.xcorr_channel(torch.tensor(moved[None]), torch.tensor(big[None]))[0].numpy()
    t, d = np.unravel_index(np.argmax(c), c.shape)
    assert (t, d - 32) == (5, 3)


def test_fine_cells_cover_the_coarse_rotation():
    from zcmir_amd import prior

    assert prior.fine_cells([(0, 4)]) == [(71, 4), (0, 4), (1, 4)]          # −180° ± 5° wraps
    assert prior.fine_cells([(5, 2), (5, 2), (6, 2)]) == [(14, 2), (15, 2), (16, 2), (17, 2), (18, 2), (19, 2)]
    assert prior.fine_cells([(0, 4)], margin=1) == [(70, 4), (71, 4), (0, 4), (1, 4), (2, 4)]
    assert prior.truth_cell(-180.0, 1.0) == (0, 4) and prior.truth_cell(172.6, 0.8) == (0, 0)
    assert prior.truth_cell(37.0, 1.12) == (14, 6)


def test_fine_cells_follow_the_grid_by_angle_and_scale():
    """The proposals map onto any sweep grid by value: at the registrar's default 90 rotations
    (4° steps) class 0 (−180°) covers the rotations within ±7.5° (−184°, −180°, −176°), class 23
    (165°) four rotations (160°..172°), never the 12°-per-class cells an index rule would give; a
    5-scale grid over the same range maps a class to the scale at its value, a class between two
    grid scales to the nearer; a partial rotation range keeps only its own rotations."""
    from zcmir_amd import prior

    grid = {**prior.TRAIN_GRID, "sw_nth": 90}
    assert prior.fine_cells([(0, 4)], grid) == [(89, 4), (0, 4), (1, 4)]
    assert prior.fine_cells([(23, 4)], grid) == [(85, 4), (86, 4), (87, 4), (88, 4)]
    th, sg, wrap = prior.sweep_axes(grid)
    assert wrap and len(th) == 90 and abs(th[85] - 160.0) < 1e-12 and abs(sg[4] - 1.0) < 1e-12
    five = {**prior.TRAIN_GRID, "sw_nth": 72, "sw_ns": 5}
    assert prior.fine_cells([(12, 2)], five) == [(35, 1), (36, 1), (37, 1)]     # 0°, scale 0.8^(1/2)
    assert prior.fine_cells([(12, 8)], five) == [(35, 4), (36, 4), (37, 4)]     # scale 1.25
    assert prior.fine_cells([(12, 3)], five) == [(35, 1), (36, 1), (37, 1)]     # between: the nearer (first)
    part = {**prior.TRAIN_GRID, "sw_rot_range": True, "sw_rot_lo": -30.0, "sw_rot_hi": 30.0, "sw_nth": 13}
    cells = prior.fine_cells([(12, 4), (10, 4)], part)   # 0° and −30°
    ths = prior.sweep_axes(part)[0]
    assert [ths[j] for j, _ in cells] == [-5.0, 0.0, 5.0, -30.0, -25.0]
    assert cells == [(5, 4), (6, 4), (7, 4), (0, 4), (1, 4)]
