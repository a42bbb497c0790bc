"""GPU parity of the SMI hot path against the float64 oracle, through the C ABI.

Bars: rank transform (histogram, CDF, bins, rank values, mean/spread) bit-exact; maps within
max |Δ| / max |map| <= 1e-4 on lags clear of the overlap floor, with the same argmax lag;
FFT within 2e-6 of the largest output magnitude (float32 Stockham against float64 pocketfft).
"""
import numpy as np
import pytest

from oracle import smi_ref as O
from tests._maps import compare
from zcmir_amd import synth

pytestmark = pytest.mark.gpu

MAP_TOL = 1e-4


def _pair(size, shift, seed=0, quantize=False):
    return synth.shift_pair(size, shift, seed=seed, quantize=quantize)


@pytest.mark.parametrize("seed,size,half", [(0, (256, 256), False), (1, (97, 131), False), (2, (64, 200), True),
                                             (3, (300, 257), True)])
def test_rank_bit_exact(reg_factory, seed, size, half):
    rng = np.random.default_rng(seed)
    img = rng.uniform(0, 1, size).astype(np.float32)
    img[:5] = np.round(img[:5] * 255) / 255      # ties in some rows
    reg = reg_factory(half=half, clahe=False)
    reg.set_images(img, img[::-1].copy(), preprocess=False)
    g = O.round_half(img) if half else img
    r, hist, cdf, bins = O.rank_transform(g)
    mean, spread = O.image_stats(g)
    gr, ghist, gcdf, (gm, gs) = reg.rank(0)
    assert gm.tobytes() == np.float32(mean).tobytes() and gs.tobytes() == np.float32(spread).tobytes()
    assert np.array_equal(ghist, hist)
    assert np.array_equal(gcdf, cdf)
    assert np.array_equal(gr.view(np.uint32), r.view(np.uint32))


def test_features_match_oracle(reg_factory):
    mv, fx, _ = _pair((160, 192), (7, -5), seed=4)
    for mode in ("laplacian", "three", "cubic"):
        reg = reg_factory(half=False, smi_features=mode)
        reg.set_images(mv, fx, preprocess=False)
        for side, img in ((0, mv), (1, fx)):
            p = O.prepare(img, mode)
            f = reg.features(side)
            assert np.abs(f - p["feat"]).max() < 2e-4, mode


@pytest.mark.parametrize("n,planes", [(32, 3), (64, 2), (256, 2), (1024, 1), (1536, 1), (2048, 1), (2304, 1),
                                      (3072, 1), (4096, 1)])
def test_fft2_matches_numpy(reg_factory, n, planes):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal((planes, n, n)) + 1j * rng.standard_normal((planes, n, n))).astype(np.complex64)
    reg = reg_factory()
    ref = np.fft.fft2(x.astype(np.complex128))
    y = reg.fft2(x.copy())
    assert np.abs(y - ref).max() / np.abs(ref).max() < 2e-6
    back = reg.fft2(y.copy(), inverse=True)
    assert np.abs(back - x).max() / np.abs(x).max() < 2e-6


def _map_case(reg_factory, mv, fx, H, exact=True, calibrated=False, mode="laplacian", cap=1024, min_overlap=0.0):
    reg = reg_factory(half=False, smi_features=mode, exact=exact, calibrated=calibrated, map_res=cap,
                      min_overlap=min_overlap)
    reg.set_images(mv, fx, preprocess=False)
    if min_overlap > 0:
        sm = reg.shift_map(H)
    else:
        sm = reg.shift_map(H, exact=exact, calibrated=calibrated, resolution=cap)
    pa, pb = O.prepare(mv, mode), O.prepare(fx, mode)
    om, info = O.shift_map(pa, pb, H, exact=exact, calibrated=calibrated, cap=cap, min_overlap=min_overlap)
    return sm, om, info


@pytest.mark.parametrize("exact,calibrated,mode", [(True, False, "laplacian"), (True, False, "three"),
                                                   (False, False, "laplacian"), (True, True, "laplacian"),
                                                   (True, False, "cubic")])
def test_shift_map_256_shift_17_m9(reg_factory, exact, calibrated, mode):
    mv, fx, Ht = _pair((256, 256), (17, -9))
    sm, om, info = _map_case(reg_factory, mv, fx, np.eye(3), exact, calibrated, mode)
    assert (sm.n, sm.ny, sm.cw, sm.ch) == (512, 512, 256, 256)
    err, same, d = compare(sm.map, sm, om, info)
    assert err <= MAP_TOL, (err, d)
    assert same, ((sm.dx, sm.dy), (info["dx"], info["dy"]))
    assert (sm.dx, sm.dy) == (17.0, -9.0)
    if calibrated:
        assert abs(sm.corr_area_px - info["area_px"]) < 1e-3 * info["area_px"]


def test_shift_map_min_overlap_25pct(reg_factory):
    mv, fx, _ = _pair((256, 256), (17, -9))
    sm, om, info = _map_case(reg_factory, mv, fx, np.eye(3), min_overlap=0.25, calibrated=False)
    err, same, d = compare(sm.map, sm, om, info)
    assert err <= MAP_TOL and same and (sm.dx, sm.dy) == (17.0, -9.0)


def test_shift_map_warped_resampled(reg_factory):
    """A rotated/scaled pose on a rectangular pair, map capped below the canvas (k x k resample)."""
    mv, fx, _ = _pair((180, 240), (5, 3), seed=7)
    c = np.array([[1, 0, 120.5], [0, 1, 90.5], [0, 0, 1.0]])
    t = np.deg2rad(4.0)
    R = np.array([[1.03 * np.cos(t), -1.03 * np.sin(t), 0], [1.03 * np.sin(t), 1.03 * np.cos(t), 0], [0, 0, 1]])
    H = c @ R @ np.linalg.inv(c) @ np.array([[1, 0, -6.0], [0, 1, 4.0], [0, 0, 1]])
    sm, om, info = _map_case(reg_factory, mv, fx, H, cap=256)
    assert sm.cw < sm.canvas[0]
    err, same, d = compare(sm.map, sm, om, info)
    assert err <= MAP_TOL, (err, d)
    assert same


@pytest.mark.parametrize("exact,mode", [(True, "laplacian"), (True, "three"), (False, "laplacian")])
def test_shift_map_warped_full_resolution(reg_factory, exact, mode):
    """The same warped pose at one cell per canvas pixel: the fixed rectangle sits inside a larger
    canvas (negative offsets), so the rectangle sums see clipped, shifted rectangles."""
    mv, fx, _ = _pair((180, 240), (5, 3), seed=7)
    c = np.array([[1, 0, 120.5], [0, 1, 90.5], [0, 0, 1.0]])
    t = np.deg2rad(4.0)
    R = np.array([[1.03 * np.cos(t), -1.03 * np.sin(t), 0], [1.03 * np.sin(t), 1.03 * np.cos(t), 0], [0, 0, 1]])
    H = c @ R @ np.linalg.inv(c) @ np.array([[1, 0, -6.0], [0, 1, 4.0], [0, 0, 1]])
    sm, om, info = _map_case(reg_factory, mv, fx, H, exact=exact, mode=mode, cap=1024)
    assert (sm.cw, sm.ch) == tuple(sm.canvas) and info["offset"][0] < 0 and info["offset"][1] < 0
    err, same, d = compare(sm.map, sm, om, info)
    assert err <= MAP_TOL, (err, d)
    assert same


@pytest.mark.parametrize("mode", ["laplacian", "three"])
def test_shift_map_1024_full_resolution(reg_factory, mode):
    """1024² pair, 8-bit intensities, shift (113, -57), full-resolution 2048² FFT."""
    mv, fx, _ = _pair((1024, 1024), (113, -57), seed=11, quantize=True)
    sm, om, info = _map_case(reg_factory, mv, fx, np.eye(3), mode=mode, cap=2048)
    assert (sm.n, sm.ny, sm.cw, sm.ch) == (2048, 2048, 1024, 1024)
    err, same, d = compare(sm.map, sm, om, info)
    assert err <= MAP_TOL, (err, d)
    assert same and (sm.dx, sm.dy) == (113.0, -57.0)


def test_moments_and_score(reg_factory):
    mv, fx, Ht = _pair((200, 220), (6, -4), seed=5)
    H = np.array([[1.01, 0.02, -5.0], [-0.015, 0.99, 3.5], [1e-5, -2e-5, 1.0]])
    for sym in (False, True):
        reg = reg_factory(half=False, symmetric=sym, exact=True)
        reg.set_images(mv, fx, preprocess=False)
        pa, pb = O.prepare(mv), O.prepare(fx)
        m = reg.moments(H)
        om = O.pose_moments(pa["feat"], pb["feat"], H)
        assert np.abs(m - om).max() < 1e-4 * np.abs(om).max()
        s = reg.score(H)
        mean, fwd, inv = O.pose_score(pa["feat"], pb["feat"], H, exact=True, symmetric=sym)
        assert abs(s.fwd - fwd) < 1e-4 * abs(fwd)
        assert abs(s.mean - mean) < 1e-4 * abs(mean)
        assert (s.inv is None) == (not sym)


def test_errors_fail_loudly(reg_factory):
    reg = reg_factory()
    with pytest.raises(RuntimeError, match="NoImages"):
        reg.shift_map(np.eye(3), exact=True, calibrated=False)
    with pytest.raises(RuntimeError, match="BadImage"):
        reg.set_moving(np.zeros((4, 4), np.float32), preprocess=False)
    with pytest.raises(RuntimeError, match="FeatureModeMismatch|NoImages|BadImage"):
        reg.set_moving(np.zeros((6, 16), np.float32), preprocess=True)
    with pytest.raises(RuntimeError, match="UnknownSetting"):
        reg.configure(nope=1)


@pytest.mark.parametrize("n", [256, 1024, 1536, 2048, 2304, 3072, 4096])
def test_fft2_every_column_tiling(reg_factory, n):
    """Each columns-per-workgroup setting of the column FFT (zc_fft_tune) computes the same FFT."""
    import zcmir_amd

    lib = zcmir_amd._load()
    rng = np.random.default_rng(n + 1)
    x = (rng.standard_normal((1, n, n)) + 1j * rng.standard_normal((1, n, n))).astype(np.complex64)
    ref = np.fft.fft2(x.astype(np.complex128))
    reg = reg_factory()
    ran = 0
    try:
        for cols in (1, 2, 4, 8, 16, 9, 17, 33, 68, 72, 80, 41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53):
            lib.zc_fft_tune(0, cols)
            try:
                y = reg.fft2(x.copy())
            except RuntimeError as e:
                assert "UnsupportedFftLength" in str(e)
                continue
            ran += 1
            assert np.abs(y - ref).max() / np.abs(ref).max() < 2e-6, cols
    finally:
        lib.zc_fft_tune(0, 0)
    assert ran >= 2


def test_autotuned_tiling_leaves_maps_unchanged(reg_factory):
    """The column-FFT autotuner picks among tilings that compute every column with the same
    stages: the map with the tuned tiling equals the built-in default's bit for bit."""
    mv, fx, _ = _pair((256, 256), (17, -9))
    maps = []
    for tune in (True, False):
        reg = reg_factory(half=False, fft_autotune=tune)
        reg.set_images(mv, fx, preprocess=False)
        if tune:
            assert reg.fft_tiling(512) in (0, 1, 2, 4, 8, 16, 9, 17, 33, 68, 72, 80, 41, 42, 43, 44, 45, 46, 47, 48, 49)
        maps.append(reg.shift_map(np.eye(3), exact=True, calibrated=False, resolution=1024).map)
    assert np.array_equal(maps[0].view(np.uint32), maps[1].view(np.uint32))


def test_row_tilings_of_the_inverse_pass(reg_factory):
    """The inverse row pass's tilings (zc_fft_tune rows): those with the default's stages (16·16·8
    for 2048) give the default map bit for bit; 8 points per lane (8·8·8·4) within float32
    rounding, with the same peak."""
    import zcmir_amd

    lib = zcmir_amd._load()
    mv, fx, _ = _pair((1024, 1024), (113, -57), seed=11, quantize=True)
    maps = {}
    try:
        for rows in (0, 1, 2, 3):
            lib.zc_fft_tune(rows, 0)
            reg = reg_factory(half=False)
            reg.set_images(mv, fx, preprocess=False)
            sm = reg.shift_map(np.eye(3), exact=True, calibrated=False, resolution=2048)
            assert (sm.dx, sm.dy) == (113.0, -57.0), rows
            maps[rows] = sm.map
            reg.close()
    finally:
        lib.zc_fft_tune(0, 0)
    for rows in (1, 3):
        assert np.array_equal(maps[rows].view(np.uint32), maps[0].view(np.uint32)), rows
    assert np.abs(maps[2] - maps[0]).max() <= 1e-5 * np.abs(maps[0]).max()


@pytest.mark.parametrize("mode", ["laplacian", "three"])
def test_shift_map_2048_full_resolution_sampled_lags(reg_factory, mode):
    """2048² pair (8-bit, shift (113, -57)), 45 and 28 correlations: the full-resolution map
    (4096² FFT, rectangle sums over 2048-wide tables) at a dozen lags against the float64 oracle's
    score of the sums taken directly at each lag (a whole float64 4096² map would take minutes);
    the peak at the shift."""
    mv, fx, _ = _pair((2048, 2048), (113, -57), seed=12, quantize=True)
    reg = reg_factory(half=False, smi_features=mode)
    reg.set_images(mv, fx, preprocess=False)
    sm = reg.shift_map(np.eye(3), exact=True, calibrated=False, resolution=4096)
    assert (sm.n, sm.ny, sm.cw, sm.ch) == (4096, 4096, 2048, 2048)
    assert (sm.dx, sm.dy) == (113.0, -57.0)
    pa, pb = O.prepare(mv, mode), O.prepare(fx, mode)
    ox, oy, W, Hh = O.reg_frame(np.eye(3), 2048, 2048, 2048, 2048)
    ra = O.real_planes(O.warp_stack(pa["stack"], O.canvas_h(np.eye(3), ox, oy), W, Hh), True)
    rb = O.real_planes(O.paste_stack(pb["stack"], ox, oy, W, Hh), True)
    pairs = O.corr_list(True)

    def sums(tx, ty):   # Σ_x a(x + t) b(x) for all 45 pairs at lag t
        ys, ye, xs, xe = max(0, -ty), min(Hh, Hh - ty), max(0, -tx), min(W, W - tx)
        A = ra[:, ys + ty:ye + ty, xs + tx:xe + tx].reshape(len(ra), -1)
        B = rb[:, ys:ye, xs:xe].reshape(len(rb), -1)
        G = A @ B.T
        return np.array([G[a, b] for a, b in pairs])

    lags = [(0, 0), (113, -57), (112, -57), (113, -56), (30, 40), (-700, 900), (1500, -300),
            (-1200, -1200), (1000, 1000), (-1850, 120), (400, -1700)]
    corr = np.stack([sums(tx, ty) for tx, ty in lags], axis=1)
    n0 = corr[0, 0]
    om = O.combine(corr, exact=True, nf=3 if mode == "three" else 4, zero_n=n0)
    floor = max(O.MIN_N, 0.08 * n0)
    scale = np.abs(sm.map).max()
    for k, (tx, ty) in enumerate(lags):
        if abs(corr[0, k] - floor) <= 1e-3 * floor + 1.0:
            continue   # at the overlap floor either side may round to 0
        g = sm.map[ty % sm.ny, tx % sm.n]
        assert abs(g - om[k]) <= MAP_TOL * scale, ((tx, ty), g, om[k], scale)
    assert abs(sm.map[-57 % sm.ny, 113] - scale) == 0


@pytest.mark.parametrize("size,shift,exact,mode", [((97, 131), (5, -3), True, "laplacian"),
                                                   ((131, 97), (-7, 4), False, "three"),
                                                   ((33, 250), (11, 2), True, "three")])
def test_shift_map_ragged_full_resolution(reg_factory, size, shift, exact, mode):
    """Odd, ragged pairs at one cell per pixel: odd summed-area table widths, power-of-two FFTs
    well above 2 W − 1, a thin 33-row image (bands shorter than the table)."""
    mv, fx, _ = _pair(size, shift, seed=sum(size))
    sm, om, info = _map_case(reg_factory, mv, fx, np.eye(3), exact=exact, mode=mode, cap=1024)
    assert (sm.cw, sm.ch) == tuple(sm.canvas)
    err, same, d = compare(sm.map, sm, om, info)
    assert err <= MAP_TOL, (err, d)
    assert same and (sm.dx, sm.dy) == (float(shift[0]), float(shift[1]))


def test_shift_map_on_the_reference_grid(reg_factory):
    """A full-resolution map whose reference plan takes 2^a·3^b lengths (576 x 648 for a 257 x 300
    pair): the kernels run on power-of-two lengths, and the map, its size and its peak index come
    back on the reference grid, in a buffer sized as the reference binding sizes it
    (max(resolution, map_res, 1024)² floats), matching the oracle on that grid."""
    mv, fx, _ = _pair((300, 257), (9, -4), seed=5)
    sm, om, info = _map_case(reg_factory, mv, fx, np.eye(3), cap=1500)
    assert (sm.n, sm.ny) == (576, 648) == (info["n"], info["ny"])   # 257 wide, 300 high
    assert sm.map.shape == (648, 576)
    err, same, d = compare(sm.map, sm, om, info)
    assert err <= MAP_TOL, (err, d)
    assert same and (sm.dx, sm.dy) == (9.0, -4.0)
    assert sm.map.reshape(-1).argmax() == ((-4) % 648) * 576 + 9


@pytest.mark.parametrize("size,H,exact,mode", [((180, 240), (9.0, -6.0), True, "laplacian"),
                                               ((131, 97), (-4.0, 7.0), True, "three"),
                                               ((150, 150), (0.0, 0.0), False, "laplacian"),
                                               ((1024, 1024), (0.0, 0.0), True, "laplacian")])
def test_translation_plans_match_oracle_and_rectangle_plans(reg_factory, size, H, exact, mode):
    """Full-resolution maps at integer-translation poses take the translation plans (the moving
    weight plane is a rectangle as well, so the correlations with it are rectangle sums of the
    fixed side's tables and only the feature cross terms go through the FFT): the map matches the
    float64 oracle and the rectangle plans' map (zc_map_path(0)) with the same peak, for shifted
    canvases (negative offsets, the moving rectangle clipped), three features and global SMI."""
    import zcmir_amd

    lib = zcmir_amd._load()
    mv, fx, _ = _pair(size, (5, -3), seed=sum(size), quantize=size[0] >= 1024)
    Hp = np.array([[1, 0, H[0]], [0, 1, H[1]], [0, 0, 1.0]])
    cap = 2048 if size[0] >= 1024 else 1024
    sm, om, info = _map_case(reg_factory, mv, fx, Hp, exact=exact, mode=mode, cap=cap)
    assert (sm.cw, sm.ch) == tuple(sm.canvas)
    err, same, d = compare(sm.map, sm, om, info)
    assert err <= MAP_TOL, (err, d)
    assert same
    try:
        lib.zc_map_path(0)
        reg = reg_factory(half=False, smi_features=mode)
        reg.set_images(mv, fx, preprocess=False)
        rect = reg.shift_map(Hp, exact=exact, calibrated=False, resolution=cap)
        reg.close()
    finally:
        lib.zc_map_path(1)
    assert (rect.dx, rect.dy) == (sm.dx, sm.dy)
    assert np.abs(rect.map - sm.map).max() <= MAP_TOL * np.abs(rect.map).max()


@pytest.mark.parametrize("size,H,exact,mode,calibrated", [((1024, 1024), (0.0, 0.0), True, "laplacian", False),
                                                          ((256, 256), (0.0, 0.0), True, "three", False),
                                                          ((256, 256), (0.0, 0.0), False, "laplacian", False),
                                                          ((256, 256), (0.0, 0.0), True, "laplacian", True),
                                                          ((256, 256), (3.0, -2.0), True, "laplacian", False),
                                                          ((200, 256), (0.0, 0.0), True, "laplacian", False)])
def test_translation_plan_quad_cells_bit_identical(reg_factory, size, H, exact, mode, calibrated):
    """Where both images have the same size and the FFT plane is twice as large, the translation
    plans' score kernel loads the summed-area cells of four lags once (zc_map_path(1), the
    default); every lag reading its own four cells per sum (zc_map_path(2)) gives the same map bit
    for bit, also where the quad form does not apply (a shifted pose, unequal sizes)."""
    import zcmir_amd

    lib = zcmir_amd._load()
    mv, fx, _ = _pair(size, (5, -3), seed=7 + size[1], quantize=size[0] >= 1024)
    Hp = np.array([[1, 0, H[0]], [0, 1, H[1]], [0, 0, 1.0]])
    cap = 2048 if size[0] >= 1024 else 512
    maps = []
    try:
        for path in (1, 2):
            lib.zc_map_path(path)
            reg = reg_factory(half=False, smi_features=mode)
            reg.set_images(mv, fx, preprocess=False)
            maps.append(reg.shift_map(Hp, exact=exact, calibrated=calibrated, resolution=cap))
            reg.close()
    finally:
        lib.zc_map_path(1)
    a, b = maps
    assert (a.dx, a.dy) == (b.dx, b.dy)
    assert a.map.tobytes() == b.map.tobytes()
