"""Map comparison helpers shared by the parity tests."""
import numpy as np


def lag_view(m, nx, ny, cw, ch):
    """The (2 ch - 1) x (2 cw - 1) block of lags -(c-1)..(c-1) of a wrapped map, centred."""
    ys = np.r_[ny - (ch - 1):ny, 0:ch] % ny
    xs = np.r_[nx - (cw - 1):nx, 0:cw] % nx
    return m[np.ix_(ys, xs)]


def compare(gpu_map, gpu_sm, orc_map, orc_info, floor_band=1e-3, tol=1e-4):
    """Max |Δ| / max |map| over lags whose overlap is clear of the floor, and whether the argmax
    lags agree. Returns (err, argmax_equal, details)."""
    cw, ch = orc_info["cw"], orc_info["ch"]
    assert (gpu_sm.cw, gpu_sm.ch) == (cw, ch)
    g = lag_view(gpu_map, gpu_sm.n, gpu_sm.ny, cw, ch)
    o = lag_view(orc_map, orc_info["n"], orc_info["ny"], cw, ch)
    n = lag_view(orc_info["corr"][0], orc_info["n"], orc_info["ny"], cw, ch)
    clear = np.ones(n.shape, bool)
    for th in orc_info["thresholds"]:
        clear &= np.abs(n - th) > floor_band * max(th, 1.0) + 1.0
    scale = max(np.abs(o).max(), 1e-30)
    err = float(np.abs(g - o)[clear].max() / scale) if clear.any() else 0.0
    same = (gpu_sm.dx, gpu_sm.dy) == (orc_info["dx"], orc_info["dy"])
    return err, same, dict(scale=scale, n_clear=int(clear.sum()))
