"""Float32 CPU restatement of the reference's shift-map schedule — the CPU baseline that bench.py
times beside the GPU (test / benchmark infrastructure; never part of the product path).

Follows Smi.shiftMapFrom → packFft → correlate → combine (smi.zig:483-572) with the same plan as
the HIP path: two real planes packed per complex FFT (smi.wgsl:708-728), the Hermitian unpack and
A conj(B) per correlation pair, two correlations packed per complex inverse FFT
(smi.wgsl:730-796), then the per-lag whitened score (smi.wgsl:825-931). FFTs are scipy.fft
(pocketfft) in complex64 with `workers` threads; the combine is vectorized NumPy float32.
"""
from __future__ import annotations

import math
import os
import time

import numpy as np
import scipy.fft as sfft

from . import smi_ref as O


def _mirror(Z):
    """Z(-k) for every bin k of (..., ny, nx) planes."""
    return np.roll(np.flip(Z, axis=(-2, -1)), 1, axis=(-2, -1))


def map_f32(stack_a, stack_b, nx, ny, exact=True, three=False, workers=None, min_n=O.MIN_N):
    """Shift map of two grid stacks (5, ch, cw) (identity canvas), float32 end to end."""
    workers = workers or os.cpu_count()
    pa = O.real_planes(stack_a.astype(np.float64), exact).astype(np.float32)
    pb = O.real_planes(stack_b.astype(np.float64), exact).astype(np.float32)
    lst = O.corr_list(exact)
    keep = O.kept_corrs(exact, three)
    slots = [p for p in range(len(pa)) if not (three and p in O.FOURTH_PLANES)]
    slot_of = {p: i for i, p in enumerate(slots)}
    _, ch, cw = pa.shape

    def spectra(planes):
        R = len(slots)
        P = (R + 1) // 2
        buf = np.zeros((P, ny, nx), np.complex64)
        for s, p in enumerate(slots):
            if s % 2 == 0:
                buf[s // 2, :ch, :cw].real = planes[p]
            else:
                buf[s // 2, :ch, :cw].imag = planes[p]
        return sfft.fft2(buf, axes=(-2, -1), workers=workers, overwrite_x=True)

    ZA, ZB = spectra(pa), spectra(pb)
    ZAm, ZBm = _mirror(ZA), _mirror(ZB)

    def unpack(Z, Zm, s):
        z, zm = Z[s // 2], Zm[s // 2]
        if s % 2 == 0:
            return 0.5 * (z + np.conj(zm))
        return -0.5j * (z - np.conj(zm))

    corr = {}
    pairs = [lst[c] for c in keep]
    for q in range(0, len(pairs), 2):
        a1, b1 = slot_of[pairs[q][0]], slot_of[pairs[q][1]]
        P = unpack(ZA, ZAm, a1) * np.conj(unpack(ZB, ZBm, b1))
        if q + 1 < len(pairs):
            a2, b2 = slot_of[pairs[q + 1][0]], slot_of[pairs[q + 1][1]]
            P = P + 1j * (unpack(ZA, ZAm, a2) * np.conj(unpack(ZB, ZBm, b2)))
        x = sfft.ifft2(P.astype(np.complex64), axes=(-2, -1), workers=workers, overwrite_x=True)
        corr[keep[q]] = x.real.astype(np.float32)
        if q + 1 < len(pairs):
            corr[keep[q + 1]] = x.imag.astype(np.float32)
    full = np.zeros((len(lst), ny, nx), np.float32)
    for c, v in corr.items():
        full[c] = v
    return combine_f32(full, exact, 3 if three else 4, min_n)


def combine_f32(corr, exact=True, nf=4, min_n=O.MIN_N, ridge=np.float32(1e-4), zero_n=None):
    """The per-lag score in float32 (the reference kernel's precision, smi.wgsl:825-931); the
    overlap floor is 0.08 of zero_n (default: the overlap at lag 0)."""
    with np.errstate(all="ignore"):
        return _combine_f32(corr, exact, nf, min_n, ridge, zero_n)


def _combine_f32(corr, exact, nf, min_n, ridge, zero_n=None):
    c = corr.reshape(corr.shape[0], -1)
    n = c[0]
    z = float(c[0, 0]) if zero_n is None else float(zero_n)
    floor = max(float(min_n), 0.08 * max(z, 0.0))
    ok = ~(n < floor)
    den = np.maximum(n, np.float32(1e-8))
    inv = np.float32(1.0) / den
    ma = [c[1 + i] * inv for i in range(nf)]
    mb = [c[5 + i] * inv for i in range(nf)]
    C = [[c[9 + i * 4 + j] * inv - ma[i] * mb[j] for j in range(nf)] for i in range(nf)]
    if not exact:
        s = sum(np.clip(C[i][j], -1, 1) ** 2 for i in range(nf) for j in range(nf))
        return np.where(ok, n * s, 0).reshape(corr.shape[1:])
    ga = [[c[25 + O.tri_k(i, j)] * inv - ma[i] * ma[j] for j in range(nf)] for i in range(nf)]
    gb = [[c[35 + O.tri_k(i, j)] * inv - mb[i] * mb[j] for j in range(nf)] for i in range(nf)]
    tra = sum(ga[i][i] for i in range(nf))
    trb = sum(gb[i][i] for i in range(nf))
    for i in range(nf):
        ga[i][i] = ga[i][i] + ridge * np.maximum(tra * 0.25, 1e-12)
        gb[i][i] = gb[i][i] + ridge * np.maximum(trb * 0.25, 1e-12)
    pd = np.ones(n.shape, bool)
    for G in (ga, gb):
        for i in range(nf):
            for j in range(i + 1):
                s = G[i][j] - sum(G[i][k] * G[j][k] for k in range(j))
                if i == j:
                    pd &= s > 0
                    G[i][i] = np.sqrt(np.where(s > 0, s, 1))
                else:
                    G[i][j] = s / G[j][j]
    X = [[None] * nf for _ in range(nf)]
    for j in range(nf):
        for i in range(nf):
            s = C[i][j] - sum(ga[i][k] * X[k][j] for k in range(i))
            X[i][j] = s / ga[i][i]
    acc = np.zeros(n.shape, np.float32)
    for r in range(nf):
        y = [None] * nf
        for i in range(nf):
            s = X[r][i] - sum(gb[i][k] * y[k] for k in range(i))
            y[i] = s / gb[i][i]
            acc += y[i] * y[i]
    out = np.where(ok & pd, n * acc, 0)
    out = np.where(np.isfinite(out), out, 0)
    return out.reshape(corr.shape[1:])


def time_maps(stack_a, stack_b, nx, ny, exact=True, three=False, budget_s=20.0, workers=None):
    """Maps/s of map_f32 over whole maps, running until `budget_s` has passed (at least one)."""
    t0 = time.perf_counter()
    n = 0
    last = None
    while True:
        last = map_f32(stack_a, stack_b, nx, ny, exact, three, workers)
        n += 1
        el = time.perf_counter() - t0
        if el + el / n > budget_s:
            break
    return n / el, n, el, last


def _planes_f32(stack, exact):
    return O.real_planes(stack.astype(np.float64), exact).astype(np.float32)


def _correlate_f32(pa, pb, nx, ny, pairs, slot_a, slot_b, workers, spec_b=None):
    """Float32 correlations Σ a(x + t) b(x) of the listed (a, b) plane pairs on an nx × ny plane:
    two real planes per complex forward FFT, the Hermitian unpack, A conj(B), two correlations per
    complex inverse FFT (smi.wgsl:708-796). spec_b: side B's unpacked spectra (cached by the
    caller); returns (correlations, side B's spectra)."""
    def spectra(planes, slots):
        R = len(slots)
        buf = np.zeros(((R + 1) // 2, ny, nx), np.complex64)
        _, h, w = planes.shape
        for s, p in enumerate(slots):
            if s % 2 == 0:
                buf[s // 2, :h, :w].real = planes[p]
            else:
                buf[s // 2, :h, :w].imag = planes[p]
        Z = sfft.fft2(buf, axes=(-2, -1), workers=workers, overwrite_x=True)
        Zm = _mirror(Z)
        out = {}
        for s, p in enumerate(slots):
            z, zm = Z[s // 2], Zm[s // 2]
            out[p] = 0.5 * (z + np.conj(zm)) if s % 2 == 0 else -0.5j * (z - np.conj(zm))
        return out

    SA = spectra(pa, slot_a)
    SB = spec_b if spec_b is not None else spectra(pb, slot_b)
    res = []
    for q in range(0, len(pairs), 2):
        P = SA[pairs[q][0]] * np.conj(SB[pairs[q][1]])
        if q + 1 < len(pairs):
            P = P + 1j * (SA[pairs[q + 1][0]] * np.conj(SB[pairs[q + 1][1]]))
        x = sfft.ifft2(P.astype(np.complex64), axes=(-2, -1), workers=workers, overwrite_x=True)
        res.append(x.real.astype(np.float32))
        if q + 1 < len(pairs):
            res.append(x.imag.astype(np.float32))
    return res, SB


def sweep_f32(prep_m, prep_f, H0, gv, grid=128, exact=True, workers=None, max_cands=None, min_overlap=0.05):
    """Float32 restatement of the sweep schedule (search.zig:300-494, sweep.wgsl:106-371): per
    candidate the k×k-tap window of the moving features at its pose, its FFT correlations with the
    fixed window (spectra cached per plane size), the per-lag score with the candidate's own overlap
    floor and the top-4 local maxima. Windows and geometry as the oracle's sweep; arithmetic in
    float32 with scipy.fft (pocketfft, `workers` threads). Returns (candidates done, best scores)."""
    workers = workers or os.cpu_count()
    feat_m, feat_f = prep_m["feat"], prep_f["feat"]
    three = prep_m["mode"] == "three"
    _, mh, mw = feat_m.shape
    _, fh, fw = feat_f.shape
    H0 = np.asarray(H0, np.float64).reshape(3, 3)
    cm = O.map_pt(H0, (mw + 1) / 2, (mh + 1) / 2)
    s_max = max(O.sv2(O.cand_linear(0.0, s, k, a))[0] for s in gv["sigmas"] for k in gv["shears"] for a in gv["anisos"])
    D = min(2 * max(fw, fh), max(max(fw, fh), s_max * math.hypot(mw, mh)))
    cell = D / grid
    ox, oy = (fw + 1) / 2 - D / 2, (fh + 1) / 2 - D / 2
    f32 = lambda v: float(np.float32(v))   # noqa: E731
    oxf, oyf, cellf = f32(ox), f32(oy), f32(cell)
    fcx, fcy = O._cells(0.5, fw + 0.5, ox, cell, grid), O._cells(0.5, fh + 0.5, oy, cell, grid)
    Fx, Fy = fcx[1] - fcx[0], fcy[1] - fcy[0]
    plen = lambda need: 1 << max(5, (max(need, 1) - 1).bit_length())   # noqa: E731
    lst = O.corr_list(exact)
    keep = O.kept_corrs(exact, three)
    pairs = [lst[c] for c in keep]
    slots = [p for p in range(15 if exact else 5) if not (three and p in O.FOURTH_PLANES)]
    min_n = max(O.MIN_N, O.js_round(min_overlap * (fw * fh) / (cell * cell)))
    kf = int(min(6, max(1, math.ceil(cell - 1e-3))))
    s0 = math.sqrt(abs(H0[0, 0] * H0[1, 1] - H0[0, 1] * H0[1, 0])) or 1.0
    fixed, best, done = {}, [], 0
    for th in gv["thetas"]:
        for s in gv["sigmas"]:
            for kk in gv["shears"]:
                for a in gv["anisos"]:
                    if max_cands is not None and done >= max_cands:
                        return done, np.array(best)
                    L = O.cand_linear(th, s, kk, a)
                    H = O.about_c(cm, L) @ H0
                    pts = [O.map_pt(H, qx, qy) for qx, qy in ((0.5, 0.5), (mw + 0.5, 0.5), (mw + 0.5, mh + 0.5),
                                                              (0.5, mh + 0.5))]
                    xs, ys = [p[0] for p in pts], [p[1] for p in pts]
                    mcx = O._cells(min(xs), max(xs), ox, cell, grid)
                    mcy = O._cells(min(ys), max(ys), oy, cell, grid)
                    nx, ny = plen(mcx[1] - mcx[0] + Fx - 1), plen(mcy[1] - mcy[0] + Fy - 1)
                    taps = int(min(6, max(1, math.ceil(cell / (O.sv2(L)[1] * s0) - 1e-3))))
                    if (nx, ny) not in fixed:
                        sf = O.sweep_window(feat_f, np.eye(3), fcx[0], fcy[0], Fx, Fy, grid, kf, oxf, oyf, cellf)
                        fixed[(nx, ny)] = [_planes_f32(sf, exact), None]
                    pb, spec_b = fixed[(nx, ny)]
                    Hi = np.linalg.inv(H).astype(np.float32).astype(np.float64)
                    wx, wy = nx - Fx + 1, ny - Fy + 1
                    sm = O.sweep_window(feat_m, Hi, mcx[0], mcy[0], wx, wy, grid, taps, oxf, oyf, cellf)
                    pa = _planes_f32(sm, exact)
                    ck, spec_b = _correlate_f32(pa, pb, nx, ny, pairs, slots, slots, workers, spec_b)
                    fixed[(nx, ny)][1] = spec_b
                    corr = np.zeros((len(lst), ny, nx), np.float32)
                    for c, v in zip(keep, ck):
                        corr[c] = v
                    zl = ((fcy[0] - mcy[0]) % ny) * nx + ((fcx[0] - mcx[0]) % nx)
                    # the candidate's own floor: 0.08 n at its zero displacement (sweep.wgsl:238-317)
                    m = combine_f32(corr, exact, 3 if three else 4, min_n, zero_n=float(corr[0].ravel()[zl]))
                    pk = O.sweep_peaks(m, 4)
                    best.append(float(pk[0][0]) * cell * cell if pk and pk[0][0] > 0 else 0.0)
                    done += 1
    return done, np.array(best)


def time_sweep(prep_m, prep_f, H0, gv, grid=128, exact=True, budget_s=15.0, workers=None):
    """Candidates/s of sweep_f32 over the grid's candidates, in chunks until `budget_s` has passed."""
    t0 = time.perf_counter()
    done, _ = sweep_f32(prep_m, prep_f, H0, gv, grid, exact, workers, max_cands=9)
    el = time.perf_counter() - t0
    per = el / max(done, 1)
    n = max(9, min(int(budget_s / max(per, 1e-9)), 72 * 9))
    t0 = time.perf_counter()
    done, best = sweep_f32(prep_m, prep_f, H0, gv, grid, exact, workers, max_cands=n)
    el = time.perf_counter() - t0
    return done / el, done, el, best
