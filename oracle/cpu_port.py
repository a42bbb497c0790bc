"""Float32 CPU restatement of the reference's shift-map schedule — the CPU baseline that bench.py
times beside the GPU (test / benchmark infrastructure; never part of the product path).

Follows Smi.shiftMapFrom → packFft → correlate → combine (smi.zig:483-572) with the same plan as
the HIP path: two real planes packed per complex FFT (smi.wgsl:708-728), the Hermitian unpack and
A conj(B) per correlation pair, two correlations packed per complex inverse FFT
(smi.wgsl:730-796), then the per-lag whitened score (smi.wgsl:825-931). FFTs are scipy.fft
(pocketfft) in complex64 with `workers` threads; the combine is vectorized NumPy float32.
"""
from __future__ import annotations

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


def combine_f32(corr, exact=True, nf=4, min_n=O.MIN_N, ridge=np.float32(1e-4)):
    """The per-lag score in float32 (the reference kernel's precision, smi.wgsl:825-931)."""
    with np.errstate(all="ignore"):
        return _combine_f32(corr, exact, nf, min_n, ridge)


def _combine_f32(corr, exact, nf, min_n, ridge):
    c = corr.reshape(corr.shape[0], -1)
    n = c[0]
    floor = max(float(min_n), 0.08 * max(float(c[0, 0]), 0.0))
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
