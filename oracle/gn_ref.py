"""Float64 NumPy oracle of the pose gradient and the Gauss–Newton refinement.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import this module, as the checker; the product path never does.

Restates (paths relative to the zcmir repository): the pose algebra of src/lie.zig (destN, inv3,
hat, expm3, projectGroup, composeN, vee, generators, adjointGrad, unitAscent, solveLinear), the
SMI gradient coefficients of src/moments.zig:91-172 (scoreFromMoments), the tangent gradient and
Gauss–Newton matrix of shaders/smi.wgsl:1129-1272 (tangent_grad), src/pose.zig:267-289
(gradFrom), src/pair.zig:51-53 (beats) and 241-268 (sym), src/climb.zig:15-24 (geoAmps) and
src/search.zig:539-614 (gnBatch).

Pinning: the coefficient vector g = dS/d(sums) against central differences of the score over
the 45 sums; the tangent gradient against central differences of the float64 pose score over
composeN steps; adjointGrad against the gradient of the inverse direction by differences; the
Gauss–Newton step at the true pose (≈ 0); expm3 against scipy.linalg.expm (tests/test_gn_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np

from oracle.smi_ref import MIN_N, RIDGE, TRI, moments, sample_pose, tri_k


# ── pose algebra (lie.zig) ────────────────────────────────────────────────────────────────
def n_keys(group: str) -> int:
    """lie.zig:22-24: homography 8 tangents, affine 6."""
    return 8 if group == "homography" else 6


def dest_n(w0, h0):
    """Pixel → normalized [-1, 1] coordinates of a w × h image (lie.zig:28-37)."""
    w, h = max(w0, 1), max(h0, 1)
    sx, sy = 0.5 * w, 0.5 * h
    return np.array([[1 / sx, 0, -0.5 * (w + 1) / sx], [0, 1 / sy, -0.5 * (h + 1) / sy], [0, 0, 1.0]])


def inv3(m):
    """Adjugate inverse with one Newton refinement step (lie.zig:93-117)."""
    m = np.asarray(m, np.float64).reshape(3, 3)
    a, b, c, d, e, f, g, h, i = m.ravel()
    A, B, C = e * i - f * h, f * g - d * i, d * h - e * g
    D, E, F = c * h - b * i, a * i - c * g, b * g - a * h
    G, Hh, I = b * f - c * e, c * d - a * f, a * e - b * d
    det = a * A + b * B + c * C
    sgn = 1.0 if det > 0 else (-1.0 if det < 0 else 0.0)
    with np.errstate(divide="ignore"):
        s = 1.0 / (sgn * max(abs(det), 1e-18))
    inv = np.array([[A, D, G], [B, E, Hh], [C, F, I]]) * s
    p = m @ inv
    return inv @ (2 * np.eye(3) - p)


def hat(x, group: str):
    """Tangent coordinates [tx, ty, th, sg, al, ga, px, py] → 3×3 generator (lie.zig:120-141)."""
    tx, ty, th, sg, al, ga, px, py = [float(v) for v in x]
    if group != "homography":
        px = py = 0.0
        z, iso = 0.0, sg
    else:
        iso = sg / 3
        z = -2 * iso
    return np.array([[iso + al, ga - th, tx], [ga + th, iso - al, ty], [px, py, z]])


def expm3(A):
    """Matrix exponential: scaling and squaring with a 13-term Taylor series (lie.zig:143-163)."""
    A = np.asarray(A, np.float64)
    nrm = np.abs(A).max()
    s = int(min(10, max(0, math.ceil(math.log2(nrm / 0.5))))) if nrm > 0.5 else 0
    B = A * 0.5 ** s
    T = np.eye(3)
    P = np.eye(3)
    for k in range(1, 14):
        P = P @ B / k
        T = T + P
    for _ in range(s):
        T = T @ T
    return T


def project_group(H, group: str):
    """lie.zig:165-175: divide by H[2,2]; affine drops the last row."""
    H = np.asarray(H, np.float64)
    s = H[2, 2] if abs(H[2, 2]) > 1e-12 else 1.0
    o = H / s
    if group != "homography":
        o[2] = [0, 0, 1]
    return o


def compose_n(H0, x, group: str, dw, dh):
    """exp(ξ̂) applied in normalized coordinates of a dw × dh destination (lie.zig:182-185)."""
    N = dest_n(dw, dh)
    return project_group(inv3(N) @ expm3(hat(x, group)) @ N @ np.asarray(H0, np.float64), group)


def vee(A0, group: str):
    """lie.zig:187-207, the inverse of hat on its range."""
    a, b, c, d, e, f, g, h, i = np.asarray(A0, np.float64).ravel()
    if group == "homography":
        tr = (a + e + i) / 3
        a -= tr
        e -= tr
        iso = (a + e) / 2
        return np.array([c, f, (d - b) / 2, iso * 3, (a - e) / 2, (d + b) / 2, g, h])
    return np.array([c, f, (d - b) / 2, (a + e) / 2, (a - e) / 2, (d + b) / 2, 0, 0])


def generators(group: str, dest=None):
    """One generator per tangent (zero past the group's), in pixel coordinates of a destination
    (N⁻¹ G N) when dest = (w, h) (lie.zig:209-225)."""
    nk = n_keys(group)
    out = []
    for k in range(8):
        if k >= nk:
            out.append(np.zeros((3, 3)))
            continue
        x = np.zeros(8)
        x[k] = 1
        G = hat(x, group)
        if dest is not None:
            N = dest_n(*dest)
            G = inv3(N) @ G @ N
        out.append(G)
    return out


def adjoint_grad(g_eta, H, group: str, mw, mh, fw, fh):
    """dS/dξ on H from dS/dη on H⁻¹ (η: left increment in the moving frame) (lie.zig:227-241)."""
    Hi = inv3(H)
    M = dest_n(mw, mh) @ Hi @ inv3(dest_n(fw, fh))
    Mi = inv3(M)
    gens = generators(group)
    nk = n_keys(group)
    out = np.zeros(8)
    for i in range(nk):
        cc = vee(M @ gens[i] @ Mi, group)
        out[i] = -sum(g_eta[j] * cc[j] for j in range(nk))
    return out


def unit_ascent(g, group: str):
    """lie.zig:243-255."""
    nk = n_keys(group)
    u = np.zeros(8)
    u[:nk] = g[:nk]
    n = math.sqrt(float(np.sum(u[:nk] ** 2)))
    return u if n < 1e-18 else u / n


def solve_linear(A, b):
    """Gauss–Jordan with partial pivoting (lie.zig:348-376); None if singular or non-finite."""
    M = np.array(A, np.float64, copy=True)
    x = np.array(b, np.float64, copy=True)
    n = len(x)
    for i in range(n):
        piv = i
        for r in range(i + 1, n):
            if abs(M[r, i]) > abs(M[piv, i]):
                piv = r
        if piv != i:
            M[[i, piv]] = M[[piv, i]]
            x[[i, piv]] = x[[piv, i]]
        d = M[i, i]
        if not math.isfinite(d) or abs(d) < 1e-300:
            return None
        for r in range(n):
            if r != i:
                f = M[r, i] / d
                if f == 0:
                    continue
                M[r, i:] -= f * M[i, i:]
                x[r] -= f * x[i]
    x = x / np.diag(M)
    return x if np.all(np.isfinite(x)) else None


def geo_amps(a0, ratio, n):
    """The ascent ladder a0, a0 r, a0 r², … (climb.zig:15-24)."""
    a = a0 if a0 > 0 else 0.08
    r = min(0.95, max(0.2, ratio if ratio > 0 else 0.5))
    k = max(2, min(12, n))
    out = []
    for _ in range(k):
        out.append(a)
        a *= r
    return out


def beats(new, old):
    """pair.zig:51-53: higher by more than 1e-6 of the old magnitude."""
    return new > old + max(1e-6 * abs(old), 1e-12)


# ── score coefficients (moments.zig:91-172) ──────────────────────────────────────────────
def _chol4(G):
    L = np.zeros((4, 4))
    for i in range(4):
        for j in range(i + 1):
            s = G[i, j] - np.dot(L[i, :j], L[j, :j])
            if i == j:
                if not s > 0:
                    return None
                L[i, i] = math.sqrt(s)
            else:
                L[i, j] = s / L[j, j]
    return L


def score_coefs(st, exact: bool = True, ridge: float = RIDGE) -> dict:
    """SMI score n Φ from the 45 sums and its gradient coefficients: W, M̂, the Gauss–Newton
    weight Q, the means and g = dS/d(sums) (moments.zig:91-172 scoreFromMoments)."""
    st = np.asarray(st, np.float64)
    n = st[0]
    zero = dict(score=0.0, n=n, W=np.zeros((4, 4)), Mh=np.zeros((4, 4)), Q=np.zeros((4, 4)), ma=np.zeros(4),
                mb=np.zeros(4), g=np.zeros(45))
    if not (n >= MIN_N):
        return zero
    ma, mb = st[1:5] / n, st[5:9] / n
    C = st[9:25].reshape(4, 4) / n - np.outer(ma, mb)
    W = C.copy()
    Mh = np.zeros((4, 4))
    Nh = np.zeros((4, 4))
    if not exact:
        phi = float(np.sum(C * C))
    else:
        GA = np.zeros((4, 4))
        GB = np.zeros((4, 4))
        for i in range(4):
            for j in range(4):
                k = tri_k(i, j)
                GA[i, j] = st[25 + k] / n - ma[i] * ma[j]
                GB[i, j] = st[35 + k] / n - mb[i] * mb[j]
        GA += ridge * max(np.trace(GA) / 4, 1e-12) * np.eye(4)
        GB += ridge * max(np.trace(GB) / 4, 1e-12) * np.eye(4)
        LA, LB = _chol4(GA), _chol4(GB)
        if LA is None or LB is None:
            return zero
        GAi = np.linalg.inv(LA @ LA.T)
        GBi = np.linalg.inv(LB @ LB.T)
        W = GAi @ C @ GBi
        phi = float(np.sum(W * C))
        Mh = W @ C.T @ GAi
        Nh = W.T @ C @ GBi
        Mh += np.trace(Mh) * ridge / 4 * np.eye(4)
        Nh += np.trace(Nh) * ridge / 4 * np.eye(4)
    score = n * phi
    if not math.isfinite(score):
        return zero
    g = np.zeros(45)
    for i in range(4):
        g[1 + i] = sum(-2 * W[i, j] * mb[j] + 2 * Mh[i, j] * ma[j] for j in range(4))
        g[5 + i] = sum(-2 * W[j, i] * ma[j] + 2 * Nh[i, j] * mb[j] for j in range(4))
        for j in range(4):
            g[9 + i * 4 + j] = 2 * W[i, j]
        for j in range(i, 4):
            f = 1.0 if i == j else 2.0
            g[25 + tri_k(i, j)] = -f * Mh[i, j]
            g[35 + tri_k(i, j)] = -f * Nh[i, j]
    g[0] = (score - float(np.dot(g[1:], st[1:]))) / n
    Q = Mh if exact else C @ C.T
    return dict(score=score, n=n, W=W, Mh=Mh, Q=Q, ma=ma, mb=mb, g=g)


# ── tangent gradient (smi.wgsl:1129-1272 tangent_grad; pose.zig:267-289 gradFrom) ──────────
def tangent_grad(feat_a, feat_b, H, co: dict, group: str, hess: bool = True):
    """dS/dξ (8) and the Gauss–Newton matrix Σ wt J_aᵀ Q J_a (8×8) of the forward direction at H,
    ξ the left increment in normalized coordinates of the destination (feat_b's image)."""
    _, hb, wb = feat_b.shape
    S = sample_pose(feat_a, hb, wb, H)
    ok = S["w"] > 0
    ok |= (S["dw"][0] != 0) | (S["dw"][1] != 0)
    va, gx, gy = S["va"][:, ok], S["gx"][:, ok], S["gy"][:, ok]
    vb = feat_b[:, ok]
    wt = S["w"][ok]
    dtx, dty = S["dw"][0][ok], S["dw"][1][ok]
    ph = S["ph"][:, ok]
    X = S["X"][:, ok]
    W, Mh, Q, ma, mb, g = co["W"], co["Mh"], co["Q"], co["ma"], co["mb"], co["g"]
    ta = va - ma[:, None]
    tb = vb - mb[:, None]
    ra = 2.0 * (W @ tb - Mh @ ta)
    # dS/dw per pixel: Σ_k g_k φ_k(va, vb) over the 45 monomials
    bw = g[0] + g[1:5] @ va + g[5:9] @ vb
    bw = bw + np.einsum("ij,ip,jp->p", g[9:25].reshape(4, 4), va, vb)
    for k, (i, j) in enumerate(TRI):
        bw = bw + g[25 + k] * va[i] * va[j] + g[35 + k] * vb[i] * vb[j]
    cx = wt * np.sum(ra * gx, axis=0) + bw * dtx
    cy = wt * np.sum(ra * gy, axis=0) + bw * dty
    Hi = np.linalg.inv(np.asarray(H, np.float64).reshape(3, 3))
    gens = generators(group, (wb, hb))
    iz2 = 1.0 / np.maximum(ph[2] * ph[2], 1e-12)
    px = np.zeros((8, ph.shape[1]))
    py = np.zeros((8, ph.shape[1]))
    for k in range(8):
        w = (-Hi @ gens[k]) @ X
        px[k] = (w[0] * ph[2] - ph[0] * w[2]) * iz2
        py[k] = (w[1] * ph[2] - ph[1] * w[2]) * iz2
    grad = px @ cx + py @ cy
    hm = None
    if hess:
        sxx = np.einsum("ip,ij,jp->p", gx, Q, gx)
        sxy = np.einsum("ip,ij,jp->p", gx, Q, gy)
        syx = np.einsum("ip,ij,jp->p", gy, Q, gx)
        syy = np.einsum("ip,ij,jp->p", gy, Q, gy)
        hm = (px * wt) @ (px * sxx + py * sxy).T + (py * wt) @ (px * syx + py * syy).T
    return grad, hm


def grad_dir(feat_a, feat_b, H, exact: bool, group: str, hess: bool):
    """One direction's score, overlap and tangent gradient (pose.zig:267-289 gradFrom)."""
    _, hb, wb = feat_b.shape
    S = sample_pose(feat_a, hb, wb, H)
    st = moments(S["va"], feat_b, S["w"])
    co = score_coefs(st, exact)
    nk = n_keys(group)
    out = dict(score=co["score"], n=co["n"], grad=np.zeros(8), hess=np.zeros((nk, nk)) if hess else None)
    if co["n"] >= MIN_N:
        g, h = tangent_grad(feat_a, feat_b, H, co, group, hess)
        out["grad"][:nk] = g[:nk]
        if hess:
            out["hess"] = h[:nk, :nk]
    return out


def sym(fwd, back, H, group: str, lw, lh, rw, rh):
    """Symmetric gradient: ½ (g_fwd + Pᵀ g_inv), ½ (A_fwd + Pᵀ A_inv P) (pair.zig:241-268)."""
    nk = n_keys(group)
    pulled = adjoint_grad(back["grad"], H, group, lw, lh, rw, rh)
    out = dict(score=0.5 * (fwd["score"] + back["score"]), n=fwd["n"], grad=fwd["grad"].copy(), hess=None)
    out["grad"][:nk] = 0.5 * (fwd["grad"][:nk] + pulled[:nk])
    if fwd["hess"] is not None and back["hess"] is not None:
        PT = np.zeros((nk, nk))
        for j in range(nk):
            e = np.zeros(8)
            e[j] = 1
            PT[:, j] = adjoint_grad(e, H, group, lw, lh, rw, rh)[:nk]
        out["hess"] = 0.5 * (fwd["hess"] + PT @ back["hess"] @ PT.T)
    return out


def pair_grad(feat_a, feat_b, H, exact=True, group="homography", symmetric=True, hess=True):
    """Pair.grad (pair.zig:270-306): the forward gradient, symmetrized with the inverse one."""
    H = np.asarray(H, np.float64).reshape(3, 3)
    f = grad_dir(feat_a, feat_b, H, exact, group, hess)
    if not symmetric:
        return f
    b = grad_dir(feat_b, feat_a, inv3(H), exact, group, hess)
    _, lh, lw = feat_a.shape
    _, rh, rw = feat_b.shape
    return sym(f, b, H, group, lw, lh, rw, rh)


def pair_score(feat_a, feat_b, H, exact=True, symmetric=True):
    """Pair.parts mean (pair.zig:162-184): non-finite direction scores count 0."""
    H = np.asarray(H, np.float64).reshape(3, 3)
    _, hb, wb = feat_b.shape
    _, ha, wa = feat_a.shape
    fin = lambda v: v if math.isfinite(v) else 0.0
    S = sample_pose(feat_a, hb, wb, H)
    fwd = fin(score_coefs(moments(S["va"], feat_b, S["w"]), exact)["score"])
    if not symmetric:
        return fwd
    Si = sample_pose(feat_b, ha, wa, inv3(H))
    inv = fin(score_coefs(moments(Si["va"], feat_a, Si["w"]), exact)["score"])
    return 0.5 * (fwd + inv)


def gn_batch(feat_a, feat_b, Hs, iters, group="homography", exact=True, symmetric=True, amps=()):
    """Gauss–Newton on many poses (search.zig:539-614 gnBatch): per iteration each active pose
    tries 5 lengths of the step (A + 1e-4 diag) δ = ½ g (capped at norm 0.25) and the ascent
    ladder along the unit gradient, and keeps any try that beats it; a pose that did not move
    leaves the batch. Returns [(H, score)] in input order."""
    _, lh, lw = feat_a.shape
    _, rh, rw = feat_b.shape
    nk = n_keys(group)
    cur = [[np.asarray(H, np.float64).reshape(3, 3), -math.inf] for H in Hs]
    act = list(range(len(Hs)))
    mults = (0.5, 1, 2, 4, 8)

    def move(H, v, a):
        x = np.zeros(8)
        x[:nk] = a * np.asarray(v)[:nk]
        return compose_n(H, x, group, rw, rh)

    it = 0
    while it < iters and act:
        tries = []
        for i in act:
            gi = pair_grad(feat_a, feat_b, cur[i][0], exact, group, symmetric, True)
            cur[i][1] = gi["score"]
            if amps:
                d = unit_ascent(gi["grad"], group)
                if math.sqrt(float(np.sum(d[:nk] ** 2))) > 1e-12:
                    tries += [(i, move(cur[i][0], d, a)) for a in amps]
            A = gi["hess"].copy()
            for k in range(nk):
                A[k, k] += 1e-4 * max(A[k, k], 1e-12)
            step = solve_linear(A, 0.5 * gi["grad"][:nk])
            if step is None:
                continue
            nrm = math.sqrt(float(np.sum(step ** 2)))
            for m in mults:
                s = m * (0.25 / (nrm * m) if nrm * m > 0.25 else 1)
                tries.append((i, move(cur[i][0], step, s)))
        if not tries:
            break
        moved = set()
        for i, H in tries:
            sc = pair_score(feat_a, feat_b, H, exact, symmetric)
            if beats(sc, cur[i][1]):
                cur[i] = [H, sc]
                moved.add(i)
        act = [i for i in act if i in moved]
        it += 1
    return [(H, s) for H, s in cur]


def climb(prep_a: dict, prep_b: dict, H0, n_iters: int, hop: bool = True, hop_fm: bool = False,
          group: str = "homography", exact: bool = True, symmetric: bool = True, calibrated: bool = False,
          cap: int = 1024, min_overlap: float = 0.0, smin: float = 0.2, smax: float = 5.0):
    """The climb (gclimb.zig:313-489 run; climb.wgsl:229-296 gn_cands, 320-376 hop_cand / rs_cand,
    415-468 climb_select, 840-858 iter_begin / iter_end), SMI family. Per iteration: a shift-map
    hop, a Gauss–Newton block of up to 3 steps (5 lengths each; the block ends at a step that beats
    nothing), and with hop_fm a roto-scale hop then another block. A candidate replaces the pose if
    it beats its score by max(1e-6 |s|, 1e-12) (the best such). Stop after 2 iterations without a
    move (1 when no hop is on) or n_iters. The maps' frame and roto-scale centre are fixed for
    batches of 8 iterations. Returns (H, score, iterations, moved)."""
    from oracle import smi_ref as O

    fa, fb = prep_a["feat"], prep_b["feat"]
    _, lh, lw = fa.shape
    _, rh, rw = fb.shape
    nk = n_keys(group)
    Hc = np.asarray(H0, np.float64).reshape(3, 3).copy()
    g = pair_grad(fa, fb, Hc, exact, group, symmetric, True)
    cur = g["score"]
    start = cur
    state = dict(moved=False, n_rec=0)

    def select(cands):
        nonlocal Hc, cur, g
        best, bs = None, cur
        for Hk in cands:
            sc = pair_score(fa, fb, Hk, exact, symmetric)
            if sc > cur + max(1e-6 * abs(cur), 1e-12) and (best is None or sc > bs):
                best, bs = Hk, sc
        if best is None:
            return False
        Hc, cur = best, bs
        state["moved"] = True
        state["n_rec"] += 1
        g = pair_grad(fa, fb, Hc, exact, group, symmetric, True)
        return True

    def gn_cands():
        A = g["hess"].copy()
        for k in range(nk):
            A[k, k] += 1e-4 * max(A[k, k], 1e-12)
        step = solve_linear(A, 0.5 * g["grad"][:nk])
        if step is None or not np.all(np.abs(step) < 1e30):
            return []
        sn = math.sqrt(float(np.sum(step ** 2)))
        out = []
        for mu in (0.5, 1.0, 2.0, 4.0, 8.0):
            nrm = sn * mu
            sc = mu * (0.25 / nrm if nrm > 0.25 else 1.0)
            x = np.zeros(8)
            x[:nk] = sc * step
            out.append(compose_n(Hc, x, group, rw, rh))
        return out

    def gn_block():
        for _ in range(3):
            if not select(gn_cands()):
                break

    it = stall = 0
    done = False
    while not done:
        fr = O.reg_frame(Hc, lw, lh, rw, rh)
        rs_c = O.moving_centroid(Hc, lw, lh)
        mirror = Hc[0, 0] * Hc[1, 1] - Hc[0, 1] * Hc[1, 0] > 0
        for _ in range(8):
            if done:
                break
            state["moved"] = False
            if hop:
                _, info = O.shift_map(prep_a, prep_b, Hc, exact, calibrated and exact, cap, min_overlap, frame=fr)
                T = np.array([[1, 0, -info["dx"]], [0, 1, -info["dy"]], [0, 0, 1.0]])
                select([T @ Hc])
            if hop or not hop_fm:
                gn_block()
            if hop_fm:
                fwd, p = O.rs_map_dir(prep_a, prep_b, Hc, rs_c, None, None, exact, calibrated and exact, cap, frame=fr)
                m = O.rs_finish(fwd, None, mirror, p["dlam"], smin, smax)
                n = p["n"]
                lag = O.peak_lag(O.argmax_first(m), n, n)
                dth = -(lag[0] / max(p["n_th"], 1)) * 2.0 * math.pi
                dsg = -lag[1] * p["dlam"]
                if not (abs(dth) < 2e-4 and abs(dsg) < 2e-4):
                    s_ = math.exp(dsg)
                    a, b, d, e = s_ * math.cos(dth), -s_ * math.sin(dth), s_ * math.sin(dth), s_ * math.cos(dth)
                    S = np.array([[a, b, rs_c[0] - a * rs_c[0] - b * rs_c[1]],
                                  [d, e, rs_c[1] - d * rs_c[0] - e * rs_c[1]], [0, 0, 1.0]])
                    select([project_group(S @ Hc, group)])
                gn_block()
            stall = 0 if state["moved"] else stall + 1
            it += 1
            if (not state["moved"] and (not (hop or hop_fm) or stall >= 2)) or it >= n_iters:
                done = True
    return Hc, cur, it, bool(state["n_rec"] > 0 and cur > start)
