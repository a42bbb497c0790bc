"""The learned rotation × scale prior: a small CNN that proposes where the sweep should look.

Each image of a pair becomes a log-polar map of its Fourier magnitude (zc_lp_features: 256²
downsample, rank, Sobel magnitude, FFT, 48 angles over [0, π) × 64 log radii over 4 .. 128
cycles): translation drops out, a rotation of the image is a periodic shift along the angle
axis and a scale change a shift along the radius axis. The two maps (moving, fixed) are the two
channels of the network's input, with their cross-correlation over angle and log-radius lags as a
third (a rotation and a scale change move its peak); six 3×3 convolutions (32 – 64 channels,
ReLU, three of stride 2) and a dense head give logits over 24 rotations (15° steps) × 9 scales
(0.8 – 1.25, log-spaced).
The magnitude spectrum cannot tell θ from θ + 180°, so the network spreads its belief over both;
the SMI scores of the proposed cells decide.

Every layer runs on this library's own HIP kernels (BF16 MFMA, FP32 accumulation) for both the
forward and the backward pass; PyTorch only holds the device buffers. Training pairs are warped
on the GPU by zc_warp_gray and labelled by the sweep's own argmax over the same 24 × 9 grid.
"""
from __future__ import annotations

import ctypes
import math
from pathlib import Path

import numpy as np

from . import _native

LAYERS = [(3, 32, 1), (32, 32, 2), (32, 64, 1), (64, 64, 2), (64, 64, 1), (64, 64, 2)]
T, R, C = 48, 64, 3            # angles, log radii, channels of the input map
S = 256                        # the images' downsampled side
R0, R1 = 4.0, 128.0            # radius range (cycles per 256 px)
N_TH, N_SG = 24, 9             # label grid: 24 rotations × 9 scales
THETAS = [-180.0 + 15.0 * k for k in range(N_TH)]
SIGMAS = [1.25 ** ((k - 4) / 4) for k in range(N_SG)]
WEIGHTS = Path(__file__).resolve().parent / "prior_weights.npz"


def _lib():
    from . import _load
    return _load()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: kernel launch failed (rc {rc})")


class Net:
    """Parameters (FP32 masters on the GPU), Adam state, forward and backward on HIP kernels."""

    def __init__(self, seed: int = 0):
        import torch

        rng = np.random.default_rng(seed)
        self.shapes = []
        h, w = T, R
        params = []
        for cin, cout, s in LAYERS:
            params.append(rng.standard_normal((9 * cin, cout)) * math.sqrt(2.0 / (9 * cin)))
            params.append(np.zeros(cout))
            h, w = h // s, w // s
        self.feat_hw = (h, w)
        k = h * w * LAYERS[-1][1]
        params.append(rng.standard_normal((k, N_TH * N_SG)) * math.sqrt(1.0 / k))
        params.append(np.zeros(N_TH * N_SG))
        self.p = [torch.tensor(a, dtype=torch.float32, device="cuda") for a in params]
        self.g = [torch.zeros_like(t) for t in self.p]
        self.m = [torch.zeros_like(t) for t in self.p]
        self.v = [torch.zeros_like(t) for t in self.p]
        self.t = 0

    # ── persistence ──
    def save(self, path=WEIGHTS, **meta):
        arrs = {f"p{i}": t.detach().cpu().numpy() for i, t in enumerate(self.p)}
        np.savez_compressed(path, **arrs, **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path=WEIGHTS):
        import torch

        net = cls()
        with np.load(path, allow_pickle=False) as z:
            for i in range(len(net.p)):
                net.p[i] = torch.tensor(z[f"p{i}"], dtype=torch.float32, device="cuda")
            net.meta = {k: z[k] for k in z.files if not k.startswith("p")}
        return net

    # ── passes ──
    def forward(self, x):
        """x: [n, T, R, C] float32 CUDA tensor. Returns (activations, logits [n, 216])."""
        import torch

        lib, st = _lib(), _stream()
        n = x.shape[0]
        acts = [x.contiguous()]
        h, w = T, R
        for i, (cin, cout, s) in enumerate(LAYERS):
            y = torch.empty((n, h // s, w // s, cout), dtype=torch.float32, device="cuda")
            _ok(lib.zc_cnn_conv_fwd(_p(acts[-1]), _p(self.p[2 * i]), _p(self.p[2 * i + 1]), _p(y), n, h, w, cin, cout,
                                    s, 1, st), "conv_fwd")
            acts.append(y)
            h, w = h // s, w // s
        logits = torch.empty((n, N_TH * N_SG), dtype=torch.float32, device="cuda")
        k = h * w * LAYERS[-1][1]
        _ok(lib.zc_cnn_fc_fwd(_p(acts[-1]), _p(self.p[-2]), _p(self.p[-1]), _p(logits), n, k, N_TH * N_SG, st),
            "fc_fwd")
        return acts, logits

    def backward(self, acts, dlogits):
        """Gradients into self.g from d(loss)/d(logits)."""
        import torch

        lib, st = _lib(), _stream()
        for g in self.g:
            g.zero_()
        n = dlogits.shape[0]
        f = acts[-1]
        k = f[0].numel()
        d = torch.empty_like(f)
        _ok(lib.zc_cnn_fc_bwd(_p(f), _p(self.p[-2]), _p(dlogits), _p(d), _p(self.g[-2]), _p(self.g[-1]), n, k,
                              N_TH * N_SG, st), "fc_bwd")
        for i in range(len(LAYERS) - 1, -1, -1):
            cin, cout, s = LAYERS[i]
            x = acts[i]
            h, w = x.shape[1], x.shape[2]
            dx = torch.empty_like(x) if i > 0 else None
            _ok(lib.zc_cnn_conv_bwd(_p(x), _p(self.p[2 * i]), _p(acts[i + 1]), _p(d),
                                    _p(dx) if dx is not None else None, _p(self.g[2 * i]), _p(self.g[2 * i + 1]), n, h,
                                    w, cin, cout, s, 1, st), "conv_bwd")
            d = dx

    def loss(self, logits, labels):
        """(per-sample losses, d(mean)/d(logits)) on the GPU."""
        import torch

        n = logits.shape[0]
        loss = torch.empty(n, dtype=torch.float32, device="cuda")
        d = torch.empty_like(logits)
        _ok(_lib().zc_cnn_xent(_p(logits), _p(labels), _p(loss), _p(d), n, N_TH * N_SG, _stream()), "xent")
        return loss, d

    def step(self, x, labels, lr=1e-3):
        """One Adam step on a batch; returns the mean loss (a device scalar)."""
        acts, z = self.forward(x)
        loss, d = self.loss(z, labels)
        self.backward(acts, d)
        self.t += 1
        lib, st = _lib(), _stream()
        for p, g, m, v in zip(self.p, self.g, self.m, self.v):
            _ok(lib.zc_cnn_adam(_p(p), _p(g), _p(m), _p(v), p.numel(), lr, 0.9, 0.999, 1e-8, self.t, st), "adam")
        return loss.mean()


def xcorr_channel(a, b):
    """The two log-polar maps' cross-correlation (a, b: [n, T, R] tensors): circular over the T
    angle lags, linear over the log-radius lags d = −R/2 .. R/2 − 1 (each normalized by its
    overlap T (R − |d|)), standardized per pair; [n, T, R] with lag d in column d + R/2. A
    rotation of the pair is a shift of its peak along the angle lags, a scale change one along the
    radius lags (the Fourier–Mellin correlation of the two spectra)."""
    import torch

    n, t, r = a.shape
    fa = torch.fft.fft2(torch.nn.functional.pad(a, (0, r)))
    fb = torch.fft.fft2(torch.nn.functional.pad(b, (0, r)))
    x = torch.fft.ifft2(fa * torch.conj(fb)).real                     # [n, T, 2R], lag d at d mod 2R
    d = torch.arange(-(r // 2), r - r // 2, device=a.device)
    c = x[:, :, d % (2 * r)] / (t * (r - d.abs()).to(x.dtype))
    c = c - c.mean(dim=(1, 2), keepdim=True)
    return c / c.flatten(1).std(dim=1).clamp_min(1e-12)[:, None, None]


def features(moving, fixed):
    """Network input [n, T, R, 3] from CUDA tensors moving, fixed [n, h, w] (gray, [0, 1]): the
    two log-polar maps (zc_lp_features) and their cross-correlation (xcorr_channel)."""
    import torch

    n, h, w = moving.shape
    # the two maps as a 2-channel tensor: zc_lp_features standardizes every channel once the
    # last one is written
    lp = torch.empty((n, T, R, 2), dtype=torch.float32, device="cuda")
    work = torch.empty(n * S * S * 3, dtype=torch.float32, device="cuda")
    lib, st = _lib(), _stream()
    for ch, img in ((0, moving), (1, fixed)):
        hh, ww = img.shape[1], img.shape[2]
        _ok(lib.zc_lp_features(_p(img.contiguous()), n, ww, hh, S, T, R, 2, ch, R0, R1, _p(lp), _p(work), st),
            "lp_features")
    out = torch.empty((n, T, R, C), dtype=torch.float32, device="cuda")
    out[..., :2] = lp
    out[..., 2] = xcorr_channel(lp[..., 0], lp[..., 1])
    return out


def propose(net, moving, fixed, k=8):
    """The k most likely (rotation, scale) classes of each pair, best first, each as its two cells
    (θ, θ + 180°): the magnitude spectrum cannot tell them apart, so a class's belief is the
    log-sum of its two cells' logits and the SMI scores of both cells decide."""
    import torch

    _, z = net.forward(features(moving, fixed))
    half = N_TH // 2
    zc = torch.logsumexp(z.view(-1, 2, half, N_SG), dim=1).view(-1, half * N_SG)
    top = torch.topk(zc, k, dim=1).indices.cpu().numpy()
    out = []
    for row in top:
        cells = []
        for c in row:
            t, s = int(c) // N_SG, int(c) % N_SG
            cells += [(t, s), (t + half, s)]
        out.append(cells)
    return out


# the sweep grid the network's classes are labelled on: 24 rotations over the circle, 9 scales
# log-spaced over 0.8 – 1.25
TRAIN_GRID = {"sw_rot_range": False, "sw_rot": 180.0, "sw_nth": N_TH, "sw_smin": 0.8, "sw_smax": 1.25, "sw_ns": N_SG}


def sweep_axes(settings):
    """The rotation (degrees) and scale axes of a sweep grid from sweep settings, with the clamps
    and spacing of the engine's sweep (search.zig:89-116 sweepGrid): the full circle without
    repeating ±180°, a partial range with both ends; scales log-spaced. Returns (thetas, sigmas,
    wrap)."""
    cl = lambda v, lo, hi: min(max(v, lo), hi)   # noqa: E731
    rot = cl(float(settings["sw_rot"]), 0.0, 180.0)
    if settings["sw_rot_range"]:
        lo, hi = cl(float(settings["sw_rot_lo"]), -360.0, 360.0), cl(float(settings["sw_rot_hi"]), -360.0, 360.0)
    else:
        lo, hi = -rot, rot
    if hi < lo:
        hi += 360.0
    hi = min(hi, lo + 360.0)
    n = cl(int(settings["sw_nth"]), 1, 720)
    span = abs(hi - lo)
    wrap = span >= 360.0 - 1e-9
    if wrap:
        thetas = [-180.0 + 360.0 * i / n for i in range(n)]
    elif n == 1:
        thetas = [lo + span / 2]
    else:
        thetas = [lo + span * i / (n - 1) for i in range(n)]
    smin, smax = cl(float(settings["sw_smin"]), 0.05, 1.0), cl(float(settings["sw_smax"]), 1.0, 20.0)
    ns = cl(int(settings["sw_ns"]), 1, 129)
    a, b = math.log(max(1e-3, min(smin, smax))), math.log(max(1e-3, max(smin, smax)))
    sigmas = [math.exp((a + b) / 2 if ns == 1 else a + (b - a) * k / (ns - 1)) for k in range(ns)]
    return thetas, sigmas, wrap


def fine_cells(cells, settings=None, margin=0):
    """The (rotation index, scale index) cells of a sweep grid (settings: the registrar's sweep
    settings; None: 72 rotations × the training grid's 9 scales) inside the network's proposed
    cells: for a class (it, is) every rotation within ±7.5° of −180° + 15°·it, widened by `margin`
    of the grid's rotation steps on each side (a pose near a class's edge), and every scale within
    half a class step (a factor 1.25^(1/8)) of 1.25^((is − 4)/4). Cells keep their first-seen order."""
    if settings is None:
        settings = {**TRAIN_GRID, "sw_nth": 72}
    thetas, sigmas, wrap = sweep_axes(settings)
    step = (360.0 / len(thetas)) if wrap else (abs(thetas[-1] - thetas[0]) / max(len(thetas) - 1, 1))
    half_class = math.log(1.25) / 8
    out = []
    for it, is_ in cells:
        centre = -180.0 + 15.0 * it
        reach = 7.5 + margin * step + 1e-9
        rot_idx = [j for j, t in enumerate(thetas) if abs((t - centre + 180.0) % 360.0 - 180.0) <= reach]
        # rotations in order around the class centre, from its lower side
        rot_idx.sort(key=lambda j: (thetas[j] - centre + 180.0) % 360.0 - 180.0)
        ls = math.log(1.25) * (is_ - 4) / 4
        sc_idx = [k for k, sg in enumerate(sigmas) if abs(math.log(sg) - ls) <= half_class + 1e-9]
        if not sc_idx:   # no scale of the grid inside the class: its nearest
            sc_idx = [min(range(len(sigmas)), key=lambda k: abs(math.log(sigmas[k]) - ls))]
        for j in rot_idx:
            for k in sc_idx:
                if (j, k) not in out:
                    out.append((j, k))
    return out


def check_score(reg, fine):
    """How well the cells explain the pair: the sweep over them on the 128² canvas, then the
    forward SMI per overlapping pixel (Φ = S / n, the mean squared whitened correlation, between
    0 and 4) at the best peak's pose. Unlike S, Φ does not grow with the overlap, so one
    threshold serves pairs of any size and overlap."""
    s = reg.settings
    keep = {"sw_full": s["sw_full"], "sw_grid": s["sw_grid"]}
    reg.configure(sw_full=False, sw_grid=128)
    try:
        _, peaks, _ = reg.sweep(cells=fine)
    finally:
        reg.configure(**keep)
    if len(peaks) == 0:
        return 0.0
    H = peaks[int(np.argmax(peaks["score"]))]["H"]
    n = float(reg.moments(H)[0])
    return float(reg.score(H).fwd) / n if n > 0 else 0.0


def calibrate(reg, net, moving, fixed, poses, thetas, sigmas, q=0.05):
    """The fallback threshold: the q-quantile of check_score over held-out pairs at the cells that
    hold their true poses (moving, fixed: CUDA tensors [n, h, w]; poses, thetas, sigmas: their
    candidate-convention truths)."""
    vals = []
    for i in range(len(poses)):
        reg.set_images(moving[i].cpu().numpy(), fixed[i].cpu().numpy(), preprocess=False)
        t, s_ = truth_cell(thetas[i], sigmas[i])
        vals.append(check_score(reg, fine_cells([(t, s_)], reg.settings)))
    return float(np.quantile(vals, q))


def learned_search(reg, net, moving, fixed, k=8, threshold=0.0, margin=1):
    """The search with the prior: the network's top-k classes (2k cells: θ and θ + 180°), the sweep
    over the grid cells inside them and `margin` fine rotations beyond (the registrar's sweep
    settings: 72 rotations × 9 scales), NMS and Gauss–Newton. When the proposals' check_score falls below 0.8 × threshold (calibrate), the
    full search runs from the same start instead. moving, fixed: CUDA tensors [h, w] of the
    registrar's images.

    Returns (Search, info dict: proposed cells, fine cells, check score, fell_back)."""
    H0 = reg.pose
    cells = propose(net, moving[None], fixed[None], k)[0]
    fine = fine_cells(cells, reg.settings, margin)
    score = check_score(reg, fine)
    fell_back = score < 0.8 * threshold
    reg.pose = H0
    res = reg.search() if fell_back else reg.search_cells(fine)[0]
    return res, dict(cells=cells, fine=fine, check_score=score, fell_back=fell_back)


# ── training data: pairs the method warps and labels ──────────────────────────────────────
def _about(c, A):
    T_ = np.eye(3)
    T_[:2, :2] = A
    return np.array([[1, 0, c[0]], [0, 1, c[1]], [0, 0, 1.0]]) @ T_ @ np.array([[1, 0, -c[0]], [0, 1, -c[1]], [0, 0, 1]])


def random_poses(rng, n, size, shift=16.0):
    """n candidate-convention poses (moving → fixed) about the centre of size × size images:
    rotation uniform over the circle, scale log-uniform over 0.8 – 1.25, a shift of up to
    `shift` px; with the remap of each pair (1: 1 − x², 2: √x, 3: sin πx)."""
    c = ((size + 1) / 2.0, (size + 1) / 2.0)
    th = rng.uniform(-180.0, 180.0, n)
    sg = np.exp(rng.uniform(math.log(0.8), math.log(1.25), n))
    t = rng.uniform(-shift, shift, (n, 2))
    Hs = []
    for i in range(n):
        a = math.radians(th[i])
        A = sg[i] * np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
        Hs.append(np.array([[1, 0, t[i, 0]], [0, 1, t[i, 1]], [0, 0, 1.0]]) @ _about(c, A))
    return np.stack(Hs), th, sg, rng.integers(1, 4, n)


def truth_cell(th, sg):
    """The label grid cell nearest a rotation (degrees) and scale."""
    it = int(round((((th + 180.0) % 360.0)) / 15.0)) % N_TH
    is_ = int(min(max(round(4 * math.log(sg) / math.log(1.25) + 4), 0), N_SG - 1))
    return it, is_


def gpu_textures(n, size, seed, n_blobs=20):
    """n training textures on the GPU, the recipe of synth._texture (Gaussian-filtered noise,
    sigma 4, periodic; unit spread; n_blobs Gaussian blobs; scaled to [0, 1]), each its own draw."""
    import torch

    g = torch.Generator(device="cuda").manual_seed(int(seed))
    x = torch.randn(n, size, size, device="cuda", generator=g)
    f = torch.fft.fftfreq(size, device="cuda", dtype=torch.float32)
    k = torch.exp(-2.0 * math.pi ** 2 * 16.0 * (f[:, None] ** 2 + f[None, :] ** 2))
    x = torch.fft.ifft2(torch.fft.fft2(x) * k).real
    x = x / x.flatten(1).std(dim=1).clamp_min(1e-12)[:, None, None]
    yy = torch.arange(size, device="cuda", dtype=torch.float32)[None, :, None]
    xx = torch.arange(size, device="cuda", dtype=torch.float32)[None, None, :]
    u = lambda lo, hi: lo + (hi - lo) * torch.rand(n, 1, 1, device="cuda", generator=g)
    for _ in range(n_blobs):
        cy, cx = u(0.0, size), u(0.0, size)
        sg = u(2.0, max(3.0, size / 12.0))
        amp = u(-3.0, 3.0)
        x = x + amp * torch.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sg * sg))
    x = x - x.flatten(1).min(dim=1).values[:, None, None]
    x = x / x.flatten(1).max(dim=1).values.clamp_min(1e-12)[:, None, None]
    return x.contiguous()


def make_pairs(textures, idx, Hs, remaps):
    """moving[i] = remap(fixed[i] at H_i X): fixed = textures[idx] (CUDA [k, s, s])."""
    import torch

    fixed = textures[torch.as_tensor(idx, device="cuda")].contiguous()
    n, s, _ = fixed.shape
    His = torch.tensor(np.asarray(Hs).reshape(n, 9), dtype=torch.float64, device="cuda")
    rm = torch.tensor(np.asarray(remaps, np.int32), device="cuda")
    moving = torch.empty_like(fixed)
    _ok(_lib().zc_warp_gray(_p(fixed), n, s, s, _p(His), _p(rm), _p(moving), _stream()), "warp_gray")
    return moving, fixed


def sweep_labels(reg, moving, fixed, grid=128):
    """The sweep's argmax cell (24 × 9 grid, on a grid² canvas) for each pair: the label of the
    pair, and the cell's best SMI score divided by the fixed image's area."""
    import torch

    torch.cuda.synchronize()
    reg.configure(sw_nth=N_TH, sw_smin=0.8, sw_smax=1.25, sw_ns=N_SG, sw_grid=grid, sw_rot_range=False, sw_rot=180,
                  sw_full=False)
    labels, best_norm = [], []
    n, h, w = fixed.shape
    for i in range(n):
        reg.set_image_device(0, _p(moving[i]), w, h)
        reg.set_image_device(1, _p(fixed[i]), w, h)
        _, _, best = reg.sweep()
        k = int(np.argmax(best))
        labels.append(k)
        best_norm.append(float(best[k]) / (w * h))
    return np.array(labels, np.int32), np.array(best_norm)


CHUNK = 500   # pairs generated (and their textures drawn) together


def _pair_chunk(seed, size, s0, s1, Hs, rm):
    """Pairs s0 .. s1 of a training set (s0 a multiple of CHUNK): each its own texture."""
    textures = gpu_textures(s1 - s0, size, seed * 1000003 + s0)
    return make_pairs(textures, np.arange(s1 - s0), Hs[s0:s1], rm[s0:s1])


def label_pairs(total, size=1024, seed=1, start=0, stop=None, grid=128, log=print):
    """Sweep labels of pairs [start, stop) of a training set of `total` pairs (start a multiple
    of CHUNK). Returns a dict of arrays: labels, best-score norms, whether the label is the nearest
    true cell, and the range."""
    import time

    import zcmir_amd

    stop = total if stop is None else stop
    if start % CHUNK:
        raise ValueError(f"start must be a multiple of {CHUNK}")
    rng = np.random.default_rng(seed)
    Hs, th, sg, rm = random_poses(rng, total, size)
    reg = zcmir_amd.Registrar(half=False)
    labels, norms, agree = [], [], []
    t0 = time.time()
    for s0 in range(start, stop, CHUNK):
        s1 = min(stop, s0 + CHUNK)
        mv, fx = _pair_chunk(seed, size, s0, s1, Hs, rm)
        lab, bn = sweep_labels(reg, mv, fx, grid)
        labels.append(lab)
        norms.append(bn)
        agree += [truth_cell(th[i], sg[i]) == (lab[i - s0] // N_SG, lab[i - s0] % N_SG) for i in range(s0, s1)]
        log(f"labelled {s1 - start}/{stop - start}  {time.time() - t0:.1f} s  label = nearest true cell: "
            f"{np.mean(agree):.3f}")
    reg.close()
    return {"labels": np.concatenate(labels), "norms": np.concatenate(norms), "agree": np.array(agree),
            "range": np.array([start, stop]), "total": np.array(total), "size": np.array(size),
            "seed": np.array(seed), "grid": np.array(grid)}


def train(n_pairs=16000, n_val=1000, size=512, epochs=12, batch=128, lr=1e-3, seed=1, log=print, out=WEIGHTS,
          labels=None, grid=128, budget_s=None):
    """Generate pairs on the GPU, label them by the sweep (or take `labels`: label_pairs' output
    covering every pair), train the network, save WEIGHTS. Returns metrics. With `budget_s`, the
    epochs are cut to those that fit the seconds left after the features (timed on the first 16
    steps, which count towards the first epoch)."""
    import time

    import torch

    t0 = time.time()
    total = n_pairs + n_val
    if labels is None:
        labels = label_pairs(total, size, seed, 0, total, grid, log)
    if len(labels["labels"]) != total:
        raise ValueError(f"labels for {len(labels['labels'])} pairs, the set has {total}")
    rng = np.random.default_rng(seed)
    Hs, th, sg, rm = random_poses(rng, total, size)
    X = []
    for s0 in range(0, total, CHUNK):
        s1 = min(total, s0 + CHUNK)
        mv, fx = _pair_chunk(seed, size, s0, s1, Hs, rm)
        X.append(features(mv, fx))
        del mv, fx
        if (s0 // CHUNK) % 10 == 0:
            log(f"features {s1}/{total}  {time.time() - t0:.1f} s")
    X = torch.cat(X)
    Y = torch.tensor(labels["labels"], device="cuda")
    norms = labels["norms"]
    agree = labels["agree"]
    net = Net(seed)
    xtr, ytr, xva, yva = X[:n_pairs], Y[:n_pairs], X[n_pairs:], Y[n_pairs:]
    steps = 0
    per_epoch = n_pairs // batch
    ep = 0
    best = None
    while ep < epochs:
        perm = torch.randperm(n_pairs, device="cuda")
        run = 0.0
        for b0 in range(0, n_pairs - batch + 1, batch):
            sel = perm[b0:b0 + batch]
            lr_t = lr * 0.5 * (1 + math.cos(math.pi * steps / (epochs * per_epoch)))
            if budget_s is not None and steps == 16:
                torch.cuda.synchronize()
                t_step = (time.time() - t_fit) / 16
                fit = int((budget_s - (time.time() - t0)) / (t_step * per_epoch * 1.05))
                if fit < epochs:
                    epochs = max(1, fit)
                    log(f"{t_step * 1e3:.1f} ms per step: {epochs} epochs fit the budget")
            if steps == 0:
                t_fit = time.time()
            run += float(net.step(xtr[sel], ytr[sel], lr_t))
            steps += 1
        acc = evaluate(net, xva, yva)
        log(f"epoch {ep + 1}: loss {run / per_epoch:.3f}  val top1 {acc['top1']:.3f} top8 {acc['top8']:.3f} "
            f"top8±1θ {acc['top8_near']:.3f}  {time.time() - t0:.1f} s")
        ep += 1
        # early stopping on the held-out pairs: keep the weights of the best epoch
        if best is None or acc["top8_near"] > best[0]:
            best = (acc["top8_near"], ep, [t.detach().clone() for t in net.p])
    with torch.no_grad():
        for t, b in zip(net.p, best[2]):
            t.copy_(b)
    log(f"kept epoch {best[1]} (val top8±1θ {best[0]:.3f})")
    thr = float(np.quantile(norms[n_pairs:], 0.05))
    net.save(out, threshold=thr, label_agreement=float(np.mean(agree)), n_pairs=n_pairs, size=size,
             label_grid=int(np.asarray(labels.get("grid", grid))), epochs=epochs, kept_epoch=best[1])
    return dict(evaluate(net, xva, yva), threshold=thr, label_agreement=float(np.mean(agree)),
                seconds=time.time() - t0)


def evaluate(net, x, y, k=8):
    """Top-1, top-k and top-k-within-one-rotation-step hit rates (θ and θ + 180° count as one)."""
    import torch

    with torch.no_grad():
        _, z = net.forward(x)
    top = torch.topk(z, k, dim=1).indices.cpu().numpy()
    lab = y.cpu().numpy()

    def same(c, l, near):
        dt = abs(c // N_SG - l // N_SG) % 12
        dt = min(dt, 12 - dt)
        return c % N_SG == l % N_SG and dt <= near

    t1 = np.mean([same(top[i, 0], lab[i], 0) for i in range(len(lab))])
    tk = np.mean([any(same(c, lab[i], 0) for c in top[i]) for i in range(len(lab))])
    tn = np.mean([any(same(c, lab[i], 1) for c in top[i]) for i in range(len(lab))])
    return dict(top1=float(t1), top8=float(tk), top8_near=float(tn))


def main():
    import argparse
    import json

    ap = argparse.ArgumentParser(description="train the learned rotation × scale prior on the GPU")
    ap.add_argument("--pairs", type=int, default=16000)
    ap.add_argument("--val", type=int, default=1000)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--grid", type=int, default=128, help="the labelling sweep's canvas cells per side")
    ap.add_argument("--out", default=str(WEIGHTS))
    ap.add_argument("--label-range", type=int, nargs=2, metavar=("START", "STOP"),
                    help="only label pairs [START, STOP) of the set and save them to --labels-out")
    ap.add_argument("--labels-out", default="labels.npz")
    ap.add_argument("--labels", nargs="*", help="label files (label_pairs' ranges) that together cover the set")
    ap.add_argument("--budget-s", type=float, help="cut the epochs to what fits this many seconds")
    a = ap.parse_args()
    log = lambda s: print(s, flush=True)   # noqa: E731
    total = a.pairs + a.val
    if a.label_range:
        lab = label_pairs(total, a.size, 1, a.label_range[0], a.label_range[1], a.grid, log)
        np.savez(a.labels_out, **lab)
        print(json.dumps({"labelled": int(len(lab["labels"])), "agreement": float(np.mean(lab["agree"]))}), flush=True)
        return
    labels = None
    if a.labels:
        parts = sorted((dict(np.load(f, allow_pickle=False)) for f in a.labels), key=lambda d: int(d["range"][0]))
        labels = {k: np.concatenate([p[k] for p in parts]) for k in ("labels", "norms", "agree")}
        labels["grid"] = parts[0]["grid"]
    r = train(a.pairs, a.val, a.size, a.epochs, a.batch, a.lr, log=log, out=Path(a.out), labels=labels, grid=a.grid,
              budget_s=a.budget_s)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
