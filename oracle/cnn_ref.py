"""NumPy oracle of the learned rotation × scale prior's network and input features.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import this module, as the checker; the product path never does.

The network has no counterpart in the reference (it is this project's own addition); this module
states its arithmetic independently of the HIP kernels (zcmir_amd/csrc/cnn.hip): 3×3 convolutions
(periodic along H, zero-padded along W, stride 1 or 2) with bias and ReLU, a dense head, softmax
cross-entropy, and their gradients. ``bf16=True`` rounds each product's operands to bfloat16
(round to nearest even), as the MFMA kernels do. The gradients are pinned by central differences
in float64 (tests/test_cnn_oracle.py).
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """float32 → bfloat16 (round to nearest even) → float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _q(a, on):
    return bf16(a).astype(np.float64) if on else np.asarray(a, np.float64)


def im2col(x, stride):
    """x [n, h, w, c] → [n·ho·wo, 9·c] (tap-major: (ky, kx), then channel)."""
    n, h, w, c = x.shape
    ho, wo = h // stride, w // stride
    cols = np.zeros((n, ho, wo, 9, c), x.dtype)
    for ky in range(3):
        yi = (np.arange(ho) * stride + ky - 1) % h
        for kx in range(3):
            xi = np.arange(wo) * stride + kx - 1
            ok = (xi >= 0) & (xi < w)
            patch = x[:, yi][:, :, np.clip(xi, 0, w - 1)]
            patch[:, :, ~ok] = 0
            cols[:, :, :, ky * 3 + kx] = patch
    return cols.reshape(n * ho * wo, 9 * c)


def col2im(dcols, shape, stride):
    """Adjoint of im2col: [n·ho·wo, 9·c] → [n, h, w, c] (sums overlapping taps)."""
    n, h, w, c = shape
    ho, wo = h // stride, w // stride
    d = dcols.reshape(n, ho, wo, 9, c)
    dx = np.zeros(shape, dcols.dtype)
    for ky in range(3):
        yi = (np.arange(ho) * stride + ky - 1) % h
        for kx in range(3):
            xi = np.arange(wo) * stride + kx - 1
            ok = (xi >= 0) & (xi < w)
            for a, y in enumerate(yi):   # distinct columns within one row and tap
                dx[:, y, xi[ok]] += d[:, a, ok, ky * 3 + kx]
    return dx


def conv_fwd(x, W, b, stride, relu=True, bf=False):
    n, h, w, c = x.shape
    y = _q(im2col(x, stride), bf) @ _q(W, bf) + b
    if relu:
        y = np.maximum(y, 0)
    return y.reshape(n, h // stride, w // stride, W.shape[1])


def conv_bwd(x, W, y, dy, stride, relu=True, bf=False):
    """(dx, dW, db) of conv_fwd, through the ReLU via its output y."""
    cout = W.shape[1]
    d = (dy * (y > 0) if relu else dy).reshape(-1, cout)
    cols = im2col(x, stride)
    dW = _q(cols, bf).T @ _q(d, bf)
    db = d.sum(0)
    dx = col2im(_q(d, bf) @ _q(W, bf).T, x.shape, stride)
    return dx, dW, db


def fc_fwd(f, W, b, bf=False):
    return _q(f.reshape(f.shape[0], -1), bf) @ _q(W, bf) + b


def fc_bwd(f, W, dout, bf=False):
    f2 = f.reshape(f.shape[0], -1)
    return (_q(dout, bf) @ _q(W, bf).T).reshape(f.shape), _q(f2, bf).T @ _q(dout, bf), dout.sum(0)


def xent(logits, labels):
    """Mean softmax cross-entropy and d(mean)/d(logits)."""
    z = logits - logits.max(1, keepdims=True)
    p = np.exp(z)
    p /= p.sum(1, keepdims=True)
    n = len(labels)
    loss = -np.log(p[np.arange(n), labels])
    d = p.copy()
    d[np.arange(n), labels] -= 1
    return loss, d / n


def net_forward(x, params, layers, bf=False):
    """Activations [x, a1, …, a6] and the logits."""
    acts = [x]
    for (cin, cout, s), (W, b) in zip(layers, params[:-1]):
        acts.append(conv_fwd(acts[-1], W, b, s, True, bf))
    Wf, bfc = params[-1]
    return acts, fc_fwd(acts[-1], Wf, bfc, bf)


def net_backward(acts, params, layers, dlogits, bf=False):
    """Gradients [(dW, db)] per layer (convs then the head) and d(input)."""
    Wf, _ = params[-1]
    df, dWf, dbf = fc_bwd(acts[-1], Wf, dlogits, bf)
    grads = [(dWf, dbf)]
    d = df
    for i in range(len(layers) - 1, -1, -1):
        W, _ = params[i]
        d, dW, db = conv_bwd(acts[i], W, acts[i + 1], d, layers[i][2], True, bf)
        grads.insert(0, (dW, db))
    return grads, d


# ── input features (zc_lp_features) ──────────────────────────────────────────────────────
def lp_features(img, S=256, T=48, R=64, r0=4.0, r1=128.0):
    """[T, R] map of one gray image: box downsample to S², rank (256-bin histogram equalization
    of [0, 1]), Sobel magnitude (clamped borders) × Hann window, 2-D FFT, log(1 + |F|) bilinear
    at angles π t / T and radii r0 (r1 / r0)^(j / R), standardized."""
    img = np.asarray(img, np.float64)
    h, w = img.shape
    d = np.zeros((S, S))
    for y in range(S):
        y0, y1 = y * h // S, max(y * h // S + 1, (y + 1) * h // S)
        for x in range(S):
            x0, x1 = x * w // S, max(x * w // S + 1, (x + 1) * w // S)
            d[y, x] = img[y0:y1, x0:x1].mean()
    b = np.clip((d * 256).astype(np.int64), 0, 255)
    hist = np.bincount(b.ravel(), minlength=256)
    cdf = np.concatenate([[0], np.cumsum(hist)[:-1]])
    r = (cdf[b] + 0.5 * hist[b]) / (S * S) * 2 - 1
    p = np.pad(r, 1, mode="edge")
    gx = (p[:-2, 2:] + 2 * p[1:-1, 2:] + p[2:, 2:]) - (p[:-2, :-2] + 2 * p[1:-1, :-2] + p[2:, :-2])
    gy = (p[2:, :-2] + 2 * p[2:, 1:-1] + p[2:, 2:]) - (p[:-2, :-2] + 2 * p[:-2, 1:-1] + p[:-2, 2:])
    win = 0.5 - 0.5 * np.cos(2 * np.pi * (np.arange(S) + 0.5) / S)
    g = np.hypot(gx, gy) * win[:, None] * win[None, :]
    F = np.abs(np.fft.fft2(g))
    th = np.pi * np.arange(T) / T
    rr = r0 * (r1 / r0) ** (np.arange(R) / R)
    u = rr[None, :] * np.cos(th)[:, None]
    v = rr[None, :] * np.sin(th)[:, None]
    iu, iv = np.floor(u).astype(np.int64), np.floor(v).astype(np.int64)
    au, av = u - iu, v - iv
    m = lambda vv, uu: F[vv % S, uu % S]
    val = (m(iv, iu) * (1 - au) + m(iv, iu + 1) * au) * (1 - av) + (m(iv + 1, iu) * (1 - au) + m(iv + 1, iu + 1) * au) * av
    out = np.log1p(val)
    return (out - out.mean()) / np.sqrt(max(out.var(), 1e-12))
