"""Plain functions over one engine shared by the process (the functional surface of the reference
package: python/zcmir/api.py:291-580 register, refine, score, shift_map, rs_map, search,
register_stack).

Every function takes engine settings as keyword arguments and starts from the defaults, so one
call never leaks settings into the next. Calls are serialized by a process-wide lock (one engine,
one device queue). Images are numpy arrays (uint8 gray / RGB / RGBA, uint16 gray, float gray in
[0, 1]) or file paths (read with Pillow when it is installed).

Keypoint detection is not part of this library: ``register`` needs a starting pose (``init``) and
refines it, ``search`` finds a pose from nothing with the rotation × scale × shift sweep, and
``register_stack`` registers all slices of a stack together on the GPU (see each function).
"""
from __future__ import annotations

import difflib
import json
import math
import threading
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import Registrar
from ._handle import EngineError

_LOCK = threading.RLock()


class _Shared:
    """The process's engine, its default settings and the images it holds (side → key)."""
    reg: Registrar | None = None
    defaults: dict = {}
    held: dict = {}


# settings that change a side's work image: an image is uploaded again when any of them changes
_IMAGE_KEYS = ("clahe", "clahe_grid", "clahe_bins", "band", "bp_fine", "bp_coarse", "invert", "half")
# reference names a caller may use for engine settings
_ALIASES = {"model": "group"}


def _engine() -> Registrar:
    if _Shared.reg is None:
        _Shared.reg = Registrar()
        _Shared.defaults = dict(_Shared.reg.settings)
        _Shared.held = {}
    return _Shared.reg


def close():
    """Release the shared engine and its device memory; the next call makes a new one."""
    with _LOCK:
        if _Shared.reg is not None:
            _Shared.reg.close()
        _Shared.reg = None
        _Shared.held = {}


def settings():
    """Every engine setting with its default value."""
    with _LOCK:
        _engine()
        return dict(_Shared.defaults)


def _settings_for(model="homography", spline=False, **kw) -> dict:
    """The defaults overridden by the caller's keywords; unknown names raise TypeError with the
    closest known names."""
    if model not in ("affine", "homography"):
        raise ValueError(f'model: "affine" or "homography", not {model!r}')
    if spline:
        raise EngineError("spline: Unsupported (the B-spline stage is not part of this library)")
    out = dict(_Shared.defaults)
    out["group"] = model
    for name, value in kw.items():
        key = _ALIASES.get(name, name)
        if key not in out:
            near = difflib.get_close_matches(key, list(out), n=3)
            hint = f"; did you mean {', '.join(near)}?" if near else " (zcmir_amd.settings() lists them)"
            raise TypeError(f"unknown setting {name!r}{hint}")
        out[key] = value
    return out


def _read(x) -> np.ndarray:
    if isinstance(x, (str, Path)):
        try:
            from PIL import Image
        except ImportError as e:   # pragma: no cover - Pillow is optional
            raise ImportError("reading image files needs Pillow; pass numpy arrays instead") from e
        with Image.open(x) as im:
            if im.mode.startswith("I"):
                return np.asarray(im).astype(np.uint16)
            return np.asarray(im if im.mode in ("L", "RGB", "RGBA") else im.convert("RGBA"))
    a = np.asarray(x)
    if a.ndim not in (2, 3):
        raise ValueError(f"an image is (h, w) or (h, w, channels), not shape {a.shape}")
    return a


def _image_key(a: np.ndarray, preprocess: bool, s: dict):
    """Identifies an upload: the array's identity, shape, dtype, a content sample and the settings
    that shape the work image."""
    flat = a.reshape(-1)
    sample = flat[:: max(1, flat.size // 2048)]
    return (id(a), a.shape, a.dtype.str, hash(sample.tobytes()), preprocess, tuple(s.get(k) for k in _IMAGE_KEYS))


def _with_pair(moving, fixed, preprocess, kw):
    """The engine configured for this call with both images in it; returns (engine, settings,
    moving array, fixed array)."""
    reg = _engine()
    s = _settings_for(**kw)
    reg.configure(**s)
    m, f = _read(moving), _read(fixed)
    for side, a in ((0, m), (1, f)):
        key = _image_key(a, preprocess, s)
        held = _Shared.held.get(side)
        if held is None or held[0] != key or held[1] is not a:
            (reg.set_moving if side == 0 else reg.set_fixed)(a, preprocess=preprocess)
            _Shared.held[side] = (key, a)
    return reg, s, m, f


# ── geometry on the host (results only) ──────────────────────────────────────────────────

def _apply(H, pts):
    """Points (n, 2) through a 3×3 pose."""
    q = np.c_[pts, np.ones(len(pts))] @ np.asarray(H, np.float64).T
    return q[:, :2] / q[:, 2:3]


def _corners(shape):
    h, w = shape[:2]
    return np.array([[1.0, 1.0], [w, 1.0], [w, h], [1.0, h]])


def _bilinear(img: np.ndarray, x: np.ndarray, y: np.ndarray, fill):
    """img sampled at 0-based (x, y) (pixel centres at integers), clamped taps, `fill` outside
    the half-pixel border; returns (values (..., channels), inside mask)."""
    h, w = img.shape[:2]
    src = img.reshape(h, w, -1).astype(np.float32)
    inside = (x > -0.5) & (y > -0.5) & (x < w - 0.5) & (y < h - 0.5)
    xc, yc = np.clip(x, 0, w - 1), np.clip(y, 0, h - 1)
    x0 = np.minimum(np.floor(xc).astype(np.int64), max(w - 2, 0))
    y0 = np.minimum(np.floor(yc).astype(np.int64), max(h - 2, 0))
    x1, y1 = np.minimum(x0 + 1, w - 1), np.minimum(y0 + 1, h - 1)
    ax, ay = (xc - x0)[..., None], (yc - y0)[..., None]
    top = src[y0, x0] * (1 - ax) + src[y0, x1] * ax
    bot = src[y1, x0] * (1 - ax) + src[y1, x1] * ax
    v = top * (1 - ay) + bot * ay
    v[~inside] = fill
    return v, inside


# ── results ──────────────────────────────────────────────────────────────────────────────

@dataclass
class Registration:
    """A registered pair: ``H`` takes moving pixel coordinates to fixed ones (3×3, row-major,
    1-based pixel centres) with ``score`` its SMI (both directions averaged when symmetric).
    ``method``: "refine" (climbed from a given pose), "search" (found by the sweep) or "stack"
    / "carry" (a slice of register_stack). The spline and keypoint fields of the reference's
    record stay empty (neither is part of this library)."""
    H: np.ndarray
    score: float
    model: str
    moving_shape: tuple
    fixed_shape: tuple
    spline: np.ndarray | None = None
    lattice: int = 0
    method: str = "refine"
    keypoints: tuple | None = None
    inliers: int | None = None
    iterations: int = 0
    redetected: bool = False
    _field: np.ndarray | None = field(default=None, repr=False)

    def __repr__(self):
        return f"Registration({self.model}, score {self.score:.4g}, {self.method}, {self.iterations} iterations)"

    def map_points(self, points):
        """Moving (x, y) points, 1-based (n, 2), to fixed coordinates."""
        return _apply(self.H, np.asarray(points, np.float64).reshape(-1, 2))

    def displacement(self):
        """x_moving − x_fixed over the fixed image, (h, w, 2) float32, 0-based pixels (the ITK
        displacement-field convention)."""
        if self._field is None:
            h, w = self.fixed_shape[:2]
            Hi = np.linalg.inv(np.asarray(self.H, np.float64))
            out = np.empty((h, w, 2), np.float32)
            xs = np.arange(1, w + 1, dtype=np.float64)
            for r0 in range(0, h, 512):
                ys = np.arange(r0 + 1, min(h, r0 + 512) + 1, dtype=np.float64)
                X, Y = np.meshgrid(xs, ys)
                p = _apply(Hi, np.c_[X.ravel(), Y.ravel()]).reshape(len(ys), w, 2)
                out[r0:r0 + len(ys)] = p - np.stack([X, Y], axis=-1)
            self._field = out
        return self._field

    def fixed_to_moving(self, points):
        """Fixed (x, y) points, 1-based, to moving coordinates."""
        return _apply(np.linalg.inv(np.asarray(self.H, np.float64)), np.asarray(points, np.float64).reshape(-1, 2))

    def warp(self, moving, fill=0, mask=False):
        """The moving image (any dtype, gray or with channels) resampled into the fixed frame,
        bilinear, `fill` outside; mask=True also returns where the moving image covers."""
        img = _read(moving)
        d = self.displacement()
        h, w = d.shape[:2]
        gx, gy = np.meshgrid(np.arange(w, dtype=np.float32), np.arange(h, dtype=np.float32))
        v, inside = _bilinear(img, gx + d[..., 0], gy + d[..., 1], fill)
        v = v.reshape((h, w) + img.shape[2:])
        if np.issubdtype(img.dtype, np.integer):
            lim = np.iinfo(img.dtype)
            v = np.clip(np.rint(v), lim.min, lim.max)
        v = v.astype(img.dtype)
        return (v, inside) if mask else v

    def to_dict(self):
        return {"H": np.asarray(self.H).tolist(), "score": float(self.score), "model": self.model,
                "method": self.method, "moving_shape": list(self.moving_shape[:2]),
                "fixed_shape": list(self.fixed_shape[:2]), "spline": None, "iterations": int(self.iterations),
                "redetected": bool(self.redetected)}


@dataclass
class StackRegistration:
    """Per slice pair: its Registration, how far its pose sits from where its neighbours put it
    (``off_trend``, largest corner distance in fixed pixels; None without neighbours) and whether
    that exceeds ``threshold`` (``flagged``)."""
    slices: list
    off_trend: list
    flagged: list
    threshold: float

    def __len__(self):
        return len(self.slices)

    def __getitem__(self, i):
        return self.slices[i]

    def __repr__(self):
        return f"StackRegistration({len(self.slices)} slices, {sum(self.flagged)} flagged, threshold {self.threshold:.1f} px)"

    @property
    def H(self):
        return [r.H for r in self.slices]

    def warp(self, moving, fill=0):
        """Each slice's moving image resampled into its fixed frame (a generator)."""
        for r, m in zip(self.slices, moving):
            yield r.warp(m, fill)

    def to_json(self, path=None):
        doc = json.dumps({
            "convention": "H maps moving pixel coordinates to fixed pixel coordinates, row-major 3x3, "
                          "1-based pixel centres",
            "threshold_px": self.threshold,
            "slices": [{**r.to_dict(), "off_trend_px": t, "flagged": bool(fl)}
                       for r, t, fl in zip(self.slices, self.off_trend, self.flagged)],
        }, indent=1)
        if path:
            Path(path).write_text(doc)
        return doc


# ── pairs ────────────────────────────────────────────────────────────────────────────────

def register(moving, fixed, *, model="homography", spline=False, detector="pos_gift", init=None, preprocess=True,
             **settings):
    """Refine a pose: the engine's climb (shift-map hops, roto-scale hops, Gauss–Newton blocks)
    from ``init`` (3×3, moving → fixed, 1-based). ``init=None`` would start from keypoints, which
    this library does not build: it raises EngineError (use ``search`` for a pose from nothing).
    ``detector`` is accepted for signature compatibility and not used."""
    if detector not in ("pos_gift", "gls_mift"):
        raise ValueError(f'detector: "pos_gift" or "gls_mift", not {detector!r}')
    if init is None:
        raise EngineError("register: Unsupported without init (keypoint detection is not part of this library; "
                          "pass init=H, or find a pose with search())")
    with _LOCK:
        reg, s, m, f = _with_pair(moving, fixed, preprocess, dict(model=model, spline=spline, **settings))
        reg.pose = np.asarray(init, np.float64).reshape(3, 3)
        c = reg.climb()
        H = reg.pose
        return Registration(H=H, score=reg.score(H).mean, model=model, moving_shape=m.shape, fixed_shape=f.shape,
                            method="refine", iterations=c.iterations)


def refine(moving, fixed, H, **kw):
    """register(moving, fixed, init=H, ...)."""
    return register(moving, fixed, init=H, **kw)


def score(moving, fixed, H, *, preprocess=True, **settings):
    """The SMI at pose H (both directions averaged when the symmetric setting is on)."""
    with _LOCK:
        reg, *_ = _with_pair(moving, fixed, preprocess, settings)
        return reg.score(np.asarray(H, np.float64)).mean


def shift_map(moving, fixed, H=None, *, preprocess=True, **settings):
    """The SMI at every translation of the moving image about H (default identity), with the
    settings' score, resolution and overlap floor."""
    with _LOCK:
        reg, *_ = _with_pair(moving, fixed, preprocess, settings)
        return reg.shift_map(np.eye(3) if H is None else np.asarray(H, np.float64))


def rs_map(moving, fixed, H=None, *, center=None, preprocess=True, **settings):
    """The SMI at every rotation × scale of the moving image about ``center`` (fixed pixels)."""
    with _LOCK:
        reg, *_ = _with_pair(moving, fixed, preprocess, settings)
        return reg.rs_map(np.eye(3) if H is None else np.asarray(H, np.float64), center=center)


def _sweep_settings(rotation, scale, shear, stretch) -> dict:
    """The sweep's range arguments as sw_* settings."""
    out = {}
    if rotation is not None:
        if np.ndim(rotation) == 0:
            out.update(sw_rot_range=False, sw_rot=float(rotation))
        else:
            lo, hi = rotation
            out.update(sw_rot_range=True, sw_rot_lo=float(lo), sw_rot_hi=float(hi))
    if scale is not None:
        out.update(sw_smin=float(scale[0]), sw_smax=float(scale[1]))
    for key, v in (("shear", shear), ("aniso", stretch)):
        if v is None:
            continue
        largest, count = (v, None) if np.ndim(v) == 0 else v
        out[f"sw_{key}"] = float(largest) > 0
        out[f"sw_{key}_max"] = float(largest)
        if count is not None:
            out[f"sw_n{key}"] = int(count)
    return out


def search(moving, fixed, H=None, *, mode="sweep", rotation=None, scale=None, shear=None, stretch=None,
           preprocess=True, **settings):
    """A pose from nothing (or about H with sw_rel): the sweep scores every rotation × scale (×
    shear × stretch) candidate's full shift map, keeps the best peaks after NMS and climbs the
    finalists with Gauss–Newton. Returns the best pose as a Registration (method "search").

    rotation: a half-range in degrees (30 → ±30°) or a (from, to) range (may run through ±180°);
    scale: (min, max); shear / stretch: the largest value or (largest, count). Only
    mode="sweep" is built (the reference's "cloud" search is not part of this library)."""
    if mode != "sweep":
        raise EngineError(f"search: mode {mode!r} Unsupported (only the sweep is part of this library)")
    kw = {**settings, **_sweep_settings(rotation, scale, shear, stretch), "search_mode": mode}
    with _LOCK:
        reg, s, m, f = _with_pair(moving, fixed, preprocess, kw)
        reg.pose = np.eye(3) if H is None else np.asarray(H, np.float64)
        reg.search()
        P = reg.pose
        return Registration(H=P, score=reg.score(P).mean, model=s["group"], moving_shape=m.shape,
                            fixed_shape=f.shape, method="search")


# ── stacks ───────────────────────────────────────────────────────────────────────────────

def _trend(slices, jump):
    """Each slice's distance from the pose its neighbours imply (the mean of the two neighbours'
    corner positions, or a straight-line extrapolation from two on one side), the threshold
    (``jump``, else 4× the median distance and at least 2 px) and the flags."""
    n = len(slices)
    dist = []
    for k, r in enumerate(slices):
        c = _corners(r.moving_shape)
        at = lambda j: _apply(slices[j].H, c) if 0 <= j < n else None   # noqa: E731
        prev, nxt = at(k - 1), at(k + 1)
        if prev is not None and nxt is not None:
            want = 0.5 * (prev + nxt)
        elif prev is not None:
            far = at(k - 2)
            want = 2 * prev - far if far is not None else prev
        elif nxt is not None:
            far = at(k + 2)
            want = 2 * nxt - far if far is not None else nxt
        else:
            want = None
        dist.append(None if want is None else float(np.hypot(*(_apply(r.H, c) - want).T).max()))
    known = sorted(d for d in dist if d is not None)
    thr = float(jump) if jump is not None else (max(2.0, 4.0 * known[len(known) // 2]) if known else math.inf)
    return dist, [d is not None and len(known) >= 3 and d > thr for d in dist], thr


def _gap(shape, Ha, Hb):
    c = _corners(shape)
    return float(np.hypot(*(_apply(Ha, c) - _apply(Hb, c)).T).max())


def register_stack(moving, fixed, *, init=None, start="previous", drop=0.15, jump=None, polish=False, progress=None,
                   model="homography", spline=False, detector="pos_gift", preprocess=True, iters=30, **settings):
    """Register a stack of slice pairs (serial sections, time series) — all slices at once on the
    GPU, where the reference registers them one after another (python/zcmir/api.py:522-580).

    Every slice starts from ``init`` (one 3×3 pose for all, one per slice, or None: each moving
    slice centred on its fixed one), corrected by the best peak of a small rotation × scale sweep
    about it (±12°, 0.88–1.12, the batched stand-in for the reference's keypoints), then ``iters``
    Gauss–Newton iterations run on every slice together. With ``start="previous"`` a slice whose
    score falls more than ``drop`` below the previous slice's, or whose pose jumps more than
    ``jump`` fixed px from it (None: automatic, 4× the median jump, at least 2 px), is refined
    again from the previous slice's pose and the better of the two kept (``redetected``);
    ``start="keypoints"`` keeps every slice's own result. ``polish``: afterwards refine each
    slice from the next slice's pose (last to first) and keep it when it scores higher.
    ``progress(i, n, registration)`` is called per slice as results are collected."""
    if start not in ("previous", "keypoints"):
        raise ValueError('start: "previous" or "keypoints"')
    if detector not in ("pos_gift", "gls_mift"):
        raise ValueError(f'detector: "pos_gift" or "gls_mift", not {detector!r}')
    n = min(len(moving), len(fixed))
    if n == 0:
        raise ValueError("register_stack: no slices")
    with _LOCK:
        reg = _engine()
        s = _settings_for(model=model, spline=spline, **settings)
        reg.configure(**s)
        mv = [_read(moving[i]) for i in range(n)]
        fx = [_read(fixed[i]) for i in range(n)]
        # the slices' work images (preprocessing on the device, as for pairs)
        work_m, work_f = [], []
        for m, f in zip(mv, fx):
            if preprocess or m.dtype == np.uint8 or f.dtype == np.uint8:
                reg.set_images(m, f, preprocess=True)
                work_m.append(reg.work_image(0))
                work_f.append(reg.work_image(1))
            else:
                work_m.append(np.asarray(m, np.float32))
                work_f.append(np.asarray(f, np.float32))
        _Shared.held = {}
        reg.set_stack(work_m, work_f)
        H0 = None
        if init is not None:
            a = np.asarray(init, np.float64)
            H0 = np.broadcast_to(a.reshape(-1, 3, 3), (n, 3, 3)) if a.size == 9 else a.reshape(n, 3, 3)
        Hs, scores, _, it = reg.register_stack(H0, iters=iters, init="sweep")
        out = [Registration(H=Hs[i].copy(), score=float(scores[i]), model=model, moving_shape=mv[i].shape,
                            fixed_shape=fx[i].shape, method="stack", iterations=int(it)) for i in range(n)]

        def climb_from(i, H):
            reg.set_images(work_m[i], work_f[i], preprocess=False)
            reg.pose = H
            c = reg.climb()
            P = reg.pose
            return Registration(H=P, score=reg.score(P).mean, model=model, moving_shape=mv[i].shape,
                                fixed_shape=fx[i].shape, method="carry", iterations=c.iterations)

        if start == "previous":
            jumps = []
            for i in range(1, n):
                prev, cur = out[i - 1], out[i]
                g = _gap(cur.moving_shape, cur.H, prev.H)
                med = sorted(jumps)[len(jumps) // 2] if len(jumps) >= 3 else None
                thr = float(jump) if jump is not None else (max(2.0, 4.0 * med) if med is not None else math.inf)
                if (prev.score > 0 and cur.score < prev.score * (1 - drop)) or g > thr:
                    alt = climb_from(i, prev.H)
                    if alt.score > cur.score:
                        alt.redetected = True
                        out[i] = alt
                    else:
                        cur.redetected = True
                jumps.append(_gap(out[i].moving_shape, out[i].H, out[i - 1].H))
        if polish:
            for i in range(n - 2, -1, -1):
                alt = climb_from(i, out[i + 1].H)
                if alt.score > out[i].score * (1 + 1e-4):
                    alt.redetected = out[i].redetected
                    out[i] = alt
        _Shared.held = {}
        if progress:
            for i, r in enumerate(out):
                progress(i, n, r)
        dist, flags, thr = _trend(out, jump)
        return StackRegistration(out, dist, flags, thr)
