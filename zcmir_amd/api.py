"""The functional API over one shared engine (python/zcmir/api.py:1-148, 333-352): every
function takes the engine's settings as keyword arguments and starts from the defaults, so one
call never changes another. Calls are serialized (the engine is one device).
"""
import difflib
import threading

import numpy as np

from . import Registrar

_LOCK = threading.RLock()
_ENGINE = None
_DEFAULTS = None


def _engine():
    global _ENGINE, _DEFAULTS
    if _ENGINE is None:
        _ENGINE = Registrar()
        _DEFAULTS = _ENGINE.settings
    return _ENGINE


def close():
    """Release the shared engine (its GPU memory). The next call makes a new one."""
    global _ENGINE
    with _LOCK:
        if _ENGINE is not None:
            _ENGINE.close()
            _ENGINE = None


def settings():
    """Every setting with its default value."""
    with _LOCK:
        _engine()
        return dict(_DEFAULTS)


def _resolve(**kw):
    _engine()
    s = dict(_DEFAULTS)
    for k, v in kw.items():
        if k not in s:
            near = difflib.get_close_matches(k, s, n=3)
            raise TypeError(f"unknown setting {k!r}" + (f"; did you mean {', '.join(near)}?" if near else ""))
        s[k] = v
    return s


def _prepare(moving, fixed, preprocess, kw):
    reg = _engine()
    reg.configure(**_resolve(**kw))
    reg.set_images(moving, fixed, preprocess=preprocess)
    return reg


def score(moving, fixed, H, *, preprocess=True, **settings):
    """The registration score at pose H (the mean of both directions when symmetric)."""
    with _LOCK:
        reg = _prepare(moving, fixed, preprocess, settings)
        return reg.score(np.asarray(H, np.float64)).mean


def shift_map(moving, fixed, H=None, *, preprocess=True, **settings):
    """The score at every translation of the moving image about pose H (default: identity)."""
    with _LOCK:
        reg = _prepare(moving, fixed, preprocess, settings)
        return reg.shift_map(np.eye(3) if H is None else np.asarray(H, np.float64))
