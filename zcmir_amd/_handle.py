"""One engine of libzcsmi.so seen from Python: the opaque pointer, return-code checking, the
length-prefixed text calls, and the progress-event queue.

Every C entry point returns a negative value on failure and keeps the reason for ``zc_error``;
``Handle.call`` turns that into an ``EngineError`` (a ``RuntimeError``, as the reference package
raises: python/zcmir/__init__.py:230-238) whose text is ``"<what>: <reason>"``. Events are pulled
after every call and handed to the registrar's callback.
"""
from __future__ import annotations

import ctypes
import json

from . import _native


class EngineError(RuntimeError):
    """A call into the engine failed; the message names the call and the engine's reason."""


_lib = None


def library():
    """The loaded libzcsmi.so with every signature bound (loaded once per process). Raises when the
    library has not been built: there is no CPU fallback."""
    global _lib
    if _lib is None:
        _lib = _native.load()
    return _lib


def read_text(fn, *args, first=1024) -> bytes:
    """Text from a call that copies at most `cap` bytes into a buffer and returns the full length;
    grows the buffer once when the first try was short."""
    cap = first
    while True:
        buf = ctypes.create_string_buffer(cap)
        n = int(fn(*args, buf, cap))
        if n <= cap:
            return buf.raw[:n]
        cap = n


class Handle:
    """Owns one engine (one HIP device and one image pair). Not thread-safe."""

    EVENT_CHUNK = 1 << 16

    def __init__(self, device: int = 0, on_event=None):
        self.lib = library()
        self.on_event = on_event
        self.pending: list = []
        ptr = self.lib.zc_create(str(int(device)).encode())
        if not ptr:
            why = read_text(self.lib.zc_create_error).decode(errors="replace")
            raise EngineError(why or "zcmir_amd: no engine on this device")
        self.ptr = ptr

    # ── calls ──
    def fn(self, name: str):
        return getattr(self.lib, name)

    def call(self, what: str, name: str, *args) -> int:
        """Run zc_<name>(engine, *args); raise EngineError on a negative return, else return it."""
        if not self.ptr:
            raise EngineError(f"{what}: the engine is closed")
        rc = int(self.fn(name)(self.ptr, *args))
        self.pull_events()
        if rc < 0:
            raise EngineError(f"{what}: {self.reason() or rc}")
        return rc

    def reason(self) -> str:
        buf = ctypes.create_string_buffer(1024)
        n = self.lib.zc_error(self.ptr, buf, len(buf))
        return buf.raw[: min(n, len(buf))].decode(errors="replace")

    def text(self, name: str) -> str:
        return read_text(self.fn(name), self.ptr, first=8192).decode(errors="replace")

    # ── events ──
    def pull_events(self) -> None:
        buf = ctypes.create_string_buffer(self.EVENT_CHUNK)
        while True:
            n = self.lib.zc_events(self.ptr, buf, len(buf))
            if not n:
                return
            for line in buf.raw[:n].splitlines():
                if not line.strip():
                    continue
                ev = json.loads(line)
                self.pending.append(ev)
                if self.on_event is not None:
                    self.on_event(ev)

    def take_events(self) -> list:
        if self.ptr:
            self.pull_events()
        out, self.pending = self.pending, []
        return out

    # ── lifetime ──
    def close(self) -> None:
        ptr, self.ptr = getattr(self, "ptr", None), None
        if ptr:
            self.lib.zc_destroy(ptr)
