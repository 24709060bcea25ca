"""Render server over the C ABI (tvk_server_* in include/tvk_hip.h).

Requests are submitted from any thread and batched by the library's worker
thread into persistent-kernel launches over several resident volumes; see
the header for the admission and batching rules.

    srv = RenderServer(arena_bytes=1 << 30, lanes=4)
    vid = srv.addVolume(renderer)            # renderer holds its volume + TFs
    t = srv.submit(vid, view, mode=RenderMode.RM_1DTRANS)
    img, res = srv.wait(t)                   # res.latency_ms, res.nominal_samples
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import Request, RequestResult, ServerStats, call
from .renderer import Renderer


class RenderServer:
    def __init__(self, device: int = 0, arena_bytes: int = 1 << 30, lanes: int = 4):
        self._h = C.c_void_p()
        call("tvk_server_create", int(device), int(arena_bytes), int(lanes), C.byref(self._h))
        self._out = {}     # ticket -> image buffer kept alive until wait()
        self._renderers = []

    def addVolume(self, renderer) -> int:
        vid = C.c_uint32()
        call("tvk_server_add_volume", self._h, renderer._h, C.byref(vid))
        self._renderers.append(renderer)
        return vid.value

    def submit(self, volume: int, view, mode: int = 0, readback: bool = True) -> int:
        arr, imgs = Renderer._views([view], readback)
        q = Request()
        q.volume = int(volume)
        q.render_mode = int(mode)
        q.view = arr[0]
        t = C.c_uint64()
        call("tvk_server_submit", self._h, C.byref(q), C.byref(t))
        self._out[t.value] = imgs[0] if readback else None
        return t.value

    def wait(self, ticket: int, timeout_ms: int = 0):
        """(image or None, RequestResult) of a submitted request."""
        res = RequestResult()
        call("tvk_server_wait", self._h, int(ticket), int(timeout_ms), C.byref(res))
        return self._out.pop(ticket), res

    def stats(self) -> ServerStats:
        st = ServerStats()
        call("tvk_server_get_stats", self._h, C.byref(st))
        return st

    def setPartCost(self, samples: int) -> None:
        """Requests above `samples` estimated samples render in parts (0:
        the default, 4e9; 2**64 - 1: never split)."""
        call("tvk_server_set_part_cost", self._h, int(samples))

    def close(self) -> None:
        if self._h:
            _lib.load().tvk_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
