"""Serving front-end over the C-ABI (slak_server_*): continuous batching of
subdivision requests, p50/p99 latency from submit to resident result."""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

from . import capi
from .capi import SlakMesh, SlakRequestStats, SlakServerConfig, SlakServerStats, check, lib
from .mesh import MeshInfo


class Server:
    def __init__(self, scheme: int = capi.SLAK_SCHEME_LAK, memory_budget_bytes: int = 0, max_batch: int = 0,
                 device: int = 0, slots: int = 0):
        cfg = SlakServerConfig(scheme, device, max_batch, slots, memory_budget_bytes)
        h = ctypes.c_void_p()
        check(lib().slak_server_start(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.scheme = scheme

    def submit(self, mesh: MeshInfo, level: int) -> int:
        """Queues a request; the device mesh must stay valid until it completes."""
        m = mesh.to_c()
        rid = ctypes.c_uint64()
        check(lib().slak_server_submit(self._h, ctypes.byref(m), int(level), ctypes.byref(rid)))
        return int(rid.value)

    def wait(self, request_id: int, timeout_ms: int = -1) -> Tuple[MeshInfo, SlakRequestStats]:
        """The refined mesh (caller-owned: freeAndReset) and the request's timings."""
        out = SlakMesh()
        st = SlakRequestStats()
        check(lib().slak_server_wait(self._h, request_id, timeout_ms, ctypes.byref(out), ctypes.byref(st)))
        r = MeshInfo()
        if st.levels > 0:
            r.from_c(out)
        return r, st

    def stats(self) -> SlakServerStats:
        st = SlakServerStats()
        check(lib().slak_server_get_stats(self._h, ctypes.byref(st)))
        return st

    def stop(self) -> None:
        if self._h:
            check(lib().slak_server_stop(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
