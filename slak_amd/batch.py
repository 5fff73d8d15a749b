"""Batched subdivision over the C-ABI (slak_batch_*): many meshes of uneven
size and depth refined with one launch per kernel per level, the chain
replayed as a hipGraph."""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import numpy as np

from . import capi
from .capi import SlakBatchInfo, SlakMesh, check, lib
from .mesh import MeshInfo


class Batch:
    """Owns a prepared batch; ``run()`` recomputes every mesh's refinement."""

    def __init__(self, meshes: Sequence[MeshInfo], levels: Sequence[int], scheme: int = capi.SLAK_SCHEME_LAK,
                 stream: int = 0):
        n = len(meshes)
        arr = (SlakMesh * max(n, 1))()
        for i, m in enumerate(meshes):
            arr[i] = m.to_c()
        lv = (ctypes.c_int32 * max(n, 1))(*[int(x) for x in levels])
        h = ctypes.c_void_p()
        check(lib().slak_batch_create(scheme, n, arr, lv, ctypes.byref(h), stream or None))
        self._h = h
        self.n = n
        self.scheme = scheme

    def run(self, use_graph: bool = True) -> None:
        check(lib().slak_batch_run(self._h, 1 if use_graph else 0))

    def result(self, i: int) -> MeshInfo:
        """View of mesh i's refined mesh (buffers owned by the batch)."""
        m = SlakMesh()
        check(lib().slak_batch_result(self._h, i, ctypes.byref(m)))
        out = MeshInfo()
        out.from_c(m)
        return out

    def info(self) -> SlakBatchInfo:
        inf = SlakBatchInfo()
        check(lib().slak_batch_info_get(self._h, ctypes.byref(inf)))
        return inf

    def close(self) -> None:
        if self._h:
            check(lib().slak_batch_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
