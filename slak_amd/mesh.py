"""Host-side mirror of the reference's public API over the C-ABI.

``LAKCatmullClark.subdivideIteratively`` and ``CatmullClark.subdivideIteratively``
take and fill :class:`MeshInfo` objects whose pointers are device addresses,
exactly like ``MeshInfo<int,int,int,float>`` in LibInclude/LAKCatmullClark.h:95-156
and LibInclude/CatmullClark.h:69-130. Failures raise :class:`SlakError`
(the reference throws ``const char*``, cuda_host_helpers.h:34-41).
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Optional

import numpy as np

from . import capi
from .capi import SlakError, SlakMesh, SlakStats, check, lib
from .meshgen import HostMesh


class DeviceBuffer:
    """Device allocation from the library pool; freed with slak_free."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().slak_malloc(ctypes.byref(p), max(1, self.nbytes)))
        self.ptr = int(p.value)

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(lib().slak_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes))
        return b

    def free(self) -> None:
        if self.ptr:
            check(lib().slak_free(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def to_host(ptr: int, count: int, dtype) -> np.ndarray:
    out = np.empty(count, dtype=dtype)
    if count and ptr:
        check(lib().slak_memcpy_d2h(out.ctypes.data, ptr, out.nbytes))
    return out


@dataclasses.dataclass
class MeshInfo:
    """MeshInfo<int, int, int, float>: device pointers as integers."""

    type: int = capi.SLAK_POLY
    nfaces: int = 0
    nverts: int = 0
    nnz: int = 0
    max_face_order: int = 0
    is_reduced: bool = False
    ptr: int = 0
    ids: int = 0
    vals: int = 0
    verts: int = 0
    _owned: list = dataclasses.field(default_factory=list, repr=False)

    def to_c(self) -> SlakMesh:
        return SlakMesh(self.type, int(self.is_reduced), self.nfaces, self.nverts, self.nnz, self.max_face_order,
                        self.ptr or None, self.ids or None, self.vals or None, self.verts or None)

    def from_c(self, m: SlakMesh) -> None:
        self.type = m.type
        self.is_reduced = bool(m.is_reduced)
        self.nfaces, self.nverts, self.nnz, self.max_face_order = m.nfaces, m.nverts, m.nnz, m.max_face_order
        self.ptr = m.ptr or 0
        self.ids = m.ids or 0
        self.vals = m.vals or 0
        self.verts = m.verts or 0

    def freeAndReset(self) -> None:
        """MeshInfo::freeAndReset (LAKCatmullClark.h:120-127) with slak_free."""
        for name in ("ptr", "ids", "vals", "verts"):
            p = getattr(self, name)
            if p:
                check(lib().slak_free(p))
        self._owned.clear()
        self.__init__()

    @classmethod
    def from_host(cls, h: HostMesh, with_vals: bool = True) -> "MeshInfo":
        bufs = []
        m = cls(type=h.type, nfaces=h.nfaces, nverts=h.nverts, nnz=h.nnz, max_face_order=h.max_face_order,
                is_reduced=h.vals is None)
        if h.ptr is not None:
            b = DeviceBuffer.from_array(h.ptr.astype(np.int32))
            bufs.append(b)
            m.ptr = b.ptr
        b = DeviceBuffer.from_array(h.ids.astype(np.int32))
        bufs.append(b)
        m.ids = b.ptr
        if h.vals is not None and with_vals:
            b = DeviceBuffer.from_array(h.vals.astype(np.int32))
            bufs.append(b)
            m.vals = b.ptr
        b = DeviceBuffer.from_array(h.verts.astype(np.float32))
        bufs.append(b)
        m.verts = b.ptr
        for b in bufs:  # ownership moves to the MeshInfo (freeAndReset)
            b.ptr = 0
        return m

    def reduce(self, stream: int = 0) -> None:
        """GPU::CSCMesh::reduce (include/Mesh.h:305-334) on the device."""
        m = self.to_c()
        check(lib().slak_mesh_reduce(ctypes.byref(m), stream or None))
        self.from_c(m)

    def unreduce(self, stream: int = 0) -> None:
        """GPU::CSCMesh::unreduce (include/Mesh.h:336-377) on the device; a
        missing ptr / vals is allocated (freed by freeAndReset)."""
        m = self.to_c()
        check(lib().slak_mesh_unreduce(ctypes.byref(m), stream or None))
        self.from_c(m)

    def to_host(self) -> HostMesh:
        ptr = to_host(self.ptr, self.nfaces + 1, np.int32) if self.ptr else None
        vals = to_host(self.vals, self.nnz, np.int32) if self.vals else None
        ids = to_host(self.ids, self.nnz, np.int32)
        verts = to_host(self.verts, 4 * self.nverts, np.float32).reshape(-1, 4)
        return HostMesh(self.type, ptr, ids, vals, verts)


def _subdivide(scheme: int, cmesh: MeshInfo, rmesh: MeshInfo, target_level: int, stream: int = 0) -> SlakStats:
    cin = cmesh.to_c()
    cout = rmesh.to_c()
    st = SlakStats()
    check(lib().slak_subdivide(scheme, ctypes.byref(cin), ctypes.byref(cout), int(target_level), ctypes.byref(st),
                               stream or None))
    rmesh.from_c(cout)
    return st


class LAKCatmullClark:
    """LAKCatmullClark (LibInclude/LAKCatmullClark.h:33-192)."""

    @staticmethod
    def subdivideIteratively(cmesh: MeshInfo, rmesh: MeshInfo, target_level: int, stream: int = 0) -> SlakStats:
        return _subdivide(capi.SLAK_SCHEME_LAK, cmesh, rmesh, target_level, stream)


class CatmullClark:
    """CatmullClark, the specialised scheme (LibInclude/CatmullClark.h:33-165)."""

    @staticmethod
    def subdivideIteratively(cmesh: MeshInfo, rmesh: MeshInfo, target_level: int, stream: int = 0) -> SlakStats:
        return _subdivide(capi.SLAK_SCHEME_SLAK, cmesh, rmesh, target_level, stream)


def subdivide_host(h: HostMesh, levels: int, scheme: int = capi.SLAK_SCHEME_LAK):
    """Uploads a host mesh, refines it on the GPU and downloads the result."""
    c = MeshInfo.from_host(h)
    r = MeshInfo()
    try:
        st = _subdivide(scheme, c, r, levels)
        out = r.to_host() if r.ids else None
    finally:
        c.freeAndReset()
        if r.ids:
            r.freeAndReset()
    return out, st


@dataclasses.dataclass
class SpGEMMResult:
    F_ptr: np.ndarray
    F_ids: np.ndarray
    F_vals: np.ndarray
    D_ptr: np.ndarray
    D_ids: np.ndarray
    D_vals: np.ndarray
    E_ptr: np.ndarray
    E_ids: np.ndarray
    E_vals: np.ndarray
    nedges: int
    nextern: int


def spgemm_mapped_host(h: HostMesh) -> SpGEMMResult:
    """Runs the mapped SpGEMM operator (slak_spgemm_mapped) on a host mesh."""
    c = MeshInfo.from_host(h)
    cm = c.to_c()
    ptrs = [ctypes.c_void_p() for _ in range(9)]
    ne = ctypes.c_uint64()
    nx = ctypes.c_uint64()
    try:
        check(lib().slak_spgemm_mapped(ctypes.byref(cm), *[ctypes.byref(p) for p in ptrs], ctypes.byref(ne),
                                       ctypes.byref(nx), None))
        nv, nnz, e = h.nverts, h.nnz, int(ne.value)
        sizes = [nv + 1, nnz, nnz, nv + 1, nnz, nnz, nv + 1, e, e]
        arrs = [to_host(p.value, n, np.int32) for p, n in zip(ptrs, sizes)]
        for p in ptrs:
            check(lib().slak_free(p.value))
        return SpGEMMResult(*arrs, e, int(nx.value))
    finally:
        c.freeAndReset()
