"""ctypes front-end of the CPU oracle (oracle/lib/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``, as the checker. The product
(``slak_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import subprocess
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")

LAK, SLAK = 0, 1

_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        i32p = ctypes.POINTER(ctypes.c_int32)
        L.oracle_run.restype = ctypes.c_void_p
        L.oracle_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, i32p, i32p, i32p,
                                 ctypes.POINTER(ctypes.c_float), ctypes.c_int]
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_info.restype = ctypes.c_int64
        L.oracle_info.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.oracle_array_len.restype = ctypes.c_int64
        L.oracle_array_len.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]
        L.oracle_array_copy.restype = ctypes.c_int
        L.oracle_array_copy.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p]
        L.oracle_level_scalar.restype = ctypes.c_int64
        L.oracle_level_scalar.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]
        L.oracle_release.argtypes = [ctypes.c_void_p]
        L.oracle_time_ms.restype = ctypes.c_double
        L.oracle_time_ms.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_int64, i32p, i32p, i32p, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        L.oracle_max_threads.restype = ctypes.c_int
        pp = ctypes.POINTER(ctypes.c_void_p)
        L.oracle_time_batch_ms.restype = ctypes.c_double
        L.oracle_time_batch_ms.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int64), pp, pp, pp, pp,
                                           ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int64)]
        _lib = L
    return _lib


def _p(a: Optional[np.ndarray], ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


@dataclasses.dataclass
class OracleResult:
    levels: int
    nf: int
    nv: int
    ptr: Optional[np.ndarray]
    ids: np.ndarray
    vals: Optional[np.ndarray]
    P: np.ndarray  # float64 [nv, 4]
    debug: list  # per level dict of intermediate arrays


_DEBUG_NAMES = ["order", "Mt_ptr", "Mt_ids", "Mt_vals", "F_ptr", "F_ids", "F_vals", "D_ptr", "D_ids", "D_vals",
                "E_ptr", "E_ids", "E_vals", "int0", "int1", "intid", "f0", "f1", "ext0", "ext1", "extid",
                "S_ptr", "S_ids", "S_epids"]


def _inputs(mesh):
    ptr = None if mesh.ptr is None else np.ascontiguousarray(mesh.ptr, dtype=np.int32)
    ids = np.ascontiguousarray(mesh.ids, dtype=np.int32)
    vals = None if mesh.vals is None else np.ascontiguousarray(mesh.vals, dtype=np.int32)
    verts = np.ascontiguousarray(mesh.verts, dtype=np.float32)
    return ptr, ids, vals, verts


def subdivide(mesh, levels: int, scheme: int = LAK, precision: int = 64, nthreads: int = 1,
              keep_debug: bool = False) -> OracleResult:
    """Run the oracle's subdivideIteratively on a host mesh (slak_amd.meshgen.HostMesh)."""
    L = lib()
    ptr, ids, vals, verts = _inputs(mesh)
    h = L.oracle_run(scheme, precision, nthreads, int(keep_debug), mesh.type, mesh.nfaces, mesh.nverts, mesh.nnz,
                     _p(ptr, ctypes.c_int32), _p(ids, ctypes.c_int32), _p(vals, ctypes.c_int32),
                     _p(verts, ctypes.c_float), levels)
    if not h:
        raise ValueError(L.oracle_last_error().decode())
    try:
        def arr(level, name, dtype=np.int32):
            n = L.oracle_array_len(h, level, name.encode())
            if n < 0:
                raise KeyError(name)
            out = np.zeros(n, dtype=dtype)
            if n:
                L.oracle_array_copy(h, level, name.encode(), out.ctypes.data_as(ctypes.c_void_p))
            return out

        lv = L.oracle_info(h, b"levels")
        nf = L.oracle_info(h, b"nf")
        nv = L.oracle_info(h, b"nv")
        P = arr(-1, "P", np.float64).reshape(-1, 4)
        ptr_o = arr(-1, "ptr") if L.oracle_info(h, b"has_ptr") else None
        vals_o = arr(-1, "vals") if L.oracle_info(h, b"has_vals") else None
        ids_o = arr(-1, "ids")
        debug = []
        for lev in range(L.oracle_info(h, b"ndebug")):
            d: Dict[str, object] = {n: arr(lev, n) for n in _DEBUG_NAMES}
            d["nedges"] = L.oracle_level_scalar(h, lev, b"nedges")
            d["nextern"] = L.oracle_level_scalar(h, lev, b"nextern")
            debug.append(d)
        return OracleResult(int(lv), int(nf), int(nv), ptr_o, ids_o, vals_o, P, debug)
    finally:
        L.oracle_release(h)


def time_cpu(mesh, levels: int, nthreads: int, scheme: int = LAK, warmup: int = 1, reps: int = 5):
    """Median wall ms of the FP32 LAK algorithm on ``nthreads`` host threads."""
    L = lib()
    ptr, ids, vals, verts = _inputs(mesh)
    faces = ctypes.c_int64(0)
    ms = L.oracle_time_ms(scheme, nthreads, mesh.type, mesh.nfaces, mesh.nverts, mesh.nnz,
                          _p(ptr, ctypes.c_int32), _p(ids, ctypes.c_int32), _p(vals, ctypes.c_int32),
                          _p(verts, ctypes.c_float), levels, warmup, reps, ctypes.byref(faces))
    if ms < 0:
        raise ValueError(L.oracle_last_error().decode())
    return ms, int(faces.value)


def time_cpu_batch(meshes, levels, nthreads: int, scheme: int = LAK, warmup: int = 1, reps: int = 3):
    """Median wall ms of the FP32 LAK algorithm over a list of meshes, each at
    its own level, spread over ``nthreads`` host threads (one mesh per thread
    at a time). Returns (ms, total final faces)."""
    L = lib()
    n = len(meshes)
    keep = [_inputs(m) for m in meshes]
    VP = ctypes.c_void_p * n

    def ptrs(k):
        return VP(*[None if a[k] is None else a[k].ctypes.data for a in keep])

    has_vals = any(a[2] is not None for a in keep)
    types = (ctypes.c_int * n)(*[m.type for m in meshes])
    nf = (ctypes.c_int64 * n)(*[m.nfaces for m in meshes])
    nv = (ctypes.c_int64 * n)(*[m.nverts for m in meshes])
    nnz = (ctypes.c_int64 * n)(*[m.nnz for m in meshes])
    lv = (ctypes.c_int * n)(*[int(x) for x in levels])
    faces = ctypes.c_int64(0)
    cast = lambda x: ctypes.cast(x, ctypes.POINTER(ctypes.c_void_p))
    ms = L.oracle_time_batch_ms(scheme, nthreads, n, types, nf, nv, nnz, cast(ptrs(0)), cast(ptrs(1)),
                                cast(ptrs(2)) if has_vals else None, cast(ptrs(3)), lv, warmup, reps,
                                ctypes.byref(faces))
    if ms < 0:
        raise ValueError(L.oracle_last_error().decode())
    return ms, int(faces.value)


def max_threads() -> int:
    return int(lib().oracle_max_threads())
