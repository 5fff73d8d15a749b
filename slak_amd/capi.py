"""ctypes binding of libslak.so (include/slak_c.h).

The product path: every call goes through the C-ABI into the HIP library.
If the library is missing the import of :func:`lib` raises; there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libslak.so")

SLAK_OK = 0
SLAK_TRI, SLAK_QUAD, SLAK_POLY = 0, 1, 2
SLAK_SCHEME_LAK, SLAK_SCHEME_SLAK = 0, 1

i32p = ctypes.POINTER(ctypes.c_int32)


class SlakMesh(ctypes.Structure):
    """struct slak_mesh (MeshInfo<int,int,int,float>)."""

    _fields_ = [
        ("type", ctypes.c_int32),
        ("is_reduced", ctypes.c_int32),
        ("nfaces", ctypes.c_uint64),
        ("nverts", ctypes.c_uint64),
        ("nnz", ctypes.c_uint64),
        ("max_face_order", ctypes.c_uint64),
        ("ptr", ctypes.c_void_p),
        ("ids", ctypes.c_void_p),
        ("vals", ctypes.c_void_p),
        ("verts", ctypes.c_void_p),
    ]


class SlakStats(ctypes.Structure):
    _fields_ = [
        ("wall_ms", ctypes.c_double),
        ("peak_bytes", ctypes.c_uint64),
        ("levels", ctypes.c_int32),
        ("host_syncs", ctypes.c_int32),
        ("out_faces", ctypes.c_uint64),
        ("out_verts", ctypes.c_uint64),
    ]


class SlakBatchInfo(ctypes.Structure):
    _fields_ = [
        ("nmeshes", ctypes.c_int32),
        ("max_levels", ctypes.c_int32),
        ("host_syncs", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("device_bytes", ctypes.c_uint64),
        ("out_faces", ctypes.c_uint64),
        ("out_verts", ctypes.c_uint64),
    ]


class SlakServerConfig(ctypes.Structure):
    _fields_ = [
        ("scheme", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("max_batch", ctypes.c_int32),
        ("slots", ctypes.c_int32),
        ("memory_budget_bytes", ctypes.c_uint64),
    ]


class SlakRequestStats(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("levels", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("submit_us", ctypes.c_double),
        ("admit_us", ctypes.c_double),
        ("done_us", ctypes.c_double),
        ("latency_us", ctypes.c_double),
        ("out_faces", ctypes.c_uint64),
    ]


class SlakServerStats(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("pending", ctypes.c_uint64),
        ("active", ctypes.c_uint64),
        ("inflight_bytes", ctypes.c_uint64),
        ("faces_done", ctypes.c_uint64),
        ("busy_us", ctypes.c_double),
        ("max_batch", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class SlakError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"slak error {code}: {msg}")
        self.code = code


# (name, restype, argtypes) of every exported entry point declared in include/slak_c.h
EXPORTS = [
    ("slak_subdivide", ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(SlakMesh), ctypes.POINTER(SlakMesh),
                                      ctypes.c_int32, ctypes.POINTER(SlakStats), ctypes.c_void_p]),
    ("slak_free", ctypes.c_int, [ctypes.c_void_p]),
    ("slak_mesh_reduce", ctypes.c_int, [ctypes.POINTER(SlakMesh), ctypes.c_void_p]),
    ("slak_mesh_unreduce", ctypes.c_int, [ctypes.POINTER(SlakMesh), ctypes.c_void_p]),
    ("slak_last_error", ctypes.c_char_p, []),
    ("slak_version", ctypes.c_int32, []),
    ("slak_malloc", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint64]),
    ("slak_memcpy_h2d", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
    ("slak_memcpy_d2h", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
    ("slak_memcpy_d2d", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
    ("slak_device_synchronize", ctypes.c_int, []),
    ("slak_pool_trim", ctypes.c_int, []),
    ("slak_pool_stats", ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64)] * 3),
    ("slak_spgemm_mapped", ctypes.c_int, [ctypes.POINTER(SlakMesh)] + [ctypes.POINTER(ctypes.c_void_p)] * 9
     + [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p]),
    ("slak_profile_enable", ctypes.c_int, [ctypes.c_int32]),
    ("slak_profile_report", ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64]),
    ("slak_copy_probe", ctypes.c_int, [ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    ("slak_scan_exclusive", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]),
    ("slak_batch_create", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(SlakMesh),
                                         ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.c_void_p]),
    ("slak_scan_exclusive_noalloc", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                                   ctypes.c_void_p]),
    ("slak_digest", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                   ctypes.c_void_p]),
    ("slak_batch_run", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    ("slak_batch_result", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(SlakMesh)]),
    ("slak_batch_info_get", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SlakBatchInfo)]),
    ("slak_batch_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("slak_server_start", ctypes.c_int, [ctypes.POINTER(SlakServerConfig), ctypes.POINTER(ctypes.c_void_p)]),
    ("slak_server_submit", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SlakMesh), ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_uint64)]),
    ("slak_server_wait", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(SlakMesh),
                                        ctypes.POINTER(SlakRequestStats)]),
    ("slak_server_get_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SlakServerStats)]),
    ("slak_server_stop", ctypes.c_int, [ctypes.c_void_p]),
]

_lib: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    """Loads libslak.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libslak.so not found at {LIB_PATH}; build it with `python -m slak_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in EXPORTS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def profile_enable(on: bool) -> None:
    check(lib().slak_profile_enable(1 if on else 0))


def profile_report() -> dict:
    import json
    buf = ctypes.create_string_buffer(1 << 20)
    check(lib().slak_profile_report(buf, len(buf)))
    return json.loads(buf.value.decode())


def digest(ptr: int, nbytes: int) -> int:
    """slak_digest: 64-bit order-sensitive digest of device memory."""
    out = ctypes.c_uint64()
    check(lib().slak_digest(ptr or None, int(nbytes), ctypes.byref(out), None))
    return int(out.value)


def check(code: int) -> None:
    if code != SLAK_OK:
        raise SlakError(code, lib().slak_last_error().decode())
