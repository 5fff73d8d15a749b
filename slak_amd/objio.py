"""Wavefront OBJ I/O with the reference's conventions (include/obj.h:107-297,
include/Mesh.h:153-178): 1-based and negative face indices, v/vt/vn forms,
TRI/QUAD/POLY detection, vals = 1-based corner position, w = 1, the vertex
array sized by the largest referenced index, vertices written with 6
significant digits (default stream precision) and 1-based quads."""
from __future__ import annotations

import numpy as np

from .meshgen import POLY, QUAD, TRI, HostMesh, _vals_from_ptr


def read_obj(path: str) -> HostMesh:
    verts, offs, ids = [], [0], []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                verts.append([float(x) for x in parts[1:4]])
            elif parts[0] == "f":
                nv = len(verts)
                n = 0
                for tok in parts[1:]:
                    head = tok.split("/")[0]
                    try:
                        vi = int(head)
                    except ValueError:
                        break
                    ids.append(vi + nv if vi < 0 else vi - 1)
                    n += 1
                offs.append(offs[-1] + n)
    ptr = np.asarray(offs, dtype=np.int32)
    orders = np.diff(ptr)
    if orders.size == 0 or orders.min() != orders.max() or orders.max() > 4:
        typ = POLY
    elif orders.max() == 4:
        typ = QUAD
    elif orders.max() == 3:
        typ = TRI
    else:
        raise ValueError("Error loading obj: Face with less than three vertices detected.")
    ids_a = np.asarray(ids, dtype=np.int32)
    nv = int(ids_a.max()) + 1 if ids_a.size else 0
    v = np.ones((nv, 4), dtype=np.float32)
    src = np.asarray(verts, dtype=np.float32).reshape(-1, 3)
    v[: min(nv, src.shape[0]), :3] = src[:nv]
    return HostMesh(typ, ptr, ids_a, _vals_from_ptr(ptr), v, label=path)


def write_obj(mesh: HostMesh, path: str) -> None:
    ptr = mesh.face_ptr()
    with open(path, "w") as fh:
        for x, y, z in mesh.verts[:, :3]:
            fh.write(f"v {float(x):g} {float(y):g} {float(z):g}\n")
        for f in range(mesh.nfaces):
            fh.write("f " + " ".join(str(int(i) + 1) for i in mesh.ids[ptr[f]:ptr[f + 1]]) + "\n")
