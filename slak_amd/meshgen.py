"""Seeded synthetic control cages for the Catmull-Clark workloads.

Every generator returns a :class:`HostMesh` in the reference's incidence-matrix
form (one CSR row per face: ``ptr[nf+1]``, ``ids[nnz]`` vertex ids, ``vals[nnz]``
the 1-based corner position, ``verts[nv,4]`` float32 positions with w=1).
This is the layout ``GPU::CSCMesh`` builds from an OBJ file
(include/Mesh.h:153-178) and hands to ``subdivideIteratively``.

Randomness is counter-based splitmix64, so CPU, GPU and any re-implementation
see identical inputs:

* stream state ``s0 = mix64(mix64(seed + G) ^ (mesh * C1) ^ (purpose * C2))``
  and draw k (k = 0, 1, ...) is ``mix64(s0 + (k + 1) * G)``;
* a uniform on [0, 1) is the top 53 bits of a draw times 2**-53;
* Gaussians come from Box-Muller on consecutive uniform pairs
  ``(u1, u2)``: ``r = sqrt(-2 ln(1 - u1))``, ``z0 = r cos(2 pi u2)``,
  ``z1 = r sin(2 pi u2)``.

Purposes: 0 = position jitter, 1 = triangulation, 2 = pentagon placement,
3 = per-mesh shape parameters of a batch, 4 = request stream of a server run.
"""
from __future__ import annotations

import dataclasses
import hashlib
import math
from typing import List, Optional, Sequence

import numpy as np

G = 0x9E3779B97F4A7C15
C1 = 0xD1B54A32D192ED03
C2 = 0xCA5A826395121157
M1 = 0xBF58476D1CE4E5B9
M2 = 0x94D049BB133111EB
MASK = (1 << 64) - 1

TRI, QUAD, POLY = 0, 1, 2


def _mix64_int(z: int) -> int:
    z &= MASK
    z = ((z ^ (z >> 30)) * M1) & MASK
    z = ((z ^ (z >> 27)) * M2) & MASK
    return z ^ (z >> 31)


def _mix64_arr(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(M1)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(M2)
        return z ^ (z >> np.uint64(31))


def stream_state(seed: int, mesh: int, purpose: int) -> int:
    a = _mix64_int(seed + G)
    return _mix64_int(a ^ ((mesh * C1) & MASK) ^ ((purpose * C2) & MASK))


def draws(seed: int, mesh: int, purpose: int, n: int) -> np.ndarray:
    """The first ``n`` raw 64-bit draws of stream (seed, mesh, purpose)."""
    s0 = np.uint64(stream_state(seed, mesh, purpose))
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix64_arr(s0 + k * np.uint64(G))


def uniforms(seed: int, mesh: int, purpose: int, n: int) -> np.ndarray:
    return (draws(seed, mesh, purpose, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def gaussians(seed: int, mesh: int, purpose: int, n: int) -> np.ndarray:
    npairs = (n + 1) // 2
    u = uniforms(seed, mesh, purpose, 2 * npairs)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log1p(-u1))
    z = np.empty(2 * npairs, dtype=np.float64)
    z[0::2] = r * np.cos(2.0 * math.pi * u2)
    z[1::2] = r * np.sin(2.0 * math.pi * u2)
    return z[:n]


@dataclasses.dataclass
class HostMesh:
    """Host copy of a mesh in ``MeshInfo`` layout (LibInclude/LAKCatmullClark.h:95-156)."""

    type: int
    ptr: Optional[np.ndarray]  # int32 [nf+1] or None (homogeneous, reduced)
    ids: np.ndarray  # int32 [nnz]
    vals: Optional[np.ndarray]  # int32 [nnz] or None
    verts: np.ndarray  # float32 [nv, 4]
    label: str = ""

    @property
    def nverts(self) -> int:
        return int(self.verts.shape[0])

    @property
    def nnz(self) -> int:
        return int(self.ids.shape[0])

    @property
    def nfaces(self) -> int:
        if self.ptr is not None:
            return int(self.ptr.shape[0] - 1)
        k = 4 if self.type == QUAD else 3
        return self.nnz // k

    @property
    def max_face_order(self) -> int:
        if self.ptr is None:
            return 4 if self.type == QUAD else 3
        if self.nfaces == 0:
            return 0
        return int(np.max(np.diff(self.ptr)))

    def face_ptr(self) -> np.ndarray:
        if self.ptr is not None:
            return self.ptr
        k = 4 if self.type == QUAD else 3
        return (np.arange(self.nfaces + 1, dtype=np.int64) * k).astype(np.int32)

    def digest(self) -> str:
        h = hashlib.sha256()
        for a in (self.face_ptr(), self.ids, self.verts):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()[:16]

    def bbox_diag(self) -> float:
        p = self.verts[:, :3].astype(np.float64)
        if p.shape[0] == 0:
            return 0.0
        return float(np.linalg.norm(p.max(axis=0) - p.min(axis=0)))


def classify(ptr: np.ndarray) -> int:
    """Face-type detection of the OBJ reader (include/obj.h:237-253)."""
    orders = np.diff(ptr)
    if orders.size == 0:
        return POLY
    lo, hi = int(orders.min()), int(orders.max())
    if lo < 3:
        raise ValueError("face with less than three vertices")
    if lo != hi or hi > 4:
        return POLY
    return QUAD if hi == 4 else TRI


def _vals_from_ptr(ptr: np.ndarray) -> np.ndarray:
    nnz = int(ptr[-1])
    starts = np.repeat(ptr[:-1], np.diff(ptr))
    return (np.arange(nnz, dtype=np.int64) - starts + 1).astype(np.int32)


def _torus_positions(a: int, b: int, R: float, r: float, sigma: float, seed: int, mesh: int) -> np.ndarray:
    i = np.tile(np.arange(a, dtype=np.float64), b)
    j = np.repeat(np.arange(b, dtype=np.float64), a)
    u = 2.0 * math.pi * i / a
    v = 2.0 * math.pi * j / b
    p = np.empty((a * b, 3), dtype=np.float64)
    p[:, 0] = (R + r * np.cos(v)) * np.cos(u)
    p[:, 1] = (R + r * np.cos(v)) * np.sin(u)
    p[:, 2] = r * np.sin(v)
    if sigma != 0.0:
        p += sigma * gaussians(seed, mesh, 0, 3 * a * b).reshape(a * b, 3)
    out = np.ones((a * b, 4), dtype=np.float32)
    out[:, :3] = p.astype(np.float32)
    return out


def _quad_corners(a: int, b: int) -> np.ndarray:
    """[a*b, 4] corner vertex ids of quad (i, j) = face j*a + i."""
    i = np.tile(np.arange(a, dtype=np.int64), b)
    j = np.repeat(np.arange(b, dtype=np.int64), a)
    i1 = (i + 1) % a
    j1 = (j + 1) % b
    q = np.stack([j * a + i, j * a + i1, j1 * a + i1, j1 * a + i], axis=1)
    return q.astype(np.int32)


def torus(a: int, b: int, R: float = 1.0, r: float = 0.3, sigma: float = 1e-3,
          seed: int = 1, mesh: int = 0) -> HostMesh:
    """Closed a x b quad torus: vertex (i, j) = j*a + i, quad (i, j) =
    [(i,j), (i+1,j), (i+1,j+1), (i,j+1)] mod (a, b), FP32 positions with
    per-coordinate N(0, sigma^2) jitter from stream (seed, mesh, 0)."""
    if a < 3 or b < 3:
        raise ValueError("torus needs a, b >= 3")
    q = _quad_corners(a, b)
    nf = a * b
    ptr = (np.arange(nf + 1, dtype=np.int64) * 4).astype(np.int32)
    ids = q.reshape(-1).copy()
    return HostMesh(QUAD, ptr, ids, _vals_from_ptr(ptr), _torus_positions(a, b, R, r, sigma, seed, mesh),
                    label=f"torus{a}x{b}")


def _split_tris(q: np.ndarray, diag: np.ndarray):
    """Two triangles per quad; diag 0 cuts (c0,c2), diag 1 cuts (c1,c3).
    Returns (t_first [n,3], t_second [n,3]); t_second always holds edge (c2,c3)
    i.e. the quad's top edge (i+1,j+1)-(i,j+1)."""
    c0, c1, c2, c3 = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    d0 = diag == 0
    first = np.where(d0[:, None], np.stack([c0, c1, c2], 1), np.stack([c0, c1, c3], 1))
    second = np.where(d0[:, None], np.stack([c0, c2, c3], 1), np.stack([c1, c2, c3], 1))
    return first.astype(np.int32), second.astype(np.int32)


def torus_mixed(a: int, b: int, p_tri: float, pentagons: bool = False, R: float = 1.0, r: float = 0.3,
                sigma: float = 1e-3, seed: int = 1, mesh: int = 0) -> HostMesh:
    """Torus with each quad split into 2 triangles with probability ``p_tri``
    along a seeded diagonal (stream purpose 1: draw 2q decides the split of
    quad q, draw 2q+1 its diagonal), triangles replacing the quad in place in
    face order.

    With ``pentagons``, every grid column i gets one pentagon (stream purpose 2,
    draw i picks the row j): quad (i, j) is forced to be split and quad
    (i, j+1) to stay whole; the triangle of (i, j) that holds the shared edge
    (i,j+1)-(i+1,j+1) is merged with quad (i, j+1) into the 5-gon
    [(i,j+1), c, (i+1,j+1), (i+1,j+2), (i,j+2)], c the triangle's third vertex.
    The pentagon takes the place of quad (i, j+1); the merged triangle is
    dropped. This gives extraordinary vertices of valence 3..6."""
    q = _quad_corners(a, b)
    nq = a * b
    u = uniforms(seed, mesh, 1, 2 * nq)
    split = u[0::2] < p_tri
    diag = (u[1::2] < 0.5).astype(np.int64)
    pent_row = np.full(a, -1, dtype=np.int64)
    if pentagons:
        if b < 3:
            raise ValueError("pentagons need b >= 3")
        pu = uniforms(seed, mesh, 2, a)
        pent_row = np.minimum((pu * b).astype(np.int64), b - 1)
        for i in range(a):
            j = int(pent_row[i])
            split[j * a + i] = True
            split[((j + 1) % b) * a + i] = False
    first, second = _split_tris(q, diag)
    merged_below = np.zeros(nq, dtype=bool)  # quad whose second triangle joins the quad above
    pent_at = np.zeros(nq, dtype=bool)  # quad replaced by a pentagon
    if pentagons:
        cols = np.arange(a, dtype=np.int64)
        merged_below[pent_row * a + cols] = True
        pent_at[((pent_row + 1) % b) * a + cols] = True

    # slot A: pentagon, first triangle or the quad; slot B: the second triangle
    A = np.full((nq, 5), -1, dtype=np.int64)
    A_size = np.where(pent_at, 5, np.where(split, 3, 4))
    A[:, :4] = q
    A[split, :3] = first[split]
    A[split, 3] = -1
    if pentagons:
        qi = np.nonzero(pent_at)[0]
        i, j = qi % a, qi // a
        below = ((j - 1) % b) * a + i
        c = np.where(diag[below] == 0, q[below, 0], q[below, 1])
        A[qi] = np.stack([q[qi, 0], c, q[qi, 1], q[qi, 2], q[qi, 3]], 1)
    B = np.full((nq, 5), -1, dtype=np.int64)
    B[:, :3] = second
    B_has = split & ~merged_below

    faces = np.stack([A, B], 1).reshape(2 * nq, 5)
    sizes = np.stack([A_size, np.where(B_has, 3, 0)], 1).reshape(-1)
    keep = sizes > 0
    faces, sizes = faces[keep], sizes[keep]
    mask = np.arange(5)[None, :] < sizes[:, None]
    ids = faces[mask].astype(np.int32)
    ptr = np.zeros(sizes.shape[0] + 1, dtype=np.int64)
    np.cumsum(sizes, out=ptr[1:])
    ptr = ptr.astype(np.int32)
    verts = _torus_positions(a, b, R, r, sigma, seed, mesh)
    return HostMesh(classify(ptr), ptr, ids, _vals_from_ptr(ptr), verts,
                    label=f"torus{a}x{b}_tri{p_tri}{'_pent' if pentagons else ''}")


def torus_tri(a: int, b: int, p_tri: float, **kw) -> HostMesh:
    return torus_mixed(a, b, p_tri, pentagons=False, **kw)


# ---------------------------------------------------------------------------
# small known-answer meshes (Appendix-B style closed forms)

def cube() -> HostMesh:
    """[-1,1]^3 cube, 8 vertices, 6 outward-oriented quads."""
    v = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float32)
    f = np.array([[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]],
                 dtype=np.int32)
    ptr = (np.arange(7) * 4).astype(np.int32)
    verts = np.ones((8, 4), np.float32)
    verts[:, :3] = v
    return HostMesh(QUAD, ptr, f.reshape(-1).copy(), _vals_from_ptr(ptr), verts, label="cube")


def tetrahedron() -> HostMesh:
    v = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=np.float32)
    f = np.array([[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]], dtype=np.int32)
    ptr = (np.arange(5) * 3).astype(np.int32)
    verts = np.ones((4, 4), np.float32)
    verts[:, :3] = v
    return HostMesh(TRI, ptr, f.reshape(-1).copy(), _vals_from_ptr(ptr), verts, label="tetrahedron")


def single_quad() -> HostMesh:
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], dtype=np.float32)
    ptr = np.array([0, 4], np.int32)
    verts = np.ones((4, 4), np.float32)
    verts[:, :3] = v
    return HostMesh(QUAD, ptr, np.arange(4, dtype=np.int32), _vals_from_ptr(ptr), verts, label="quad")


def open_grid(a: int, b: int, sigma: float = 0.0, seed: int = 1, mesh: int = 0) -> HostMesh:
    """Planar a x b quad patch ((a+1)(b+1) vertices) with an open boundary."""
    nv = (a + 1) * (b + 1)
    i = np.tile(np.arange(a, dtype=np.int64), b)
    j = np.repeat(np.arange(b, dtype=np.int64), a)
    w = a + 1
    q = np.stack([j * w + i, j * w + i + 1, (j + 1) * w + i + 1, (j + 1) * w + i], 1).astype(np.int32)
    ptr = (np.arange(a * b + 1) * 4).astype(np.int32)
    vi = np.tile(np.arange(w, dtype=np.float64), b + 1)
    vj = np.repeat(np.arange(b + 1, dtype=np.float64), w)
    p = np.stack([vi / a, vj / b, np.zeros(nv)], 1)
    if sigma:
        p += sigma * gaussians(seed, mesh, 0, 3 * nv).reshape(nv, 3)
    verts = np.ones((nv, 4), np.float32)
    verts[:, :3] = p.astype(np.float32)
    return HostMesh(QUAD, ptr, q.reshape(-1).copy(), _vals_from_ptr(ptr), verts, label=f"grid{a}x{b}")


def cylinder(a: int, b: int, R: float = 1.0, height: float = 1.0) -> HostMesh:
    """Open cylinder: a quads around, b rows, two boundary loops."""
    nv = a * (b + 1)
    i = np.tile(np.arange(a, dtype=np.int64), b)
    j = np.repeat(np.arange(b, dtype=np.int64), a)
    i1 = (i + 1) % a
    q = np.stack([j * a + i, j * a + i1, (j + 1) * a + i1, (j + 1) * a + i], 1).astype(np.int32)
    ptr = (np.arange(a * b + 1) * 4).astype(np.int32)
    vi = np.tile(np.arange(a, dtype=np.float64), b + 1)
    vj = np.repeat(np.arange(b + 1, dtype=np.float64), a)
    u = 2 * math.pi * vi / a
    verts = np.ones((nv, 4), np.float32)
    verts[:, 0] = (R * np.cos(u)).astype(np.float32)
    verts[:, 1] = (R * np.sin(u)).astype(np.float32)
    verts[:, 2] = (height * vj / b).astype(np.float32)
    return HostMesh(QUAD, ptr, q.reshape(-1).copy(), _vals_from_ptr(ptr), verts, label=f"cyl{a}x{b}")


def _mesh_from_faces(faces, pos: np.ndarray, sigma: float, seed: int, mesh: int, label: str) -> HostMesh:
    """HostMesh from a list of faces (vertex id lists, consistently oriented)
    and float64 positions [nv, 3], with N(0, sigma^2) jitter from stream
    (seed, mesh, 0)."""
    sizes = np.array([len(f) for f in faces], dtype=np.int64)
    ptr = np.zeros(len(faces) + 1, dtype=np.int64)
    np.cumsum(sizes, out=ptr[1:])
    ids = np.concatenate([np.asarray(f, dtype=np.int64) for f in faces]).astype(np.int32)
    nv = pos.shape[0]
    p = pos.astype(np.float64).copy()
    if sigma:
        p += sigma * gaussians(seed, mesh, 0, 3 * nv).reshape(nv, 3)
    verts = np.ones((nv, 4), np.float32)
    verts[:, :3] = p.astype(np.float32)
    ptr = ptr.astype(np.int32)
    return HostMesh(classify(ptr), ptr, ids, _vals_from_ptr(ptr), verts, label=label)


def uv_sphere(segments: int, rings: int, sigma: float = 1e-3, seed: int = 1, mesh: int = 0) -> HostMesh:
    """Closed UV sphere: ``rings - 1`` rings of ``segments`` vertices between
    two poles, quads between rings and a triangle fan at each pole, so each
    pole has valence ``segments`` (the high-valence rows of the mapped SpGEMM).
    Vertex 0 is the north pole, ring k (k = 1 .. rings-1) holds vertices
    1 + (k-1) segments + i, the south pole is last."""
    if segments < 3 or rings < 2:
        raise ValueError("uv_sphere needs segments >= 3, rings >= 2")
    n = segments
    nring = rings - 1
    nv = 2 + nring * n
    south = nv - 1

    def r(k, i):
        return 1 + (k - 1) * n + (i % n)

    faces = [[0, r(1, i), r(1, i + 1)] for i in range(n)]
    for k in range(1, nring):
        faces += [[r(k, i), r(k + 1, i), r(k + 1, i + 1), r(k, i + 1)] for i in range(n)]
    faces += [[south, r(nring, i + 1), r(nring, i)] for i in range(n)]
    pos = np.zeros((nv, 3))
    pos[0] = (0.0, 0.0, 1.0)
    pos[south] = (0.0, 0.0, -1.0)
    for k in range(1, nring + 1):
        th = math.pi * k / rings
        for i in range(n):
            ph = 2.0 * math.pi * i / n
            pos[r(k, i)] = (math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th))
    return _mesh_from_faces(faces, pos, sigma, seed, mesh, f"uvsphere{segments}x{rings}")


def capped_cylinder(segments: int, rows: int, sigma: float = 1e-3, seed: int = 1, mesh: int = 0) -> HostMesh:
    """Closed prism-like cylinder: ``rows`` rings of quads around, closed by
    two faces of order ``segments`` (the long faces); every vertex has
    valence 3."""
    n = segments
    nv = n * (rows + 1)
    faces = [[j * n + i, j * n + (i + 1) % n, (j + 1) * n + (i + 1) % n, (j + 1) * n + i]
             for j in range(rows) for i in range(n)]
    faces.append([n - 1 - i for i in range(n)])                  # bottom cap
    faces.append([rows * n + i for i in range(n)])               # top cap
    pos = np.zeros((nv, 3))
    for j in range(rows + 1):
        for i in range(n):
            ph = 2.0 * math.pi * i / n
            pos[j * n + i] = (math.cos(ph), math.sin(ph), j / max(rows, 1) * 2.0 - 1.0)
    return _mesh_from_faces(faces, pos, sigma, seed, mesh, f"cappedcyl{segments}x{rows}")


def fan_disk(segments: int, sigma: float = 1e-3, seed: int = 1, mesh: int = 0) -> HostMesh:
    """Open disk: one ring of ``segments`` triangles around a centre vertex of
    valence ``segments``; the rim is a boundary loop."""
    n = segments
    faces = [[0, 1 + i, 1 + (i + 1) % n] for i in range(n)]
    pos = np.zeros((n + 1, 3))
    for i in range(n):
        ph = 2.0 * math.pi * i / n
        pos[1 + i] = (math.cos(ph), math.sin(ph), 0.0)
    return _mesh_from_faces(faces, pos, sigma, seed, mesh, f"fandisk{segments}")


def high_valence_mix(seed: int = 1) -> HostMesh:
    """One closed mesh with every row length the first level meets from 3 to
    24 and faces of order 3 to 12: a UV sphere with 24-triangle fans at the
    poles, a 12-gon-capped cylinder and a 9-segment sphere as separate
    components of one mesh."""
    parts = [uv_sphere(24, 6, seed=seed, mesh=0), capped_cylinder(12, 3, seed=seed, mesh=1),
             uv_sphere(9, 4, seed=seed, mesh=2)]
    ptr, ids, verts, off, cbase = [np.zeros(1, np.int64)], [], [], 0, 0
    for k, m in enumerate(parts):
        ptr.append(m.ptr[1:].astype(np.int64) + cbase)
        ids.append(m.ids.astype(np.int64) + off)
        v = m.verts.copy()
        v[:, 0] += 3.0 * k
        verts.append(v)
        off += m.nverts
        cbase += m.nnz
    ptr = np.concatenate(ptr).astype(np.int32)
    return HostMesh(classify(ptr), ptr, np.concatenate(ids).astype(np.int32), _vals_from_ptr(ptr),
                    np.concatenate(verts), label="high_valence_mix")


# ---------------------------------------------------------------------------
# workload configurations

def config1(seed: int = 1) -> HostMesh:
    """32x16 closed quad torus, R=1, r=0.3, jitter sigma=1e-3."""
    m = torus(32, 16, seed=seed)
    m.label = "cfg1_torus32x16"
    return m


def config2(seed: int = 2) -> HostMesh:
    """256x128 torus, quads split with p=0.25, one pentagon per grid column."""
    m = torus_mixed(256, 128, 0.25, pentagons=True, seed=seed)
    m.label = "cfg2_mixed256x128"
    return m


def config3(seed: int = 3) -> HostMesh:
    m = torus(1024, 1024, seed=seed)
    m.label = "cfg3_torus1024x1024"
    return m


@dataclasses.dataclass
class BatchSpec:
    meshes: List[HostMesh]
    levels: List[int]


def config4(n_meshes: int = 4096, seed: int = 4, p_tri: float = 0.2,
            dim_lo: int = 4, dim_hi: int = 64, depth_lo: int = 1, depth_hi: int = 4) -> BatchSpec:
    """Batch of independent seeded tori: grid dims uniform in [dim_lo, dim_hi]^2,
    a fraction p_tri of quads triangulated, depth uniform in [depth_lo, depth_hi].
    Shape parameters of mesh m come from stream (seed, m, 3)."""
    meshes, levels = [], []
    span = dim_hi - dim_lo + 1
    dspan = depth_hi - depth_lo + 1
    for m in range(n_meshes):
        u = uniforms(seed, m, 3, 3)
        a = dim_lo + min(int(u[0] * span), span - 1)
        b = dim_lo + min(int(u[1] * span), span - 1)
        d = depth_lo + min(int(u[2] * dspan), dspan - 1)
        mesh = torus_mixed(a, b, p_tri, pentagons=False, seed=seed, mesh=m) if p_tri > 0 else torus(a, b, seed=seed, mesh=m)
        meshes.append(mesh)
        levels.append(d)
    return BatchSpec(meshes, levels)


def request_mesh(seed: int, index: int, faces_lo: float = 100.0, faces_hi: float = 1.0e6,
                 p_tri: float = 0.1, level_lo: int = 1, level_hi: int = 3):
    """One served request: torus whose face count is log-uniform in
    [faces_lo, faces_hi] (b = round(sqrt(F/2)), a = round(F/b), both >= 4),
    a fraction p_tri of quads triangulated, level uniform in [level_lo, level_hi]."""
    u = uniforms(seed, index, 4, 3)
    F = math.exp(math.log(faces_lo) + u[0] * (math.log(faces_hi) - math.log(faces_lo)))
    b = max(4, int(round(math.sqrt(F / 2.0))))
    a = max(4, int(round(F / b)))
    span = level_hi - level_lo + 1
    lvl = level_lo + min(int(u[1] * span), span - 1)
    mesh = torus_mixed(a, b, p_tri, pentagons=False, seed=seed, mesh=index)
    return mesh, lvl


def random_poly_torus(a: int, b: int, seed: int = 1, p_merge: float = 0.15, p_split: float = 0.2,
                      p_hole: float = 0.3, sigma: float = 1e-3, mesh: int = 0) -> HostMesh:
    """Torus grid (a x b quads, R = 1, r = 0.3) reworked at random: pairs of
    quads along the major direction merged into hexagons (two corners of each
    on a straight side), quads split into two triangles along a random
    diagonal, and isolated holes (faces (i, j) with i, j = 1 mod 3 removed
    with probability p_hole, so no two holes share a vertex and every
    boundary vertex has one boundary fan). A mix of face orders 3, 4 and 6,
    valences 2 to 8 and boundary loops. Decisions come from a NumPy
    generator seeded with (seed, mesh)."""
    if a < 3 or b < 3:
        raise ValueError("random_poly_torus needs a, b >= 3")
    rng = np.random.default_rng([seed, mesh, 7])
    vid = lambda i, j: (j % b) * a + (i % a)  # noqa: E731
    hole = np.zeros((a, b), bool)
    for i in range(1, a - 1, 3):
        for j in range(1, b - 1, 3):
            hole[i, j] = rng.random() < p_hole
    used = hole.copy()
    faces = []
    for j in range(b):
        for i in range(a):
            if used[i, j]:
                continue
            used[i, j] = True
            q = [vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)]
            r = rng.random()
            i2 = (i + 1) % a
            if r < p_merge and not used[i2, j] and i2 != 0:
                used[i2, j] = True
                faces.append([vid(i, j), vid(i + 1, j), vid(i + 2, j), vid(i + 2, j + 1), vid(i + 1, j + 1),
                              vid(i, j + 1)])
            elif r < p_merge + p_split:
                if rng.random() < 0.5:
                    faces += [[q[0], q[1], q[2]], [q[0], q[2], q[3]]]
                else:
                    faces += [[q[0], q[1], q[3]], [q[1], q[2], q[3]]]
            else:
                faces.append(q)
    u = 2.0 * np.pi * (np.arange(a * b) % a) / a
    v = 2.0 * np.pi * (np.arange(a * b) // a) / b
    pos = np.stack([(1.0 + 0.3 * np.cos(v)) * np.cos(u), (1.0 + 0.3 * np.cos(v)) * np.sin(u), 0.3 * np.sin(v)], 1)
    return _mesh_from_faces(faces, pos, sigma, seed, mesh, f"random_poly_torus{a}x{b}_s{seed}")
