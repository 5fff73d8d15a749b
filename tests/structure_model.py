"""Host restatement of the refined-mesh structure rules the HIP refine and
vertex kernels apply (slak_amd/csrc/slak_kernels.hip, k_refine / k_vertex).

TEST INFRASTRUCTURE: pins the closed forms by which level l writes the
structure of level l+1 (edge ids of every corner's outgoing half-edge, twin
corners, sorted M^T rows with in-neighbours, each corner's M^T storage
position, boundary flags per edge) against
the same structure computed by brute force on the refined mesh, whose faces
come from the oracle. The brute force (``structure_of``) follows the
canonical numberings the reference's kernels produce when their atomics are
serialized: LAK ranks edges by (max, min) vertex (spgemm.cuh:71-100), SLAK by
the flagged half-edge in (vertex, corner) order (CatmullClarkKernels.cuh:147-236).
"""
from __future__ import annotations

from collections import defaultdict
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

LAK, SLAK = 0, 1
UNSET = -(1 << 30)


@dataclass
class Structure:
    ptr: np.ndarray          # nf + 1 face offsets
    ids: np.ndarray          # nnz corner -> vertex (position order)
    nv: int
    eout: np.ndarray         # nnz: edge id of the corner's outgoing half-edge
    twin: np.ndarray         # nnz: corner of the twin half-edge, -1 on the boundary
    bnd: np.ndarray          # ne: boundary flag by edge id
    tptr: np.ndarray         # nv + 1
    rows: List[Tuple[int, int]]  # nnz (corner, in-neighbour), rows sorted by corner
    mpos: np.ndarray         # nnz: position of the corner in the M^T storage


def _corner_tables(ptr: np.ndarray):
    nf = len(ptr) - 1
    nnz = int(ptr[-1])
    cface = np.zeros(nnz, dtype=np.int64)
    nextc = np.zeros(nnz, dtype=np.int64)
    prevc = np.zeros(nnz, dtype=np.int64)
    for f in range(nf):
        s, e = int(ptr[f]), int(ptr[f + 1])
        for t in range(s, e):
            cface[t] = f
            nextc[t] = s if t + 1 == e else t + 1
            prevc[t] = e - 1 if t == s else t - 1
    return cface, nextc, prevc


def structure_of(ptr: np.ndarray, ids: np.ndarray, nv: int, scheme: int) -> Structure:
    """Brute-force structure of a mesh from its faces."""
    ptr = np.asarray(ptr, dtype=np.int64)
    ids = np.asarray(ids, dtype=np.int64)
    cface, nextc, prevc = _corner_tables(ptr)
    nnz = len(ids)
    he = {}
    for t in range(nnz):
        he[(int(ids[t]), int(ids[nextc[t]]))] = t
    twin = np.full(nnz, -1, dtype=np.int64)
    for t in range(nnz):
        v, w = int(ids[t]), int(ids[nextc[t]])
        twin[t] = he.get((w, v), -1)
    rows_by_v = defaultdict(list)
    for t in range(nnz):
        rows_by_v[int(ids[t])].append(t)
    rank = {}
    if scheme == LAK:
        edges = sorted({(min(v, w), max(v, w)) for (v, w) in he}, key=lambda e: (e[1], e[0]))
        rank = {e: i for i, e in enumerate(edges)}
    else:
        k = 0
        for v in range(nv):
            for t in sorted(rows_by_v[v]):
                w = int(ids[nextc[t]])
                if twin[t] < 0 or v < w:
                    rank[(min(v, w), max(v, w))] = k
                    k += 1
    eout = np.array([rank[(min(int(ids[t]), int(ids[nextc[t]])), max(int(ids[t]), int(ids[nextc[t]])))]
                     for t in range(nnz)], dtype=np.int64)
    bnd = np.zeros(len(rank), dtype=np.int64)
    for t in range(nnz):
        if twin[t] < 0:
            bnd[eout[t]] = 1
    tptr = np.zeros(nv + 1, dtype=np.int64)
    rows = []
    mpos = np.zeros(nnz, dtype=np.int64)
    for v in range(nv):
        r = sorted(rows_by_v[v])
        tptr[v + 1] = tptr[v] + len(r)
        for i, t in enumerate(r):
            mpos[t] = tptr[v] + i
        rows.extend((t, int(ids[prevc[t]])) for t in r)
    return Structure(ptr, ids, nv, eout, twin, bnd, tptr, rows, mpos)


def refine_structure(S: Structure, scheme: int) -> Structure:
    """The structure of the refined mesh as the refine (per corner) and
    vertex (per vertex) kernels write it, in closed form from S."""
    ptr, ids, nv = S.ptr, S.ids, S.nv
    nf = len(ptr) - 1
    nnz = len(ids)
    ne = len(S.bnd)
    n_ext = int(S.bnd.sum())
    cface, nextc, prevc = _corner_tables(ptr)
    blk = np.concatenate([[0], np.cumsum(S.bnd)]).astype(np.int64)
    nv2, nnz2, ne2 = nv + nf + ne, 4 * nnz, nnz + 2 * ne
    nci = np.full((nnz2, 2), UNSET, dtype=np.int64)
    ntrow: List[Optional[Tuple[int, int]]] = [None] * nnz2
    ntptr = np.full(nv2 + 1, UNSET, dtype=np.int64)
    nbnd = np.full(ne2, UNSET, dtype=np.int64)
    nids = np.zeros(nnz2, dtype=np.int64)
    nmpos = np.full(nnz2, UNSET, dtype=np.int64)
    # ---- refine: one step per corner t
    for t in range(nnz):
        f, tn, tp = int(cface[t]), int(nextc[t]), int(prevc[t])
        v, w, u = int(ids[t]), int(ids[tn]), int(ids[tp])
        e1, e0 = int(S.eout[t]), int(S.eout[tp])
        s1, s0 = int(S.twin[t]), int(S.twin[tp])
        bl1, bl0 = int(blk[e1]), int(blk[e0])
        nids[4 * t:4 * t + 4] = [v, nv + nf + e1, nv + f, nv + nf + e0]
        g1 = int(cface[s1]) if s1 >= 0 else f
        n1 = int(nextc[s1]) if s1 >= 0 else 0
        g0 = int(cface[s0]) if s0 >= 0 else f
        own = s1 < 0 or t < s1
        tw = [4 * n1 + 3 if s1 >= 0 else -1, 4 * tn + 2, 4 * tp + 1, 4 * s0 if s0 >= 0 else -1]
        if scheme == LAK:
            b1, b0 = 4 * e1 - bl1, 4 * e0 - bl0
            eo = [b1 + (v > w), b1 + 2 + (s1 >= 0 and f > g1), b0 + 2 + (s0 >= 0 and f > g0), b0 + (v > u)]
        else:
            # vertex-owned edges take the owning corner's M^T position
            eo = [int(S.mpos[t]), nnz + tn, nnz + t, 2 * nnz + bl0 if s0 < 0 else int(S.mpos[s0])]
        for k in range(4):
            nci[4 * t + k] = (eo[k], tw[k])
        # M^T positions of the refined corners: vertex rows keep their order,
        # face point rows hold corner 2 in corner order, edge point rows are
        # sorted (rank among the row's 2 or 4 corners)
        r1 = [4 * tn + 3] + ([4 * s1 + 1, 4 * n1 + 3] if s1 >= 0 else [])
        ns0 = int(nextc[s0]) if s0 >= 0 else 0
        r0 = [4 * tp + 1] + ([4 * s0 + 1, 4 * ns0 + 3] if s0 >= 0 else [])
        nmpos[4 * t] = S.mpos[t]
        nmpos[4 * t + 1] = 2 * nnz + 4 * e1 - 2 * bl1 + sum(x < 4 * t + 1 for x in r1)
        nmpos[4 * t + 2] = nnz + t
        nmpos[4 * t + 3] = 2 * nnz + 4 * e0 - 2 * bl0 + sum(x < 4 * t + 3 for x in r0)
        ntrow[nnz + t] = (4 * t + 2, nv + nf + e1)
        if t == int(ptr[f]):
            ntptr[nv + f] = nnz + t
        if scheme == SLAK:
            nbnd[nnz + t] = 0
        if own:
            row = 2 * nnz + 4 * e1 - 2 * bl1
            ntptr[nv + nf + e1] = row
            ent = [(4 * t + 1, v), (4 * tn + 3, nv + f)]
            if s1 >= 0:
                ent += [(4 * s1 + 1, w), (4 * n1 + 3, nv + g1)]
            for i, x in enumerate(sorted(ent)):
                ntrow[row + i] = x
            if scheme == LAK:
                flags = [1, 1, 0] if s1 < 0 else [0, 0, 0, 0]
                for k, fl in enumerate(flags):
                    nbnd[b1_of(e1, bl1) + k] = fl
            elif s1 < 0:
                nbnd[2 * nnz + bl1] = 1
    # ---- vertex: one step per vertex
    for v in range(nv):
        pos = int(S.tptr[v])
        ntptr[v] = pos
        for i in range(int(S.tptr[v + 1]) - pos):
            t, _ = S.rows[pos + i]
            ntrow[pos + i] = (4 * t, nv + nf + int(S.eout[prevc[t]]))
            if scheme == SLAK:
                nbnd[pos + i] = 1 if int(S.twin[t]) < 0 else 0
    ntptr[nv2] = nnz2
    assert (nci != UNSET).all() and all(r is not None for r in ntrow)
    assert (ntptr != UNSET).all() and (nbnd != UNSET).all() and (nmpos != UNSET).all()
    assert int(nbnd.sum()) == 2 * n_ext
    return Structure(np.arange(nnz + 1, dtype=np.int64) * 4, nids, nv2, nci[:, 0], nci[:, 1], nbnd, ntptr,
                     [tuple(map(int, r)) for r in ntrow], nmpos)


def b1_of(e: int, bl: int) -> int:
    """First edge id of the refined mesh owned by the edge point of e (LAK)."""
    return 4 * e - bl


@dataclass
class FinalForm:
    ids2: np.ndarray                   # 16 per corner of M: the output quads of M1 face t
    slots: dict                        # M1 edge id -> (owner corner of M, which edge)
    fp_rows: List[List[Tuple[int, int]]]   # per face F of M: the M1 row of fp(F)
    ep_rows: List[List[Tuple[int, int]]]   # per edge of M (owner order): the M1 row of ep(e)


def final_closed_form(S: Structure, scheme: int) -> FinalForm:
    """The last level in closed form from a quad mesh M, as the final-level
    kernel (slak_amd/csrc/slak_kernels.hip, final_corner) computes it: the
    output quads of every M1 corner, the M1 edge ids of the edge points each
    owner of an M edge writes, and the M^T rows of the M1 vertices fp(F) and
    ep(e) whose vertex points it sums in row order."""
    ptr, ids, nv = S.ptr, S.ids, S.nv
    nf, nnz, ne = len(ptr) - 1, len(ids), len(S.bnd)
    assert nnz == 4 * nf
    blk = np.concatenate([[0], np.cumsum(S.bnd)]).astype(np.int64)
    ep0, nv1 = nv + nf, nv + nf + ne
    epb2 = nv1 + nnz
    ids2 = np.zeros(16 * nnz, dtype=np.int64)
    slots = {}
    fp_rows: List[List[Tuple[int, int]]] = [[] for _ in range(nf)]
    ep_rows: List[List[Tuple[int, int]]] = []
    for t in range(nnz):
        F, q0 = t >> 2, t & ~3
        tn, tp = q0 | ((t + 1) & 3), q0 | ((t + 3) & 3)
        v, w, u = int(ids[t]), int(ids[tn]), int(ids[tp])
        e1, e0 = int(S.eout[t]), int(S.eout[tp])
        s1, s0 = int(S.twin[t]), int(S.twin[tp])
        g1 = s1 >> 2 if s1 >= 0 else F
        g0 = s0 >> 2 if s0 >= 0 else F
        n1 = (s1 & ~3) | ((s1 + 1) & 3) if s1 >= 0 else 0
        bl1, bl0 = int(blk[e1]), int(blk[e0])
        if scheme == LAK:
            b1, b0 = 4 * e1 - bl1, 4 * e0 - bl0
            ex, ey = b1 + (v > w), b1 + 2 + (s1 >= 0 and F > g1)
            ez, ew = b0 + 2 + (s0 >= 0 and F > g0), b0 + (v > u)
        else:
            ex, ey, ez = int(S.mpos[t]), nnz + tn, nnz + t
            ew = 2 * nnz + bl0 if s0 < 0 else int(S.mpos[s0])
        x = [v, ep0 + e1, nv + F, ep0 + e0]
        E1 = [ex, ey, ez, ew]
        E0 = [ew, ex, ey, ez]
        for k in range(4):
            ids2[16 * t + 4 * k: 16 * t + 4 * k + 4] = [x[k], epb2 + E1[k], nv1 + t, epb2 + E0[k]]
        fp_rows[F].append((4 * t + 2, ep0 + e1))
        if s1 < 0 or t < s1:
            if scheme == LAK:
                i_vh, i_wh = b1 + (v > w), b1 + (w > v)
                i_fF, i_fG = b1 + 2 + (s1 >= 0 and F > g1), b1 + 2 + (g1 > F)
            else:
                i_vh = int(S.mpos[t])
                i_wh = int(S.mpos[s1]) if s1 >= 0 else 2 * nnz + bl1
                i_fF, i_fG = nnz + tn, nnz + n1
            for name, i in (("vh", i_vh), ("wh", i_wh), ("fF", i_fF)) + ((("fG", i_fG),) if s1 >= 0 else ()):
                assert i not in slots, (name, i)
                slots[i] = (t, name)
            ent = [(4 * t + 1, v), (4 * tn + 3, nv + F)]
            if s1 >= 0:
                ent += [(4 * s1 + 1, w), (4 * n1 + 3, nv + g1)]
            ep_rows.append(sorted(ent))
    return FinalForm(ids2, slots, fp_rows, ep_rows)
