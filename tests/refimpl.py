"""Independent textbook restatements used to pin the oracle (tests only).

* ``lak_edge_rank`` / ``slak_edge_rank``: canonical edge numbering written
  from its definition (sort undirected edges by (max, min); rank flagged
  half-edges in (vertex, corner) order).
* ``cc_textbook``: one Catmull-Clark level in the textbook form
  vp = (Q + 2R + (n-3) S) / n, Q = mean of adjacent face points, R = mean of
  adjacent edge midpoints, S = old position; boundary rules 3/4, 1/8 and
  edge midpoints. Different algebra from the reference's
  (n-2)/n P + sum(P_i)/n^2 + sum(fp)/n^2, so a wrong term fails the match.
* ``dense_mapped_products``: F, D, E from dense n_v x n_v matrices with the
  reference predicate (tiny meshes only).
* ``bspline_refine``: tensor-product cubic B-spline subdivision of a periodic
  grid, which Catmull-Clark equals on an all-valence-4 closed quad grid.
"""
from __future__ import annotations

from collections import defaultdict
from typing import Dict, List, Tuple

import numpy as np


def faces_of(ptr, ids) -> List[List[int]]:
    return [list(map(int, ids[ptr[f]:ptr[f + 1]])) for f in range(len(ptr) - 1)]


def half_edges(faces):
    """(v, w, face, corner) for every corner, corner = global CSR index."""
    out = []
    t = 0
    for f, fv in enumerate(faces):
        n = len(fv)
        for k in range(n):
            out.append((fv[k], fv[(k + 1) % n], f, t))
            t += 1
    return out


def lak_edge_rank(faces) -> Dict[Tuple[int, int], int]:
    edges = set()
    for v, w, _, _ in half_edges(faces):
        edges.add((min(v, w), max(v, w)))
    order = sorted(edges, key=lambda e: (e[1], e[0]))
    return {e: i for i, e in enumerate(order)}


def slak_edge_rank(faces) -> Dict[Tuple[int, int], int]:
    he = half_edges(faces)
    directed = {(v, w) for v, w, _, _ in he}
    # per vertex, outgoing half-edges in corner order
    rows = defaultdict(list)
    for v, w, f, t in he:
        rows[v].append((t, w))
    rank = {}
    k = 0
    for v in sorted(rows):
        for t, w in sorted(rows[v]):
            twin = (w, v) in directed
            if (twin and v < w) or not twin:
                rank[(min(v, w), max(v, w))] = k
                k += 1
    return rank


def refined_topology(faces, nv, rank) -> np.ndarray:
    nf = len(faces)
    base = nv + nf
    out = []
    for f, fv in enumerate(faces):
        n = len(fv)
        for k in range(n):
            v, w, u = fv[k], fv[(k + 1) % n], fv[(k - 1) % n]
            out.append([v, base + rank[(min(v, w), max(v, w))], nv + f, base + rank[(min(u, v), max(u, v))]])
    return np.asarray(out, dtype=np.int64).reshape(-1)


def cc_textbook(faces, P: np.ndarray, rank) -> np.ndarray:
    """Refined positions [vertex points | face points | edge points] in FP64."""
    nv = P.shape[0]
    nf = len(faces)
    P = P[:, :3].astype(np.float64)
    fp = np.array([P[fv].mean(axis=0) for fv in faces])
    he = half_edges(faces)
    face_of = {}
    for v, w, f, _ in he:
        face_of[(v, w)] = f
    ne = len(rank)
    ep = np.zeros((ne, 3))
    boundary = {}
    for (a, b), e in rank.items():
        f0 = face_of.get((a, b))
        f1 = face_of.get((b, a))
        if f0 is not None and f1 is not None:
            ep[e] = (P[a] + P[b] + fp[f0] + fp[f1]) / 4.0
        else:
            ep[e] = (P[a] + P[b]) / 2.0
            boundary[(a, b)] = True
    vp = np.zeros((nv, 3))
    adj_faces = defaultdict(list)
    nbrs = defaultdict(set)
    for v, w, f, _ in he:
        adj_faces[v].append(f)
        nbrs[v].add(w)
        nbrs[w].add(v)
    bnd_nbrs = defaultdict(list)
    for (a, b) in boundary:
        bnd_nbrs[a].append(b)
        bnd_nbrs[b].append(a)
    for v in range(nv):
        if bnd_nbrs[v]:
            vp[v] = 0.75 * P[v] + 0.125 * sum(P[x] for x in bnd_nbrs[v])
            continue
        n = len(adj_faces[v])
        if n == 0:
            vp[v] = np.nan
            continue
        Q = np.mean([fp[f] for f in adj_faces[v]], axis=0)
        R = np.mean([(P[v] + P[x]) / 2.0 for x in nbrs[v]], axis=0)
        vp[v] = (Q + 2.0 * R + (n - 3.0) * P[v]) / n
    return np.concatenate([vp, fp, ep], axis=0)


def dense_mapped_products(faces, nv):
    """Dense F, D (column v, row w) by the reference predicate
    (spgemm.cuh:62-69, bhsparse ESC value = f+1) and E from them."""
    F = np.zeros((nv, nv), dtype=np.int64)  # F[w, v]
    D = np.zeros((nv, nv), dtype=np.int64)
    Pat = np.zeros((nv, nv), dtype=bool)
    for f, fv in enumerate(faces):
        n = len(fv)
        for kv in range(n):
            pv = kv + 1
            for kw in range(n):
                pw = kw + 1
                v, w = fv[kv], fv[kw]
                Pat[w, v] = True
                if pw == (pv % n) + 1:
                    F[w, v] += f + 1
                if pw == ((pv + n - 2) % n) + 1:
                    D[w, v] += f + 1
    E = np.zeros((nv, nv), dtype=np.int64)
    k = 0
    for c in range(nv):
        for r in range(c):
            if not Pat[r, c]:
                continue
            if F[r, c] or D[r, c]:
                k += 1
                E[r, c] = -k if (F[r, c] and D[r, c]) else k
    return F, D, E


def dense_to_csc(A):
    ptr = [0]
    ids, vals = [], []
    for c in range(A.shape[1]):
        nz = np.nonzero(A[:, c])[0]
        ids.extend(nz.tolist())
        vals.extend(A[nz, c].tolist())
        ptr.append(len(ids))
    return np.array(ptr), np.array(ids, dtype=np.int64), np.array(vals, dtype=np.int64)


# ---------------------------------------------------------------------------
# regular closed grids: Catmull-Clark == tensor-product cubic B-spline

def bspline_refine(G: np.ndarray) -> np.ndarray:
    """G[A, B, d] periodic in both axes -> [2A, 2B, d]."""

    def axis(X, ax):
        Xm = np.roll(X, 1, axis=ax)
        Xp = np.roll(X, -1, axis=ax)
        even = (Xm + 6.0 * X + Xp) / 8.0
        odd = (X + Xp) / 2.0
        shape = list(X.shape)
        shape[ax] *= 2
        out = np.empty(shape, dtype=X.dtype)
        sl_e = [slice(None)] * X.ndim
        sl_o = [slice(None)] * X.ndim
        sl_e[ax] = slice(0, None, 2)
        sl_o[ax] = slice(1, None, 2)
        out[tuple(sl_e)] = even
        out[tuple(sl_o)] = odd
        return out

    return axis(axis(G, 0), 1)


def _step(d, size):
    d = d % size
    return np.where(d == size - 1, -1, d)


def refine_grid_coords(coords: np.ndarray, quads: np.ndarray, refined_ids: np.ndarray, A: int, B: int,
                       nv_new: int) -> np.ndarray:
    """Grid coordinates (on the 2A x 2B grid) of every refined vertex of one
    Catmull-Clark level of a closed quad grid mesh.

    coords[v] = (i, j) of coarse vertex v on the periodic A x B grid, quads[f]
    its 4 corners (a unit cell, in face order), refined_ids the refined quads
    (4 per coarse corner: v, e_out, fp, e_in). Face point: 2 c0 + step(c0->c1)
    + step(c0->c3); edge point of corner k: 2 c_k + step(c_k -> c_{k+1})."""
    size = np.array([A, B])
    c = coords[quads]  # [nf, 4, 2]
    nv, nf = coords.shape[0], quads.shape[0]
    out = np.full((nv_new, 2), -1, dtype=np.int64)
    out[:nv] = 2 * coords
    s01 = _step(c[:, 1] - c[:, 0], size)
    s03 = _step(c[:, 3] - c[:, 0], size)
    out[nv:nv + nf] = (2 * c[:, 0] + s01 + s03) % (2 * size)
    e_out = refined_ids.reshape(-1, 4)[:, 1].reshape(nf, 4)
    nxt = np.roll(c, -1, axis=1)
    ecoord = (2 * c + _step(nxt - c, size)) % (2 * size)
    out[e_out.reshape(-1)] = ecoord.reshape(-1, 2)
    return out


def torus_coords(a: int, b: int) -> np.ndarray:
    i = np.tile(np.arange(a), b)
    j = np.repeat(np.arange(b), a)
    return np.stack([i, j], 1)
