"""Reduced / unreduced forms of homogeneous meshes (GPU::CSCMesh::reduce /
unreduce, include/Mesh.h:305-377) and homogeneous TRI inputs.

reduce reorders each face's (ids, vals) by vals: ids end in corner-position
order. unreduce fills a missing ptr (stride offsets) and vals (1-based
positions) and reorders each face by ids. Round trips restore the reduced
ids; LAK refinement reads vals, so both forms refine to the same bits."""
import numpy as np
import pytest

from slak_amd import capi, meshgen as g
from slak_amd.capi import SlakError
from slak_amd.mesh import MeshInfo, subdivide_host, to_host

pytestmark = pytest.mark.gpu

LAK, SLAK = capi.SLAK_SCHEME_LAK, capi.SLAK_SCHEME_SLAK
POS_TOL = 1e-5


def permuted(h: g.HostMesh, seed: int) -> g.HostMesh:
    """The same faces stored in a shuffled corner order per face, with vals
    recording each stored corner's position (an unreduced mesh)."""
    rng = np.random.default_rng(seed)
    ptr = h.face_ptr()
    ids = np.empty_like(h.ids)
    vals = np.empty_like(h.ids)
    for f in range(h.nfaces):
        s, e = int(ptr[f]), int(ptr[f + 1])
        perm = rng.permutation(e - s)
        ids[s:e] = h.ids[s:e][perm]
        vals[s:e] = perm + 1
    return g.HostMesh(h.type, ptr.astype(np.int32), ids, vals, h.verts)


def fetch(m: MeshInfo):
    ids = to_host(m.ids, m.nnz, np.int32)
    vals = to_host(m.vals, m.nnz, np.int32) if m.vals else None
    ptr = to_host(m.ptr, m.nfaces + 1, np.int32) if m.ptr else None
    return ptr, ids, vals


@pytest.mark.parametrize("make", [g.cube, g.tetrahedron, lambda: g.torus(9, 7, seed=2)],
                         ids=["cube", "tetrahedron", "torus"])
def test_reduce_puts_ids_in_position_order(make):
    h = make()
    u = permuted(h, 11)
    m = MeshInfo.from_host(u)
    m.is_reduced = False
    try:
        m.reduce()
        assert m.is_reduced
        ptr, ids, vals = fetch(m)
        np.testing.assert_array_equal(ids, h.ids)  # the builders store position order
        order = h.max_face_order
        np.testing.assert_array_equal(vals, np.tile(np.arange(1, order + 1, dtype=np.int32), h.nfaces))
        m.reduce()  # already reduced: unchanged
        np.testing.assert_array_equal(fetch(m)[1], h.ids)
    finally:
        m.freeAndReset()


def test_unreduce_then_reduce_round_trips_the_slak_form():
    """A SLAK result (ids only, is_reduced) unreduces to ptr = 4i, ids sorted
    per face with their positions in vals, and reduces back to itself."""
    out, _ = subdivide_host(g.torus_mixed(10, 6, 0.3, seed=4), 2, SLAK)
    assert out.ptr is None and out.vals is None
    m = MeshInfo.from_host(out)
    assert m.is_reduced
    try:
        m.unreduce()
        assert not m.is_reduced and m.ptr and m.vals
        ptr, ids, vals = fetch(m)
        np.testing.assert_array_equal(ptr, 4 * np.arange(out.nfaces + 1, dtype=np.int32))
        q = out.ids.reshape(-1, 4)
        ords = np.argsort(q, axis=1, kind="stable")
        np.testing.assert_array_equal(ids.reshape(-1, 4), np.take_along_axis(q, ords, axis=1))
        np.testing.assert_array_equal(vals.reshape(-1, 4), ords + 1)
        m.reduce()
        np.testing.assert_array_equal(fetch(m)[1], out.ids)
    finally:
        m.freeAndReset()


@pytest.mark.parametrize("make", [g.cube, lambda: g.torus(8, 5, seed=3)], ids=["cube", "torus"])
def test_lak_refines_both_forms_to_the_same_bits(make):
    h = make()
    red, _ = subdivide_host(h, 2, LAK)
    m = MeshInfo.from_host(h)
    m.is_reduced = True  # the builders store ids in position order
    try:
        m.unreduce()  # ids sorted per face, vals carry the positions
        un = m.to_host()
        assert not np.array_equal(un.ids, h.ids)
        got, _ = subdivide_host(un, 2, LAK)
        np.testing.assert_array_equal(got.ids, red.ids)
        np.testing.assert_array_equal(got.verts, red.verts)
    finally:
        m.freeAndReset()


def test_poly_meshes_are_not_homogeneous():
    m = MeshInfo.from_host(g.torus_mixed(8, 6, 0.3, pentagons=True, seed=1))
    try:
        with pytest.raises(SlakError, match="not homogeneous"):
            m.unreduce()
        with pytest.raises(SlakError, match="not homogeneous"):
            m.reduce()
    finally:
        m.freeAndReset()


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_homogeneous_tri_torus(oracle, scheme):
    """An all-triangle torus is a TRI mesh: refined like the oracle, and its
    reduced form (no ptr, no vals) refines the same."""
    h = g.torus_tri(12, 8, 1.0, seed=6)
    assert h.type == g.TRI
    out, _ = subdivide_host(h, 2, scheme)
    ref = oracle.subdivide(h, 2, scheme=scheme, precision=64)
    np.testing.assert_array_equal(out.ids, ref.ids)
    assert float(np.abs(out.verts.astype(np.float64) - ref.P).max() / h.bbox_diag()) <= POS_TOL
    bare = g.HostMesh(g.TRI, None, h.ids, None, h.verts)
    out2, _ = subdivide_host(bare, 2, scheme)
    np.testing.assert_array_equal(out2.ids, out.ids)
    np.testing.assert_array_equal(out2.verts, out.verts)
