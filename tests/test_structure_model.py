"""The closed forms by which a level writes the next level's structure (see
tests/structure_model.py) agree with brute force on the oracle's refined
meshes, over two refinements, for closed, open, quad and polygonal inputs."""
import numpy as np
import pytest

from slak_amd import meshgen as g
from tests import structure_model as M

MESHES = [
    ("cube", lambda: g.cube()),
    ("tetrahedron", lambda: g.tetrahedron()),
    ("quad", lambda: g.single_quad()),
    ("open_grid", lambda: g.open_grid(5, 3, sigma=0.01)),
    ("cylinder", lambda: g.cylinder(7, 3)),
    ("mixed_pent", lambda: g.torus_mixed(12, 8, 0.3, pentagons=True, seed=7)),
    ("mixed_tri", lambda: g.torus_mixed(9, 7, 0.5, seed=3)),
]


def _same(a: M.Structure, b: M.Structure):
    np.testing.assert_array_equal(a.ids, b.ids)
    assert a.nv == b.nv
    np.testing.assert_array_equal(a.eout, b.eout)
    np.testing.assert_array_equal(a.twin, b.twin)
    np.testing.assert_array_equal(a.bnd, b.bnd)
    np.testing.assert_array_equal(a.tptr, b.tptr)
    assert a.rows == b.rows
    np.testing.assert_array_equal(a.mpos, b.mpos)


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
@pytest.mark.parametrize("scheme", [M.LAK, M.SLAK], ids=["LAK", "SLAK"])
def test_closed_forms_match_brute_force(oracle, name, make, scheme):
    mesh = make()
    ptr = mesh.face_ptr()
    S = M.structure_of(ptr, mesh.ids, mesh.nverts, scheme)
    for level in (1, 2):
        S = M.refine_structure(S, scheme)
        ref = oracle.subdivide(mesh, level, scheme=scheme, precision=64)
        # the refined faces are the oracle's (bit-exact numbering)
        np.testing.assert_array_equal(S.ids, ref.ids)
        brute = M.structure_of(np.arange(ref.nf + 1) * 4, ref.ids, ref.nv, scheme)
        _same(S, brute)



QUAD_MESHES = [
    ("cube", lambda: g.cube()),
    ("quad", lambda: g.single_quad()),
    ("open_grid", lambda: g.open_grid(5, 3, sigma=0.01)),
    ("cylinder", lambda: g.cylinder(7, 3)),
    ("mixed_pent_l1", lambda: g.torus_mixed(12, 8, 0.3, pentagons=True, seed=7)),
]


@pytest.mark.parametrize("name,make", QUAD_MESHES, ids=[m[0] for m in QUAD_MESHES])
@pytest.mark.parametrize("scheme", [M.LAK, M.SLAK], ids=["LAK", "SLAK"])
def test_final_closed_form_matches_two_refinements(oracle, name, make, scheme):
    """The last level computed from the quad mesh M two levels back (output
    quads, edge point slots, the M1 rows of face points and edge points)
    equals the one computed from the brute-force structure of M1."""
    mesh = make()
    if mesh.type != g.QUAD:  # M = the first refinement
        r = oracle.subdivide(mesh, 1, scheme=scheme, precision=64)
        ptr, ids, nv = np.arange(r.nf + 1) * 4, r.ids, r.nv
    else:
        ptr, ids, nv = mesh.face_ptr(), mesh.ids, mesh.nverts
    S = M.structure_of(ptr, ids, nv, scheme)
    ff = M.final_closed_form(S, scheme)
    S1 = M.refine_structure(S, scheme)
    brute = M.structure_of(np.arange(len(S1.ids) // 4 + 1) * 4, S1.ids, S1.nv, scheme)
    nnz1 = len(S1.ids)
    # output quads: (vertex, ep(out-edge), fp(face), ep(in-edge)) of every M1 corner
    exp = np.zeros(4 * nnz1, dtype=np.int64)
    epb2 = S1.nv + nnz1 // 4
    for c in range(nnz1):
        pc = (c & ~3) | ((c + 3) & 3)
        exp[4 * c: 4 * c + 4] = [S1.ids[c], epb2 + brute.eout[c], S1.nv + (c >> 2), epb2 + brute.eout[pc]]
    np.testing.assert_array_equal(ff.ids2, exp)
    # every M1 edge point written exactly once, each at the id of the M1 edge
    # it belongs to: v-half = out-edge of 4t, w-half = out-edge of 4 next(t) + 3,
    # edge to fp(F) = out-edge of 4t + 1, edge to fp(G) = out-edge of 4 s1 + 1
    assert sorted(ff.slots) == list(range(len(S1.bnd)))
    for e, (t, name) in ff.slots.items():
        tn = (t & ~3) | ((t + 1) & 3)
        c = {"vh": 4 * t, "wh": 4 * tn + 3, "fF": 4 * t + 1, "fG": 4 * int(S.twin[t]) + 1}[name]
        assert int(brute.eout[c]) == e, (name, t)
    # face point and edge point rows in M^T order
    rows = {}
    for vtx in range(S1.nv):
        rows[vtx] = brute.rows[int(brute.tptr[vtx]):int(brute.tptr[vtx + 1])]
    nv0, nf0 = S.nv, len(S.ptr) - 1
    for F, r in enumerate(ff.fp_rows):
        assert rows[nv0 + F] == r
    owners = [t for t in range(len(S.ids)) if S.twin[t] < 0 or t < S.twin[t]]
    for t, r in zip(owners, ff.ep_rows):
        assert rows[nv0 + nf0 + int(S.eout[t])] == r
