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

