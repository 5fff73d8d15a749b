"""Open meshes: both schemes apply the boundary rule of the reference with no
corner rule (LAKCatmullClarkImpl.cu:191-198,223-235; CatmullClarkKernels.cuh:
200-403,688-740): a boundary vertex moves to 3/4 P + 1/8 of its two boundary
neighbours and a boundary edge point is the midpoint. Along a boundary loop
that is exactly cubic B-spline curve subdivision, which these tests use as a
closed-form check at any depth and size, besides the oracle on small cases."""
import numpy as np
import pytest

from slak_amd import capi, meshgen as g
from slak_amd.mesh import subdivide_host

pytestmark = pytest.mark.gpu

LAK, SLAK = capi.SLAK_SCHEME_LAK, capi.SLAK_SCHEME_SLAK
POS_TOL = 1e-5


def boundary_loops(ptr, ids):
    """Boundary loops as vertex sequences along the boundary half-edges (the
    half-edges v -> w of a face with no w -> v); each loop starts at its
    smallest vertex id."""
    ptr = np.asarray(ptr, dtype=np.int64)
    ids = np.asarray(ids, dtype=np.int64)
    order = np.diff(ptr)
    t = np.arange(len(ids))
    if np.all(order == 4):
        succ = (t & ~3) | ((t + 1) & 3)
    else:
        face = np.repeat(np.arange(len(order)), order)
        succ = ptr[face] + (t - ptr[face] + 1) % order[face]
    v, w = ids, ids[succ]
    n = int(ids.max()) + 1 if len(ids) else 0
    key = np.sort(v * n + w)
    rkey = w * n + v
    i = np.minimum(np.searchsorted(key, rkey), len(key) - 1)
    bnd = key[i] != rkey
    nxt = dict(zip(v[bnd].tolist(), w[bnd].tolist()))
    assert len(nxt) == int(bnd.sum()), "non-manifold boundary vertex"
    loops, seen = [], set()
    for v0 in sorted(nxt):
        if v0 in seen:
            continue
        loop, x = [], v0
        while x not in seen:
            seen.add(x)
            loop.append(x)
            x = nxt[x]
        assert x == v0
        loops.append(loop)
    return loops


def bspline_closed(P, levels):
    """Cubic B-spline subdivision of a closed polygon (float64): vertex points
    (p[i-1] + 6 p[i] + p[i+1]) / 8, edge points (p[i] + p[i+1]) / 2."""
    P = np.asarray(P, dtype=np.float64)
    for _ in range(levels):
        prev, nxt = np.roll(P, 1, axis=0), np.roll(P, -1, axis=0)
        Q = np.empty((2 * len(P), P.shape[1]))
        Q[0::2] = (prev + 6 * P + nxt) / 8
        Q[1::2] = (P + nxt) / 2
        P = Q
    return P


def check_loops(mesh, levels, scheme):
    out, _ = subdivide_host(mesh, levels, scheme)
    loops0 = boundary_loops(mesh.face_ptr(), mesh.ids)
    nf = len(out.ids) // 4
    loops = boundary_loops(np.arange(nf + 1) * 4, out.ids)
    assert len(loops) == len(loops0)
    # original vertices keep their ids: match the refined loop starting at the
    # same vertex as each original loop
    by_start = {lp[0]: lp for lp in loops}
    diag = mesh.bbox_diag()
    for lp0 in loops0:
        lp = by_start[lp0[0]]
        assert len(lp) == len(lp0) << levels
        ref = bspline_closed(mesh.verts[lp0, :3], levels)
        err = np.abs(out.verts[lp, :3].astype(np.float64) - ref).max() / diag
        assert err <= POS_TOL, err
    return out


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
@pytest.mark.parametrize("levels", [1, 2, 3, 4])
@pytest.mark.parametrize("make", [g.single_quad, lambda: g.open_grid(5, 3, sigma=0.02, seed=3),
                                  lambda: g.cylinder(9, 4)], ids=["quad", "open_grid", "cylinder"])
def test_boundary_loops_are_bspline_curves(make, levels, scheme):
    check_loops(make(), levels, scheme)


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_boundary_deep_levels_match_oracle(oracle, scheme):
    """Five levels of an open grid with noise and of a cylinder: the boundary
    variants of every kernel, the last level's included."""
    for mesh in (g.open_grid(4, 3, sigma=0.02, seed=5), g.cylinder(6, 2)):
        out, _ = subdivide_host(mesh, 5, scheme)
        ref = oracle.subdivide(mesh, 5, scheme=scheme, precision=64)
        np.testing.assert_array_equal(out.ids, ref.ids)
        err = np.abs(out.verts[:, :3].astype(np.float64) - ref.P[:, :3]).max() / mesh.bbox_diag()
        assert err <= POS_TOL, err


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_flat_open_grid_full_size(scheme):
    """The 256 x 128 open grid refined 4 times (8.4 M faces): the patch stays
    exactly in its plane and inside its square (every rule is a convex
    combination), with the closed-form face count."""
    mesh = g.open_grid(256, 128)
    out, _ = subdivide_host(mesh, 4, scheme)
    assert np.all(out.verts[:, 2] == 0.0)
    assert out.verts[:, :2].min() >= 0.0 and out.verts[:, :2].max() <= 1.0
    assert np.all(out.verts[:, 3] == 1.0)
    assert len(out.ids) == 4 * 256 * 128 * 4 ** 4
