"""Per-stage timing (slak_profile_enable / slak_profile_report, the
counterpart of TimeProfileIndividual, LibInclude/profiling.h:182-200): the
stages of one refinement, their launch counts and their algorithmic bytes,
written out from the per-kernel byte model (compulsory reads + writes of the
distinct arrays each launch touches, int32 indices, float4 positions)."""
import pytest

from slak_amd import capi, meshgen as g
from slak_amd.mesh import subdivide_host

pytestmark = pytest.mark.gpu


def expected_cfg1_lak():
    # 32 x 16 closed quad torus: sizes of the input of every level
    nv0, nf0, nnz0, ne0 = 512, 512, 2048, 1024
    nv1, nnz1, ne1 = nv0 + nf0 + ne0, 4 * nnz0, 2 * nnz0
    nf1 = nnz0
    nv2, nf2, nnz2, ne2 = nv1 + nf1 + ne1, nnz1, 4 * nnz1, 2 * nnz1
    return {
        # first level: validation + position-order gather
        "gather": (1, 4 * (nf0 + 1) + 12 * nnz0),
        # valence count, the two scans (valences, owned edges), M -> M^T
        "count_quad": (1, 4 * nnz0 + 8 * nv0),
        "scan": (2, 2 * (8 * nv0 + 4)),
        "transpose_quad": (1, 8 * nnz0 + 12 * nv0),
        # mapped SpGEMM row sizing, then edge numbering with the face points
        "spgemm_symbolic_quad": (1, 16 * nnz0 + 8 * nv0),
        "number_facepoints_quad": (1, 12 * nv0 + 20 * nnz0 + 4 * ne0 + 4 * nnz0 + 16 * nv0 + 16 * nf0),
        # level 1: refine + vertex items, writing level 1's structure; only
        # the last level's kernel reads it, so its M^T rows are corners alone
        # (4 refined corners x (8 B (edge, twin) + 4 B row entry))
        "refine_vertex_quad": (1, (20 * nnz0 + 20 * nv0 + 16 * nf0)
                               + (16 * nnz0 + 16 * ne0 + 16 * nv0 + 48 * nnz0 + 4 * (nv0 + nf0 + ne0))),
        # level 2: positions only, from the input's structure and level 1's positions
        "refine_vertex_pos_quad": (1, (20 * nnz0 + 4 * nv0 + 16 * nv1) + 16 * (nv1 + nf1 + ne1)),
        # level 3: from level 1's structure (ids, (edge, twin), row corners)
        # and level 2's positions, with quads and LAK form
        "refine_vertex_final_quad": (1, (16 * nnz1 + 4 * nv1 + 16 * nv2) + 16 * (nv2 + nf2 + ne2)
                                     + 16 * nnz2 + 4 * (nnz2 + 1) + 16 * nnz2),
    }


def test_profile_report_cfg1_lak():
    m = g.config1()
    subdivide_host(m, 3, capi.SLAK_SCHEME_LAK)  # warm-up
    capi.profile_enable(True)
    try:
        out, _ = subdivide_host(m, 3, capi.SLAK_SCHEME_LAK)
        rep = capi.profile_report()
    finally:
        capi.profile_enable(False)
    assert out.nfaces == 32 * 16 * 64
    got = {s["name"]: (s["launches"], int(round(s["bytes"]))) for s in rep["stages"]}
    assert got == expected_cfg1_lak()
    assert all(s["total_ms"] > 0 for s in rep["stages"])


def test_profile_disabled_records_nothing():
    capi.profile_enable(True)
    capi.profile_enable(False)
    subdivide_host(g.cube(), 2, capi.SLAK_SCHEME_LAK)
    assert capi.profile_report()["stages"] == []
