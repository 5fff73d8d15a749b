"""OBJ I/O and the command-line driver with the reference demo's interface."""
import os
import subprocess

import numpy as np
import pytest

from slak_amd import meshgen as g
from slak_amd.objio import read_obj, write_obj

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "slak_amd", "lib", "slak_cli")


def test_obj_round_trip_and_type_detection(tmp_path):
    for make, typ in ((g.cube, g.QUAD), (g.tetrahedron, g.TRI), (lambda: g.torus_mixed(8, 6, 0.3, pentagons=True), g.POLY)):
        m = make()
        p = tmp_path / "m.obj"
        write_obj(m, str(p))
        r = read_obj(str(p))
        assert r.type == typ
        np.testing.assert_array_equal(r.ptr, m.face_ptr())
        np.testing.assert_array_equal(r.ids, m.ids)
        np.testing.assert_array_equal(r.vals, m.vals)
        assert np.abs(r.verts - m.verts).max() < 1e-5 * max(1.0, np.abs(m.verts).max())


def test_obj_negative_and_slashed_indices(tmp_path):
    p = tmp_path / "q.obj"
    p.write_text("# quad\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
                 "f -4/1/1 -3/1/1 -2//1 -1\n")
    r = read_obj(str(p))
    assert r.type == g.QUAD and list(r.ids) == [0, 1, 2, 3]


def test_cli_parse_only_matches_python_reader(tmp_path):
    assert os.path.exists(CLI), "build with python -m slak_amd.build"
    m = g.torus_mixed(12, 8, 0.3, pentagons=True, seed=4)
    p = tmp_path / "mixed.obj"
    write_obj(m, str(p))
    out = subprocess.run([CLI, "--parse-only", str(p)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    r = read_obj(str(p))
    assert f"type={r.type} nf={r.nfaces} nnz={r.nnz} nv={r.nverts} max_face_order={r.max_face_order}" in out.stdout


def _parse_only(path):
    out = subprocess.run([CLI, "--parse-only", str(path)], capture_output=True, text=True, timeout=60)
    return out


def test_obj_reader_elements_known_answer(tmp_path):
    """OBJ::read (include/obj.h of this repository, the reference's
    include/obj.h:107-253): normals with z negated, texture and normal indices
    resolved (1-based, negative = relative), tags, skipped '#', 'g', blank and
    unknown lines. Expected values follow from the format rules by hand."""
    p = tmp_path / "e.obj"
    p.write_text("# comment\n"
                 "g group\n"
                 "\n"
                 "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 0.5 1\n"
                 "vt 0 0\nvt 1 0\nvt 1 1\n"
                 "vn 0 0 1\nvn 0.25 -0.5 -2\n"
                 "usemtl stone\n"
                 "f 1/1/1 2/2/1 3/3/2 4//2\n"
                 "f -5/-3/-2 -4/-2/-1 5\n"
                 "t crease 2/1/0 3 4 0.75\n")
    out = _parse_only(p)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert "type=2 nf=2 nnz=7 nv=5 max_face_order=4 v=5 vt=3 vn=2 f_vti=5 f_vni=6 tags=1" in lines[0]
    assert lines[1:3] == ["vn 0 0 -1", "vn 0.25 -0.5 2"]
    vti = [int(l.split()[1]) for l in lines if l.startswith("vti ")]
    vni = [int(l.split()[1]) for l in lines if l.startswith("vni ")]
    assert vti == [0, 1, 2, 0, 1]               # 4//2 has none; -3, -2 with 3 vt read -> 0, 1
    assert vni == [0, 0, 1, 1, 0, 1]            # 4//2 -> 1; -2, -1 with 2 vn read -> 0, 1
    assert [l for l in lines if l.startswith("tag ")] == ["tag crease i3 i4 f0.75"]


@pytest.mark.parametrize("faces,typ", [("f 1 2 3\nf 1 3 4\n", 0), ("f 1 2 3 4\n", 1), ("f 1 2 3\nf 1 2 3 4\n", 2)],
                         ids=["tri", "quad", "mixed"])
def test_obj_reader_face_type(tmp_path, faces, typ):
    p = tmp_path / "t.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n" + faces)
    out = _parse_only(p)
    assert out.returncode == 0 and f"type={typ} " in out.stdout


def test_obj_reader_rejects_degenerate_faces(tmp_path):
    p = tmp_path / "d.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nf 1 2\n")
    out = _parse_only(p)
    assert out.returncode != 0 and "less than three vertices" in out.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("scheme,tag", [(0, "LAK"), (1, "SLAK")])
def test_cli_refines_and_writes_obj(oracle, tmp_path, scheme, tag):
    m = g.torus_mixed(16, 10, 0.3, pentagons=True, seed=6)
    p = tmp_path / "cage.obj"
    write_obj(m, str(p))
    cage = read_obj(str(p))  # what the CLI sees (6 significant digits)
    out = subprocess.run([CLI, str(p), str(scheme), "2", "1"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "Subdivision to level 2" in out.stdout
    res = read_obj(str(tmp_path / f"cage_2_{tag}.obj"))
    ref = oracle.subdivide(cage, 2, scheme=scheme)
    np.testing.assert_array_equal(res.ids, ref.ids)
    assert res.type == g.QUAD and res.nverts == ref.nv
    assert np.abs(res.verts[:, :3] - ref.P[:, :3]).max() < 1e-5 * max(1.0, np.abs(ref.P).max())


@pytest.mark.gpu
def test_cli_config_file(tmp_path):
    a, b = tmp_path / "a.obj", tmp_path / "b.obj"
    write_obj(g.cube(), str(a))
    write_obj(g.tetrahedron(), str(b))
    cfg = tmp_path / "run.cfg"
    cfg.write_text(f"# mesh scheme level save\n{a} 0 2 1\n{b} 1 1 1\n")
    out = subprocess.run([CLI, str(cfg)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert read_obj(str(tmp_path / "a_2_LAK.obj")).nfaces == 6 * 16
    assert read_obj(str(tmp_path / "b_1_SLAK.obj")).nfaces == 12


@pytest.mark.gpu
@pytest.mark.parametrize("make", [g.cube, lambda: g.torus(9, 6, seed=2), lambda: g.torus_mixed(10, 6, 0.3, pentagons=True, seed=5)],
                         ids=["cube", "torus", "mixed_poly"])
def test_cscmesh_container(oracle, tmp_path, make):
    """GPU::CSCMesh: OBJ construction, copies, unreduce/reduce round trip,
    MeshInfo conversion, LAK refinement and toObj (tests/cpp/mesh_test.cpp)."""
    prog = os.path.join(ROOT, "slak_amd", "lib", "slak_mesh_test")
    m = make()
    p, o = tmp_path / "in.obj", tmp_path / "out.obj"
    write_obj(m, str(p))
    out = subprocess.run([prog, str(p), str(o)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "mesh_test ok" in out.stdout, out.stdout + out.stderr
    res, ref = read_obj(str(o)), oracle.subdivide(read_obj(str(p)), 2, scheme=0)
    np.testing.assert_array_equal(res.ids, ref.ids)
    assert np.abs(res.verts[:, :3] - ref.P[:, :3]).max() < 1e-5 * max(1.0, np.abs(ref.P).max())
