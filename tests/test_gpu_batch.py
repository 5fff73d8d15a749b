"""Batched subdivision (slak_batch_*) against per-mesh calls and the oracle.

The batched chain runs the same kernels over a union of meshes, so each
mesh's result must equal its single-mesh result bit for bit, graph or not,
replay after replay."""
import numpy as np
import pytest

from slak_amd import capi, meshgen as g
from slak_amd.batch import Batch
from slak_amd.capi import SlakError
from slak_amd.mesh import MeshInfo, subdivide_host, to_host

pytestmark = pytest.mark.gpu

LAK, SLAK = capi.SLAK_SCHEME_LAK, capi.SLAK_SCHEME_SLAK


def fetch(m: MeshInfo):
    ids = to_host(m.ids, m.nnz, np.int32)
    verts = to_host(m.verts, 4 * m.nverts, np.float32).reshape(-1, 4)
    ptr = to_host(m.ptr, m.nfaces + 1, np.int32) if m.ptr else None
    vals = to_host(m.vals, m.nnz, np.int32) if m.vals else None
    return ids, verts, ptr, vals


def mixed_batch(n, seed):
    rng = np.random.default_rng(seed)
    meshes, levels = [], []
    for i in range(n):
        kind = i % 5
        if kind == 0:
            m = g.torus(int(rng.integers(3, 12)), int(rng.integers(3, 10)), seed=seed, mesh=i)
        elif kind == 1:
            m = g.torus_mixed(int(rng.integers(4, 14)), int(rng.integers(3, 9)), 0.3, pentagons=True, seed=seed, mesh=i)
        elif kind == 2:
            m = g.open_grid(int(rng.integers(1, 6)), int(rng.integers(1, 5)), sigma=0.01, seed=seed, mesh=i)
        elif kind == 3:
            m = g.cube() if i % 2 else g.tetrahedron()
        else:
            m = g.torus_mixed(int(rng.integers(4, 10)), int(rng.integers(4, 8)), 0.5, seed=seed, mesh=i)
        meshes.append(m)
        levels.append(int(rng.integers(0, 4)))
    return meshes, levels


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_batch_equals_single_mesh_calls(oracle, scheme, monkeypatch):
    """Mixed depths 0-3: the meshes whose last level is formed from the quad
    mesh two levels back are taken out of that level's union and formed
    together in one launch; the results must not depend on it."""
    hosts, levels = mixed_batch(40, seed=3)
    dev = [MeshInfo.from_host(h) for h in hosts]
    try:
        monkeypatch.setenv("SLAK_FUSED_FINAL", "0")
        with Batch(dev, levels, scheme) as b:
            b.run(use_graph=False)
            chain = [None if r.ids == 0 else fetch(r) for r in (b.result(i) for i in range(len(hosts)))]
        monkeypatch.delenv("SLAK_FUSED_FINAL")
        with Batch(dev, levels, scheme) as b:
            b.run(use_graph=False)
            for i, (h, lv) in enumerate(zip(hosts, levels)):
                r = b.result(i)
                if h.type == g.QUAD and lv == 0:
                    assert r.ids == 0  # untouched, as subdivideIteratively leaves it
                    continue
                ids, verts, ptr, vals = fetch(r)
                for got, want in zip((ids, verts, ptr, vals), chain[i]):
                    np.testing.assert_array_equal(got, want, err_msg=f"mesh {i}")
                single, _ = subdivide_host(h, lv, scheme)
                np.testing.assert_array_equal(ids, single.ids, err_msg=f"mesh {i}")
                np.testing.assert_array_equal(verts, single.verts, err_msg=f"mesh {i}")
                if scheme == LAK:
                    np.testing.assert_array_equal(ptr, single.ptr)
                    np.testing.assert_array_equal(vals, single.vals)
                else:
                    assert ptr is None and vals is None
                if i % 7 == 0:
                    ref = oracle.subdivide(h, lv, scheme=scheme)
                    np.testing.assert_array_equal(ids, ref.ids)
                    assert np.abs(verts - ref.P).max() / h.bbox_diag() <= 1e-5
    finally:
        for d in dev:
            d.freeAndReset()


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
@pytest.mark.parametrize("depth,quad_only", [(2, True), (3, False), (3, True), (4, True)],
                         ids=["d2", "d3_poly", "d3_quad", "d4_quad"])
@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
def test_same_depth_batch_last_level_in_one_launch(oracle, scheme, depth, quad_only, use_graph, monkeypatch):
    """Every mesh refined to the same level: the batch's last level is formed
    for all meshes at once from the quad meshes two levels back (several
    meshes per corner block, open and closed, polygonal inputs at depth 3),
    and with quad inputs at depth 3 and 4 the level before it too (positions
    only, from the inputs' own structure at depth 3). Each mesh must equal its
    single-mesh result bit for bit."""
    hosts, _ = mixed_batch(30, seed=11)
    if quad_only:
        hosts = [h for h in hosts if h.type == g.QUAD]
    if depth == 4:  # keep the outputs small
        hosts = hosts[:8]
    hosts = hosts + [g.open_grid(3, 2, sigma=0.01, seed=5, mesh=99), g.torus(5, 4, seed=6, mesh=98)]
    levels = [depth] * len(hosts)
    dev = [MeshInfo.from_host(h) for h in hosts]
    try:
        with Batch(dev, levels, scheme) as b:
            for rep in range(2 if use_graph else 1):
                b.run(use_graph=use_graph)
                for i, h in enumerate(hosts):
                    ids, verts, ptr, vals = fetch(b.result(i))
                    single, _ = subdivide_host(h, depth, scheme)
                    np.testing.assert_array_equal(ids, single.ids, err_msg=f"mesh {i}")
                    np.testing.assert_array_equal(verts, single.verts, err_msg=f"mesh {i}")
                    if scheme == LAK:
                        np.testing.assert_array_equal(ptr, single.ptr, err_msg=f"mesh {i}")
                        np.testing.assert_array_equal(vals, single.vals, err_msg=f"mesh {i}")
                    if rep == 0 and i % 6 == 0:
                        ref = oracle.subdivide(h, depth, scheme=scheme)
                        np.testing.assert_array_equal(ids, ref.ids)
                        assert np.abs(verts - ref.P).max() / h.bbox_diag() <= 1e-5
            fused = [fetch(b.result(i)) for i in range(len(hosts))]
        if not use_graph:  # the level-by-level chain gives the same bits
            monkeypatch.setenv("SLAK_FUSED_FINAL", "0")
            with Batch(dev, levels, scheme) as b:
                b.run(use_graph=False)
                for i in range(len(hosts)):
                    for got, want in zip(fetch(b.result(i)), fused[i]):
                        np.testing.assert_array_equal(got, want, err_msg=f"mesh {i}")
    finally:
        for d in dev:
            d.freeAndReset()


def test_graph_replays_match_eager_and_track_new_positions():
    hosts, levels = mixed_batch(25, seed=8)
    levels = [max(1, x) for x in levels]
    dev = [MeshInfo.from_host(h) for h in hosts]
    try:
        with Batch(dev, levels, LAK) as b:
            b.run(use_graph=False)
            eager = [fetch(b.result(i)) for i in range(len(hosts))]
            for rep in range(3):
                b.run(use_graph=True)
                # eager work between replays
                subdivide_host(g.cube(), 1, LAK)
                for i in range(len(hosts)):
                    got = fetch(b.result(i))
                    np.testing.assert_array_equal(got[0], eager[i][0])
                    np.testing.assert_array_equal(got[1], eager[i][1])
                    # the LAK form is written once, before the capture: replays keep it
                    nf = len(got[0]) // 4
                    np.testing.assert_array_equal(got[2], 4 * np.arange(nf + 1, dtype=np.int32))
                    np.testing.assert_array_equal(got[3], np.tile(np.arange(1, 5, dtype=np.int32), nf))
            # new positions, same topology: the replay recomputes from the inputs
            for d, h in zip(dev, hosts):
                moved = (h.verts.copy())
                moved[:, :3] *= 2.0
                capi.check(capi.lib().slak_memcpy_h2d(d.verts, moved.ctypes.data, moved.nbytes))
            b.run(use_graph=True)
            for i, h in enumerate(hosts):
                moved = g.HostMesh(h.type, h.ptr, h.ids, h.vals, h.verts.copy())
                moved.verts[:, :3] *= 2.0
                single, _ = subdivide_host(moved, levels[i], LAK)
                got = fetch(b.result(i))
                np.testing.assert_array_equal(got[1], single.verts)
    finally:
        for d in dev:
            d.freeAndReset()


def test_config4_sample_batch_matches_per_mesh():
    spec = g.config4(n_meshes=48, seed=4)
    dev = [MeshInfo.from_host(h) for h in spec.meshes]
    try:
        with Batch(dev, spec.levels, LAK) as b:
            b.run(use_graph=True)
            inf = b.info()
            assert inf.nmeshes == 48 and inf.max_levels == max(spec.levels)
            tot = 0
            for i, (h, lv) in enumerate(zip(spec.meshes, spec.levels)):
                ids, verts, _, _ = fetch(b.result(i))
                single, _ = subdivide_host(h, lv, LAK)
                np.testing.assert_array_equal(ids, single.ids)
                np.testing.assert_array_equal(verts, single.verts)
                tot += single.nfaces
            assert inf.out_faces == tot
    finally:
        for d in dev:
            d.freeAndReset()


@pytest.mark.parametrize("seed", [21, 22, 23, 24])
def test_random_batches_equal_single_mesh_calls(seed):
    """Random compositions (tori, polygonal tori with pentagons, open grids,
    cubes, tetrahedra; depths 0-4), both schemes, eager and graph: the
    several-mesh paths (work-unit tables, LAK form in its own launch, rows
    read by the last level alone written as corners) give each mesh its
    single-mesh result bit for bit."""
    rng = np.random.default_rng(seed)
    hosts, _ = mixed_batch(int(rng.integers(12, 36)), seed=seed)
    levels = [int(rng.integers(0, 5)) for _ in hosts]
    for scheme in (LAK, SLAK):
        dev = [MeshInfo.from_host(h) for h in hosts]
        try:
            with Batch(dev, levels, scheme) as b:
                b.run(use_graph=bool(seed % 2))
                b.run(use_graph=bool(seed % 2))
                for i, (h, lv) in enumerate(zip(hosts, levels)):
                    r = b.result(i)
                    if h.type == g.QUAD and lv == 0:
                        assert r.ids == 0
                        continue
                    ids, verts, ptr, vals = fetch(r)
                    single, _ = subdivide_host(h, lv, scheme)
                    np.testing.assert_array_equal(ids, single.ids, err_msg=f"mesh {i} level {lv}")
                    np.testing.assert_array_equal(verts, single.verts, err_msg=f"mesh {i} level {lv}")
                    if scheme == LAK:
                        np.testing.assert_array_equal(ptr, single.ptr, err_msg=f"mesh {i} level {lv}")
                        np.testing.assert_array_equal(vals, single.vals, err_msg=f"mesh {i} level {lv}")
        finally:
            for d in dev:
                d.freeAndReset()


def test_batch_reports_the_invalid_mesh():
    hosts = [g.cube(), g.tetrahedron(), g.torus(6, 5)]
    bad = g.HostMesh(hosts[2].type, hosts[2].ptr, hosts[2].ids.copy(), hosts[2].vals, hosts[2].verts)
    bad.ids[7] = 10_000
    hosts[2] = bad
    dev = [MeshInfo.from_host(h) for h in hosts]
    try:
        with pytest.raises(SlakError) as ei:
            Batch(dev, [1, 1, 1], LAK)
        assert "mesh 2" in str(ei.value)
    finally:
        for d in dev:
            d.freeAndReset()


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_graph_replay_rewrites_every_output(scheme):
    """Every output of the batch (ids, positions and the LAK form's ptr and
    vals, whichever path forms a mesh's last level) is overwritten by each
    graph replay: poison them all between replays and the next replay must
    restore the first replay's bits."""
    hosts, levels = mixed_batch(40, seed=21)
    hosts += [g.torus(6, 5, seed=2, mesh=90), g.torus(7, 4, seed=3, mesh=91)]
    levels += [3, 4]
    dev = [MeshInfo.from_host(h) for h in hosts]
    try:
        with Batch(dev, levels, scheme) as b:
            b.run(use_graph=True)
            live = [i for i in range(len(hosts)) if b.result(i).ids != 0]
            first = {i: fetch(b.result(i)) for i in live}
            for rep in range(2):
                for i in live:
                    r = b.result(i)
                    for ptr, n in ((r.ids, r.nnz), (r.verts, 4 * r.nverts), (r.ptr, r.nfaces + 1), (r.vals, r.nnz)):
                        if ptr:
                            junk = np.full(n, -7, dtype=np.int32)
                            capi.check(capi.lib().slak_memcpy_h2d(ptr, junk.ctypes.data, junk.nbytes))
                b.run(use_graph=True)
                for i in live:
                    for got, want in zip(fetch(b.result(i)), first[i]):
                        if want is None:
                            assert got is None
                        else:
                            np.testing.assert_array_equal(got, want, err_msg=f"mesh {i}, replay {rep}")
    finally:
        for d in dev:
            d.freeAndReset()
