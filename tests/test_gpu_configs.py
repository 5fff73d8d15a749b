"""Parity at the sizes and depths the five workload configurations name.

Bar: refined topology bit-exact against the FP64 oracle's canonical LAK /
SLAK numbering (ids, and ptr/vals in LAK form); FP32 positions within
POS_TOL of the control cage's bounding-box diagonal. Batched and served
results are also compared bit for bit with the same mesh refined alone
(slak_subdivide), through device-side digests where the outputs are too large
to copy back (reference driver: LAKCatmullClarkImpl.cu:512-571,
CatmullClarkImpl.cu:457-515)."""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from slak_amd import capi, meshgen as g
from slak_amd.batch import Batch
from slak_amd.mesh import MeshInfo, _subdivide, spgemm_mapped_host, subdivide_host, to_host
from slak_amd.serve import Server

pytestmark = pytest.mark.gpu

LAK, SLAK = capi.SLAK_SCHEME_LAK, capi.SLAK_SCHEME_SLAK
POS_TOL = 1e-5  # max |p_gpu - p_fp64| / bbox diagonal
WORKERS = max(1, min(16, os.cpu_count() or 1))


def fetch(m: MeshInfo):
    ids = to_host(m.ids, m.nnz, np.int32)
    verts = to_host(m.verts, 4 * m.nverts, np.float32).reshape(-1, 4)
    ptr = to_host(m.ptr, m.nfaces + 1, np.int32) if m.ptr else None
    vals = to_host(m.vals, m.nnz, np.int32) if m.vals else None
    return ids, verts, ptr, vals


def digests(m: MeshInfo):
    """Device digests of a refined mesh's arrays (ids, verts, ptr, vals)."""
    out = [capi.digest(m.ids, 4 * m.nnz), capi.digest(m.verts, 16 * m.nverts)]
    out.append(capi.digest(m.ptr, 4 * (m.nfaces + 1)) if m.ptr else None)
    out.append(capi.digest(m.vals, 4 * m.nnz) if m.vals else None)
    return tuple(out)


def check_oracle(oracle, host, levels, scheme, got, label, nthreads=1):
    """got = (ids, verts, ptr, vals) of the GPU; returns max_err / diag."""
    ids, verts, ptr, vals = got
    ref = oracle.subdivide(host, levels, scheme=scheme, precision=64, nthreads=nthreads)
    assert len(ids) == len(ref.ids) and verts.shape[0] == ref.nv, label
    np.testing.assert_array_equal(ids, ref.ids, err_msg=label)
    if scheme == LAK:
        np.testing.assert_array_equal(ptr, ref.ptr, err_msg=label)
        np.testing.assert_array_equal(vals, ref.vals, err_msg=label)
    else:
        assert ptr is None and vals is None, label
    err = float(np.abs(verts.astype(np.float64) - ref.P).max()) / host.bbox_diag()
    assert err <= POS_TOL, (label, err)
    assert np.all(verts[:, 3] == 1.0), label
    return err


# ---------------------------------------------------------------------------
# cfg2: 256 x 128 mixed-polygon torus at its full 4 levels


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_config2_full_depth_vs_oracle(oracle, scheme):
    m = g.config2()
    out, st = subdivide_host(m, 4, scheme)
    assert st.levels == 4 and out.nfaces == m.nnz * 4 ** 3
    got = (out.ids, out.verts, out.ptr, out.vals)
    err = check_oracle(oracle, m, 4, scheme, got, "cfg2", nthreads=WORKERS)
    print(f"cfg2 4 levels scheme={scheme}: faces={out.nfaces} max_err/diag={err:.3e}")


# ---------------------------------------------------------------------------
# cfg4: the batched uneven workload


def config4_sample(n):
    spec = g.config4(n_meshes=n, seed=4)
    return spec.meshes, spec.levels


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_config4_sample_graph_vs_oracle(oracle, scheme):
    """The first 256 meshes of the 4096-mesh batch (every depth 1-4, grid
    dims up to 64), refined in one batch replayed as a hipGraph: every mesh
    against the oracle, on the capture run and on a replay."""
    hosts, levels = config4_sample(256)
    assert set(levels) == {1, 2, 3, 4}
    dims = [int(x) for h in hosts for x in h.label.split("torus")[1].split("_")[0].split("x")]
    assert max(dims) >= 60 and min(dims) <= 8
    dev = [MeshInfo.from_host(h) for h in hosts]
    errs = []
    try:
        with Batch(dev, levels, scheme) as b:
            b.run(use_graph=True)  # capture
            first = [digests(b.result(i)) for i in range(len(hosts))]
            b.run(use_graph=True)  # replay
            results = [fetch(b.result(i)) for i in range(len(hosts))]
            for i in range(len(hosts)):
                assert digests(b.result(i)) == first[i], f"mesh {i}: replay differs from the capture run"

        def one(i):
            return check_oracle(oracle, hosts[i], levels[i], scheme, results[i], f"cfg4 mesh {i}")

        with ThreadPoolExecutor(WORKERS) as ex:
            errs = list(ex.map(one, range(len(hosts))))
    finally:
        for d in dev:
            d.freeAndReset()
    print(f"cfg4 sample of {len(hosts)} scheme={scheme}: max_err/diag={max(errs):.3e}")


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_config4_full_batch_equals_per_mesh(scheme):
    """All 4096 meshes: the batch (graph replay) equals one slak_subdivide per
    mesh, mesh by mesh, by device digests of every output array."""
    hosts, levels = config4_sample(4096)
    dev = [MeshInfo.from_host(h) for h in hosts]
    try:
        with Batch(dev, levels, scheme) as b:
            b.run(use_graph=True)
            b.run(use_graph=True)
            batched = [digests(b.result(i)) for i in range(len(hosts))]
            faces = int(b.info().out_faces)
        total = 0
        for i, (d, lv) in enumerate(zip(dev, levels)):
            r = MeshInfo()
            st = _subdivide(scheme, d, r, lv)
            total += int(st.out_faces)
            try:
                assert digests(r) == batched[i], f"mesh {i}"
            finally:
                r.freeAndReset()
        assert total == faces
        print(f"cfg4 full batch scheme={scheme}: 4096 meshes, {faces} faces, batched == per-mesh")
    finally:
        for d in dev:
            d.freeAndReset()


# ---------------------------------------------------------------------------
# cfg5: the served stream, up to the 1e6-face end of the request sizes


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_config5_requests_alone_equal_in_mix(oracle, scheme):
    """The first 40 requests of the cfg5 stream (seed 5): among them
    requests 3 (949 K faces) and 24 (573 K faces) at level 3. Submitted with
    staggered arrivals, each result equals the request refined alone, bit
    for bit; the largest request and every fifth one are checked against the
    oracle."""
    n = 40
    reqs = [g.request_mesh(5, i) for i in range(n)]
    big = [i for i in range(n) if reqs[i][0].nfaces >= 5e5 and reqs[i][1] == 3]
    assert big, "no request of >= 5e5 faces at level 3"
    dev = [MeshInfo.from_host(m) for m, _ in reqs]
    try:
        with Server(scheme, max_batch=64) as srv:
            rids = []
            for i, (d, (m, lv)) in enumerate(zip(dev, reqs)):
                rids.append(srv.submit(d, lv))
                if i % 4 == 3:
                    time.sleep(0.003)
            served = [srv.wait(r, timeout_ms=120000) for r in rids]
            assert srv.stats().max_batch > 1, "requests were never coalesced"
        check = sorted(set(big[:1] + list(range(0, n, 5))))
        errs = {}
        for i, ((r, rs), d, (m, lv)) in enumerate(zip(served, dev, reqs)):
            assert rs.status == 0 and rs.levels == lv, i
            alone = MeshInfo()
            _subdivide(scheme, d, alone, lv)
            try:
                assert digests(r) == digests(alone), f"request {i}: served differs from alone"
            finally:
                alone.freeAndReset()
            if i in check:
                errs[i] = check_oracle(oracle, m, lv, scheme, fetch(r), f"cfg5 request {i}", nthreads=WORKERS)
            r.freeAndReset()
        print(f"cfg5 scheme={scheme}: {n} requests alone == in mix; oracle max_err/diag="
              + ", ".join(f"{i}:{e:.2e}" for i, e in errs.items()))
    finally:
        for d in dev:
            d.freeAndReset()


# ---------------------------------------------------------------------------
# high-valence rows (valence 9..24) and long faces (order 12): the mapped
# SpGEMM's in-memory row path, at both register capacities


HIGH_VALENCE = [
    ("uv_sphere24", lambda: g.uv_sphere(24, 6)),
    ("uv_sphere9", lambda: g.uv_sphere(9, 4)),
    ("capped_cyl12", lambda: g.capped_cylinder(12, 3)),
    ("fan_disk16", lambda: g.fan_disk(16)),
    ("mix", lambda: g.high_valence_mix()),
]


@pytest.mark.parametrize("cap", ["4", "8"])
@pytest.mark.parametrize("name,make", HIGH_VALENCE, ids=[h[0] for h in HIGH_VALENCE])
@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_high_valence_vs_oracle(oracle, scheme, name, make, cap, monkeypatch):
    monkeypatch.setenv("SLAK_ROW_CAP", cap)
    m = make()
    for levels in (1, 2, 3):
        out, _ = subdivide_host(m, levels, scheme)
        check_oracle(oracle, m, levels, scheme, (out.ids, out.verts, out.ptr, out.vals), f"{name} L{levels}")


ROW_MESHES = [
    ("tri_torus", lambda: g.torus_mixed(40, 30, 1.0, seed=4)),  # valence 6 everywhere
    ("uv_sphere24", lambda: g.uv_sphere(24, 6)),
    ("mix", lambda: g.high_valence_mix()),
    ("mixed_torus", lambda: g.torus_mixed(48, 32, 0.25, pentagons=True, seed=5)),
]


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_rows_per_block_choices_agree(oracle, scheme, monkeypatch):
    """The first level's row kernels take 64, 128 or 256 rows per block (the
    plan picks one from the row lengths): every choice gives the same bits,
    alone and in a batch replayed as a graph, and the oracle's topology."""
    hosts = [make() for _, make in ROW_MESHES]
    base = [subdivide_host(h, 2, scheme)[0] for h in hosts]
    for i, h in enumerate(hosts):
        check_oracle(oracle, h, 2, scheme, (base[i].ids, base[i].verts, base[i].ptr, base[i].vals), ROW_MESHES[i][0])
    for rows in ("64", "128", "256"):
        monkeypatch.setenv("SLAK_ROWS_PER_BLOCK", rows)
        for i, h in enumerate(hosts):
            out, _ = subdivide_host(h, 2, scheme)
            for got, want in zip((out.ids, out.verts, out.ptr, out.vals), (base[i].ids, base[i].verts, base[i].ptr, base[i].vals)):
                if want is None:
                    assert got is None
                else:
                    np.testing.assert_array_equal(got, want, err_msg=f"{ROW_MESHES[i][0]} rows {rows}")
        dev = [MeshInfo.from_host(h) for h in hosts]
        try:
            with Batch(dev, [2] * len(hosts), scheme) as b:
                for _ in range(2):
                    b.run(use_graph=True)
                    for i in range(len(hosts)):
                        ids, verts, ptr, vals = fetch(b.result(i))
                        np.testing.assert_array_equal(ids, base[i].ids, err_msg=f"batch {i} rows {rows}")
                        np.testing.assert_array_equal(verts, base[i].verts, err_msg=f"batch {i} rows {rows}")
        finally:
            for d in dev:
                d.freeAndReset()


@pytest.mark.parametrize("cap", ["4", "8"])
def test_high_valence_spgemm_operator(oracle, cap, monkeypatch):
    """F, D and E of the mapped SpGEMM operator on rows of up to 24 corners."""
    monkeypatch.setenv("SLAK_ROW_CAP", cap)
    for _, make in HIGH_VALENCE:
        m = make()
        r = spgemm_mapped_host(m)
        d = oracle.subdivide(m, 1, scheme=0, keep_debug=True).debug[0]
        for nm in ("F", "D", "E"):
            for part in ("ptr", "ids", "vals"):
                np.testing.assert_array_equal(getattr(r, f"{nm}_{part}"), d[f"{nm}_{part}"], err_msg=f"{nm}_{part}")
        assert r.nedges == d["nedges"] and r.nextern == d["nextern"]


@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_high_valence_batch(oracle, scheme):
    """High-valence meshes inside a batch with ordinary ones."""
    hosts = [g.uv_sphere(24, 5), g.torus(6, 5), g.capped_cylinder(12, 2), g.fan_disk(16), g.cube()]
    levels = [3, 2, 3, 2, 1]
    dev = [MeshInfo.from_host(h) for h in hosts]
    try:
        with Batch(dev, levels, scheme) as b:
            b.run(use_graph=False)
            for i, (h, lv) in enumerate(zip(hosts, levels)):
                check_oracle(oracle, h, lv, scheme, fetch(b.result(i)), f"batch mesh {i}")
    finally:
        for d in dev:
            d.freeAndReset()
