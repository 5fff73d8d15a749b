"""Serving with continuous batching: every result equals the same request
refined alone (bitwise), whatever else was in flight; admission control and
error isolation."""
import threading
import time

import numpy as np
import pytest

from slak_amd import capi, meshgen as g
from slak_amd.capi import SlakError
from slak_amd.mesh import MeshInfo, subdivide_host, to_host
from slak_amd.serve import Server

pytestmark = pytest.mark.gpu

LAK, SLAK = capi.SLAK_SCHEME_LAK, capi.SLAK_SCHEME_SLAK


def fetch(m):
    ids = to_host(m.ids, m.nnz, np.int32)
    verts = to_host(m.verts, 4 * m.nverts, np.float32).reshape(-1, 4)
    return ids, verts


@pytest.mark.parametrize("step_corners", [None, 300000], ids=["one-plan-per-step", "split-steps"])
@pytest.mark.parametrize("scheme", [LAK, SLAK], ids=["LAK", "SLAK"])
def test_alone_equals_in_mix(scheme, step_corners, monkeypatch):
    """A step whose union would exceed the per-plan index range runs as several
    plans (forced here with a small SLAK_SERVE_STEP_CORNERS); results are the
    same either way."""
    if step_corners:
        monkeypatch.setenv("SLAK_SERVE_STEP_CORNERS", str(step_corners))
    rng = np.random.default_rng(5)
    hosts, levels = [], []
    for i in range(30):
        F = int(np.exp(rng.uniform(np.log(100), np.log(20000))))
        b = max(4, int(round(np.sqrt(F / 2))))
        a = max(4, int(round(F / b)))
        hosts.append(g.torus_mixed(a, b, 0.1, seed=11, mesh=i))
        levels.append(int(rng.integers(1, 4)))
    dev = [MeshInfo.from_host(h) for h in hosts]
    try:
        with Server(scheme, max_batch=64) as srv:
            ids = []
            # staggered arrivals: later requests join while earlier ones are in flight
            for i, (d, lv) in enumerate(zip(dev, levels)):
                ids.append(srv.submit(d, lv))
                if i % 5 == 4:
                    time.sleep(0.002)
            results = [srv.wait(r, timeout_ms=60000) for r in ids]
            st = srv.stats()
            assert st.completed == len(hosts)
            assert st.max_batch > 1, "requests were never coalesced"
        for i, ((r, rs), h, lv) in enumerate(zip(results, hosts, levels)):
            assert rs.status == 0 and rs.levels == lv and rs.latency_us > 0
            got_ids, got_verts = fetch(r)
            alone, _ = subdivide_host(h, lv, scheme)
            np.testing.assert_array_equal(got_ids, alone.ids, err_msg=f"request {i}")
            np.testing.assert_array_equal(got_verts, alone.verts, err_msg=f"request {i}")
            if scheme == LAK:
                np.testing.assert_array_equal(to_host(r.ptr, r.nfaces + 1, np.int32), alone.ptr)
                np.testing.assert_array_equal(to_host(r.vals, r.nnz, np.int32), alone.vals)
            else:
                assert r.ptr == 0 and r.vals == 0
            r.freeAndReset()
    finally:
        for d in dev:
            d.freeAndReset()


def test_invalid_request_fails_alone():
    good = [g.torus(8, 6, seed=1, mesh=i) for i in range(3)]
    bad = g.HostMesh(good[0].type, good[0].ptr, good[0].ids.copy(), good[0].vals, good[0].verts)
    bad.ids[2] = 99999
    dev = [MeshInfo.from_host(h) for h in good + [bad]]
    try:
        with Server(LAK) as srv:
            ids = [srv.submit(d, 2) for d in dev]
            out = []
            for r in ids[:3]:
                m, st = srv.wait(r, timeout_ms=60000)
                out.append(m)
            with pytest.raises(SlakError) as ei:
                srv.wait(ids[3], timeout_ms=60000)
            assert ei.value.code == 2
        for m, h in zip(out, good):
            alone, _ = subdivide_host(h, 2, LAK)
            np.testing.assert_array_equal(fetch(m)[0], alone.ids)
            m.freeAndReset()
    finally:
        for d in dev:
            d.freeAndReset()


def test_memory_budget_admission():
    h = g.torus(64, 32)
    d = MeshInfo.from_host(h)
    try:
        # a budget that holds one request at a time: all still complete, one by one
        with Server(LAK, memory_budget_bytes=4 << 20) as srv:
            ids = [srv.submit(d, 2) for _ in range(4)]
            res = [srv.wait(r, timeout_ms=60000) for r in ids]
            assert srv.stats().max_batch == 1
            for m, st in res:
                m.freeAndReset()
        with Server(LAK, memory_budget_bytes=1 << 20) as srv:
            with pytest.raises(SlakError) as ei:
                srv.submit(d, 3)
            assert ei.value.code == 6
    finally:
        d.freeAndReset()


def test_concurrent_submitters():
    hosts = [g.torus_mixed(10 + i, 8, 0.2, seed=3, mesh=i) for i in range(16)]
    dev = [MeshInfo.from_host(h) for h in hosts]
    out = {}
    try:
        with Server(LAK) as srv:
            def client(k):
                for i in range(k, len(dev), 4):
                    rid = srv.submit(dev[i], 2)
                    out[i] = srv.wait(rid, timeout_ms=60000)
            th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
            for t in th:
                t.start()
            for t in th:
                t.join()
        for i, h in enumerate(hosts):
            alone, _ = subdivide_host(h, 2, LAK)
            np.testing.assert_array_equal(fetch(out[i][0])[1], alone.verts)
            out[i][0].freeAndReset()
    finally:
        for d in dev:
            d.freeAndReset()
