#!/usr/bin/env python3
"""Refined faces/s of SLAK's Catmull-Clark chain on one MI355X.

Headline workload: the batched uneven workload, 4096 independent seeded tori
(grid dims uniform in [4, 64]^2, 20 % of quads triangulated, depth uniform in
{1..4}, 437 M final faces). A step is one replay of the whole batched
level chain (one launch per kernel per level over the union of the meshes
still active at that level, captured once as a hipGraph) with the control
cages resident on the device. Prints one JSON line with the throughput, the
roofline of the dominant kernel (HIP events inside the library on the launch
stream, HBM counters from rocprofv3), the plain per-mesh form for the
speedup, the same batch in the other scheme, secondary legs (the
256 x 128 mixed torus at 4 levels, the
1024 x 1024 torus at 2 levels with per-stage GB/s, the served stream) and
the CPU baseline (the same LAK algorithm, FP32, on the host's cores) with the
per-config position error against the FP64 oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "refined faces/s + HBM GB/s; batched-vs-plain speedup; serve p50/p99 latency"
HBM_PEAK_GBPS = 8000.0

SINGLE = {
    "cfg1": ("32x16 closed quad torus, 3 levels", 3),
    "cfg2": ("256x128 mixed-polygon torus (p_tri=0.25, 1 pentagon per column), 4 levels", 4),
    "cfg3": ("1024x1024 closed quad torus, 2 levels", 2),
}
CFG4_DESC = "4096 seeded tori, dims U[4,64]^2, 20% of quads triangulated, depth U{1..4}, batched hipGraph replay"
# the stage and kernel that dominate each workload's time
CFG4_STAGE = "refine_vertex_final_quad"


def make_single(name):
    from slak_amd import meshgen

    return {"cfg1": meshgen.config1, "cfg2": meshgen.config2, "cfg3": meshgen.config3}[name]()


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def stage_roofline(prof, stage=None):
    """Roofline of one stage of a library profile (the slowest one by
    default): its algorithmic bytes per launch over its mean HIP-event launch
    duration, against the 8 TB/s HBM peak; plus the whole chain's figure."""
    stages = prof["stages"]
    total_ms = sum(s["total_ms"] for s in stages)
    total_b = sum(s["bytes"] for s in stages)
    dom = max(stages, key=lambda s: s["total_ms"]) if stage is None else next(s for s in stages if s["name"] == stage)
    ms_per_launch = dom["total_ms"] / dom["launches"]
    bytes_per_launch = dom["bytes"] / dom["launches"]
    achieved = bytes_per_launch / (ms_per_launch * 1e-3) / 1e9
    return {
        "bound": "hbm",
        "kernel": dom["name"],
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4),
        "traffic": None,
        "bytes_per_launch": int(bytes_per_launch),
        "us_per_launch": round(ms_per_launch * 1e3, 2),
        "launches": int(dom["launches"]),
        "chain_gbps": round(total_b / (total_ms * 1e-3) / 1e9, 1),
        "chain_frac": round(total_b / (total_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
    }


def output_level_roofline(prof):
    """The several-mesh output level as one unit: k_final (refine_vertex_final_quad)
    and the LAK-form launch that follows it (lak_form), their algorithmic
    bytes over their summed HIP-event durations. None when no LAK form is
    written separately (SLAK, or one mesh)."""
    st = {s["name"]: s for s in prof["stages"]}
    if CFG4_STAGE not in st or "lak_form" not in st:
        return None
    a, b = st[CFG4_STAGE], st["lak_form"]
    n = a["launches"]
    byts = (a["bytes"] + b["bytes"]) / n
    ms = (a["total_ms"] + b["total_ms"]) / n
    gbps = byts / (ms * 1e-3) / 1e9
    return {"stages": [CFG4_STAGE, "lak_form"], "bytes_per_launch": int(byts), "us_per_launch": round(ms * 1e3, 2),
            "achieved": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4)}


def stage_table(prof):
    return [{"name": s["name"], "launches": s["launches"], "us": round(s["total_ms"] * 1e3, 1),
             "bytes": int(s["bytes"]), "gbps": round(s["bytes"] / (s["total_ms"] * 1e-3) / 1e9, 1)}
            for s in prof["stages"]]


# ---------------------------------------------------------------------------
# GPU legs


def time_single(mesh, levels, scheme, steps, warmup, profile_steps=3):
    """K full subdivideIteratively calls through the C-ABI on the
    device-resident control cage (output allocation and free included)."""
    import ctypes

    from slak_amd import capi
    from slak_amd.mesh import MeshInfo

    c = MeshInfo.from_host(mesh)
    try:
        lib = capi.lib()
        cin = c.to_c()
        cout = capi.SlakMesh()
        stats = capi.SlakStats()
        pin, pout, pst = ctypes.byref(cin), ctypes.byref(cout), ctypes.byref(stats)

        def step():
            capi.check(lib.slak_subdivide(scheme, pin, pout, levels, pst, None))
            for p in (cout.ptr, cout.ids, cout.vals, cout.verts):
                if p:
                    capi.check(lib.slak_free(p))
            cout.ptr = cout.ids = cout.vals = cout.verts = None
            return stats

        for _ in range(warmup):
            step()
        capi.check(lib.slak_device_synchronize())
        t0 = time.perf_counter()
        for _ in range(steps):
            st = step()
        capi.check(lib.slak_device_synchronize())
        secs = time.perf_counter() - t0
        prof = None
        if profile_steps:
            capi.profile_enable(True)
            for _ in range(profile_steps):
                step()
            prof = capi.profile_report()
            capi.profile_enable(False)
        return secs, int(st.out_faces), st, prof
    finally:
        c.freeAndReset()


def fetch_result(m):
    import numpy as np

    from slak_amd.mesh import to_host
    ids = to_host(m.ids, m.nnz, np.int32)
    verts = to_host(m.verts, 4 * m.nverts, np.float32).reshape(-1, 4)
    ptr = to_host(m.ptr, m.nfaces + 1, np.int32) if m.ptr else None
    vals = to_host(m.vals, m.nnz, np.int32) if m.vals else None
    return ids, verts, ptr, vals


def batched_leg(spec, scheme, steps, warmup, keep_results=64, plain=True):
    """cfg4: the plain form (one subdivideIteratively per mesh; skipped with
    plain=False) and the batched chain (graph replays). Returns the headline
    numbers, the profile of one eager run and the first keep_results meshes'
    outputs (the parity sample)."""
    from slak_amd import capi
    from slak_amd.batch import Batch
    from slak_amd.mesh import MeshInfo, _subdivide

    t0 = time.perf_counter()
    dev = [MeshInfo.from_host(h) for h in spec.meshes]
    capi.check(capi.lib().slak_device_synchronize())
    h2d_s = time.perf_counter() - t0
    h2d_bytes = int(sum(a.nbytes for h in spec.meshes for a in (h.ptr, h.ids, h.vals, h.verts) if a is not None))
    try:
        plain_s, faces = None, None
        if plain:
            for d, lv in list(zip(dev, spec.levels))[:64]:  # warm the per-mesh path
                r = MeshInfo()
                _subdivide(scheme, d, r, lv)
                r.freeAndReset()
            capi.check(capi.lib().slak_device_synchronize())
            t0 = time.perf_counter()
            faces = 0
            for d, lv in zip(dev, spec.levels):
                r = MeshInfo()
                st = _subdivide(scheme, d, r, lv)
                faces += int(st.out_faces)
                r.freeAndReset()
            capi.check(capi.lib().slak_device_synchronize())
            plain_s = time.perf_counter() - t0

        t0 = time.perf_counter()
        b = Batch(dev, spec.levels, scheme)
        create_s = time.perf_counter() - t0
        try:
            t0 = time.perf_counter()
            b.run(use_graph=True)  # capture + first replay
            first_s = time.perf_counter() - t0
            for _ in range(warmup):
                b.run(use_graph=True)
            capi.check(capi.lib().slak_device_synchronize())
            t0 = time.perf_counter()
            for _ in range(steps):
                b.run(use_graph=True)
            capi.check(capi.lib().slak_device_synchronize())
            secs = time.perf_counter() - t0
            inf = b.info()
            if faces is None:
                faces = int(inf.out_faces)
            assert int(inf.out_faces) == faces, "batched and per-mesh face counts differ"
            capi.profile_enable(True)
            b.run(use_graph=False)
            prof = capi.profile_report()
            capi.profile_enable(False)
            kept = [fetch_result(b.result(i)) for i in range(min(keep_results, len(dev)))]
        finally:
            b.close()
    finally:
        for d in dev:
            d.freeAndReset()
    return {
        "h2d_s": h2d_s, "h2d_bytes": h2d_bytes,
        "faces": faces, "secs": secs, "plain_s": plain_s, "create_s": create_s, "first_s": first_s,
        "device_bytes": int(inf.device_bytes), "host_syncs": int(inf.host_syncs), "prof": prof, "kept": kept,
    }


def serve_leg(n_requests, qps_fraction, scheme, seed=5, keep=()):
    """cfg5: Poisson arrivals of requests (torus, faces log-uniform in
    [100, 1e6], 10% triangles, level uniform in {1..3}); latency from submit
    until the refined mesh is resident. QPS is stated as a fraction of the
    saturation rate measured with every request submitted at once. The
    results of the requests in `keep` are copied back (parity sample)."""
    import numpy as np

    from slak_amd import meshgen
    from slak_amd.mesh import MeshInfo
    from slak_amd.serve import Server

    reqs = [meshgen.request_mesh(seed, i) for i in range(n_requests)]
    dev = [MeshInfo.from_host(m) for m, _ in reqs]
    kept = {}
    try:
        def run(qps, collect):
            rng = np.random.default_rng(seed)
            gaps = rng.exponential(1.0 / qps, size=n_requests) if qps else np.zeros(n_requests)
            lat, faces = [], 0
            with Server(scheme) as srv:
                t0 = time.perf_counter()
                ids = []
                t_next = t0
                for i, d in enumerate(dev):
                    t_next += gaps[i]
                    now = time.perf_counter()
                    if t_next > now:
                        time.sleep(t_next - now)
                    ids.append(srv.submit(d, reqs[i][1]))
                for i, rid in enumerate(ids):
                    r, st = srv.wait(rid)
                    lat.append(st.latency_us / 1e3)
                    faces += int(st.out_faces)
                    if collect and i in keep:
                        kept[i] = fetch_result(r)
                    r.freeAndReset()
                wall = time.perf_counter() - t0
                sst = srv.stats()
            lat = np.array(lat)
            return {
                "p50_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_ms": round(float(lat.mean()), 3),
                "faces_per_s": round(faces / wall, 1),
                "wall_s": round(wall, 3),
                "steps": int(sst.steps),
                "max_batch": int(sst.max_batch),
                "out_faces": faces,
            }

        run(0.0, False)  # warm-up
        sat = run(0.0, False)
        sat_qps = n_requests / sat["wall_s"]
        qps = qps_fraction * sat_qps
        loaded = run(qps, True)
    finally:
        for d in dev:
            d.freeAndReset()
    return {
        "workload": f"{n_requests} Poisson requests, faces log-uniform [100, 1e6], 10% triangles, level U{{1..3}}",
        "qps": round(qps, 2),
        "qps_fraction_of_saturation": qps_fraction,
        "saturation_qps": round(sat_qps, 2),
        **loaded,
        "saturation": sat,
    }, reqs, kept


# ---------------------------------------------------------------------------
# HBM counters (rocprofv3 in child processes, before this process touches the GPU)

# kernel of each roofline stage ({s}: 0 LAK, 1 SLAK; multi: the table path)
STAGE_KERNELS = {
    "refine_vertex_final_quad": r"k_final<{s}, \w+, {multi}, true>",
    "refine_vertex_pos_quad": r"k_final<{s}, \w+, {multi}, false>",
    "refine_vertex_quad": r"k_level<true, {s},",
    "refine_vertex_poly": r"k_level<false, {s},",
    "facepoints_quad": r"k_faces<true, true, false>",
    "facepoints_poly": r"k_faces<false, true, false>",
    "gather": r"k_gather",
    "count_quad": r"k_count<true>",
    "count_poly": r"k_count<false>",
    "lak_form": r"k_lak_form",
    "transpose_quad": r"k_transpose<true>",
    "transpose_poly": r"k_transpose<false>",
    "spgemm_symbolic_quad": r"k_symbolic<true, {s},",
    "spgemm_symbolic_poly": r"k_symbolic<false, {s},",
    "number_facepoints_quad": r"k_number_faces<true, {s},",
    "number_facepoints_poly": r"k_number_faces<false, {s},",
}
PMC_WORKLOADS = ("cfg4", "cfg2", "cfg3")


def pmc_child(scheme_name, n_meshes):
    """The profiled program of hbm_traffic(): the 1 GiB copy probe (the
    counters' calibration), then each workload of PMC_WORKLOADS (one warm-up
    and one measured run), separated by a 64 KiB copy probe as a marker."""
    import ctypes

    from slak_amd import capi, meshgen
    from slak_amd.batch import Batch
    from slak_amd.mesh import MeshInfo
    gbps = ctypes.c_double()
    scheme = capi.SLAK_SCHEME_LAK if scheme_name == "lak" else capi.SLAK_SCHEME_SLAK
    capi.check(capi.lib().slak_copy_probe(1 << 30, 1, ctypes.byref(gbps)))
    for w in PMC_WORKLOADS:
        capi.check(capi.lib().slak_copy_probe(1 << 16, 1, ctypes.byref(gbps)))  # marker
        if w == "cfg4":
            spec = meshgen.config4(n_meshes=n_meshes)
            dev = [MeshInfo.from_host(h) for h in spec.meshes]
            with Batch(dev, spec.levels, scheme) as b:
                b.run(use_graph=False)
                b.run(use_graph=False)
            for d in dev:
                d.freeAndReset()
        else:
            time_single(make_single(w), SINGLE[w][1], scheme, 1, 1, profile_steps=0)
        print(f"pmc child: {w} done", flush=True)


def hbm_traffic(scheme_name, n_meshes):
    """HBM bytes per dispatch of every kernel of each workload, from two
    rocprofv3 --pmc runs of pmc_child (FETCH_SIZE, then WRITE_SIZE), each
    scaled by what the counter reads for the 1 GiB copy probe (1 GiB read,
    1 GiB written). None when rocprofv3 is unavailable or a run fails."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    from collections import defaultdict

    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    out = tempfile.mkdtemp(prefix="slak_pmc_")
    counts = {}
    calib = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", counter.lower(), "--",
               sys.executable, os.path.abspath(__file__), "--pmc-child", "--scheme", scheme_name,
               "--batch-meshes", str(n_meshes)]
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, env=dict(os.environ))
        except subprocess.TimeoutExpired:
            return None
        if r.returncode != 0:
            return None
        rows = []
        for f in glob.glob(os.path.join(out, "**", f"{counter.lower()}_counter_collection.csv"), recursive=True):
            with open(f) as fh:
                rows += [row for row in csv.DictReader(fh) if row.get("Counter_Name") == counter]
        rows.sort(key=lambda row: int(row["Dispatch_Id"]))
        per = {w: defaultdict(list) for w in PMC_WORKLOADS}
        # the 1 GiB copies (a warm-up and the measured one) calibrate; each
        # run of 64 KiB marker copies (two dispatches) opens the next workload
        seg, in_marker = -1, False
        for row in rows:
            name = row["Kernel_Name"]
            for pre in ("void ", "slak::(anonymous namespace)::", "slak::"):
                name = name.replace(pre, "")
            name = name.split("(")[0]
            value = float(row["Counter_Value"]) * 1024.0
            if name == "k_copy_probe":
                if int(row.get("Grid_Size", 0)) == (1 << 30) // 16:
                    calib[counter] = value
                elif not in_marker:
                    seg += 1
                    in_marker = True
                continue
            in_marker = False
            if 0 <= seg < len(PMC_WORKLOADS):
                per[PMC_WORKLOADS[seg]][name].append(value)
        counts[counter] = per
    shutil.rmtree(out, ignore_errors=True)
    if "FETCH_SIZE" not in calib or "WRITE_SIZE" not in calib:
        return None
    copy = float(1 << 30)
    cal_f, cal_w = copy / calib["FETCH_SIZE"], copy / calib["WRITE_SIZE"]
    result = {"calibration": {"fetch": round(cal_f, 4), "write": round(cal_w, 4)}, "workloads": {}}
    for w in PMC_WORKLOADS:
        fe, wr = counts["FETCH_SIZE"][w], counts["WRITE_SIZE"][w]
        kernels = {}
        for name in set(fe) | set(wr):
            f, wv = fe.get(name, []), wr.get(name, [])
            kernels[name] = {"dispatches": max(len(f), len(wv)),
                             "read": cal_f * sum(f) / max(len(f), 1), "write": cal_w * sum(wv) / max(len(wv), 1)}
        result["workloads"][w] = kernels
    return result


def stage_traffic(traffic, workload, stage, scheme_name, multi):
    """Average calibrated HBM bytes (read, write) per dispatch of a stage's kernel."""
    import re

    if not traffic or stage not in STAGE_KERNELS:
        return None
    kernels = traffic["workloads"].get(workload, {})
    pat = re.compile(STAGE_KERNELS[stage].format(s=0 if scheme_name == "lak" else 1,
                                                  multi="true" if multi else "false"))
    sel = [v for k, v in kernels.items() if pat.match(k)]
    n = sum(v["dispatches"] for v in sel)
    if not n:
        return None
    rd = sum(v["read"] * v["dispatches"] for v in sel) / n
    wt = sum(v["write"] * v["dispatches"] for v in sel) / n
    return int(rd), int(wt)


def add_traffic(roof, traffic, workload, scheme_name, multi):
    t = stage_traffic(traffic, workload, roof["kernel"], scheme_name, multi)
    if t:
        roof["traffic"] = t[0] + t[1]
        roof["traffic_read"], roof["traffic_write"] = t
        roof["traffic_over_algorithmic"] = round((t[0] + t[1]) / roof["bytes_per_launch"], 3)
        roof["traffic_calibration"] = traffic["calibration"]
        roof["traffic_source"] = "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by the 1 GiB copy probe"
    return roof


# ---------------------------------------------------------------------------
# I/O floor


def io_floor_bytes(mesh, levels, lak_form):
    """Implementation-independent I/O floor of a chain: per level, read the
    level's ids (+ ptr + vals in LAK form) and positions, write the refined
    ids (+ ptr + vals) and positions; summed over the levels. Sizes follow
    the closed forms (n_f' = nnz, nnz' = 4 nnz, n_e = (nnz - n_ext)/2 + n_ext,
    n_v' = n_v + n_f + n_e, n_ext' = 2 n_ext)."""
    import numpy as np
    ptr = mesh.face_ptr().astype(np.int64)
    ids = mesh.ids.astype(np.int64)
    nxt = np.arange(len(ids)) + 1
    nxt[ptr[1:] - 1] = ptr[:-1]  # cyclic successor within each face
    a, b = ids, ids[nxt]
    key = a * (int(mesh.nverts) + 1) + b
    rev = b * (int(mesh.nverts) + 1) + a
    n_ext = int((~np.isin(rev, key)).sum())
    nv, nf, nnz = int(mesh.nverts), int(mesh.nfaces), int(len(ids))
    total = 0.0
    for _ in range(levels):
        ne = (nnz - n_ext) // 2 + n_ext
        nv2, nf2, nnz2 = nv + nf + ne, nnz, 4 * nnz
        rd = 4 * nnz + 16 * nv + ((4 * (nf + 1) + 4 * nnz) if lak_form else 0)
        wr = 4 * nnz2 + 16 * nv2 + ((4 * (nf2 + 1) + 4 * nnz2) if lak_form else 0)
        total += rd + wr
        nv, nf, nnz, n_ext = nv2, nf2, nnz2, 2 * n_ext
    return total


def io_floor(b, secs_per_step):
    gbps = b / secs_per_step / 1e9
    return {"bytes_per_step": int(b), "gbps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4),
            "note": "per level: read ids(+ptr+vals)+positions, write refined ids(+ptr+vals)+positions"}


# ---------------------------------------------------------------------------
# CPU baseline and the per-config position error (the FP64 oracle as checker)


def cpu_baseline(spec, scheme, budget_s):
    """The oracle's FP32 LAK algorithm on the host: the headline workload
    (cfg4, every mesh at its own depth) with the meshes spread over all host
    threads, and single-threaded on a bounded prefix of the same batch."""
    from oracle import pyoracle

    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
    threads = pyoracle.max_threads()
    msn, faces = pyoracle.time_cpu_batch(spec.meshes, spec.levels, threads, scheme=scheme, warmup=0, reps=1)
    # single thread: the prefix of the batch that fits about half the budget at the OpenMP rate / threads
    per_face_1 = msn * 1e-3 * threads / max(faces, 1)
    want = max(1.0, 0.5 * budget_s / per_face_1)
    n1, acc = 0, 0
    for h, lv in zip(spec.meshes, spec.levels):
        acc += h.nnz * 4 ** (max(lv, 1) - 1)  # final faces of the mesh
        n1 += 1
        if acc >= want:
            break
    ms1, faces1 = pyoracle.time_cpu_batch(spec.meshes[:n1], spec.levels[:n1], 1, scheme=scheme, warmup=0, reps=1)
    return {
        "value": round(faces / (msn * 1e-3), 1),
        "unit": "faces/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the full cfg4 batch ({len(spec.meshes)} meshes, {faces} final faces, each mesh at its own depth), "
                  f"FP32 LAK, meshes spread over {threads} OpenMP threads, one timed pass",
        "value_1thread": round(faces1 / (ms1 * 1e-3), 1),
        "sample_1thread": f"the first {n1} meshes of the same batch ({faces1} final faces), one thread",
        "cpu_model": cpu_model(),
        "host_cores": os.cpu_count(),
    }


def position_error(ref_P, verts, diag):
    import numpy as np
    return float(np.abs(verts.astype(np.float64) - ref_P).max()) / diag


def parity_check(oracle, host, levels, scheme, got):
    """Topology bit-exact and max |p - p_fp64| / bbox diagonal of one GPU result."""
    import numpy as np
    ids, verts, ptr, vals = got
    ref = oracle.subdivide(host, levels, scheme=scheme, precision=64, nthreads=oracle.max_threads())
    exact = len(ids) == len(ref.ids) and bool(np.array_equal(ids, ref.ids))
    if scheme == 0:
        exact = exact and ptr is not None and bool(np.array_equal(ptr, ref.ptr)) and bool(np.array_equal(vals, ref.vals))
    err = position_error(ref.P, verts, host.bbox_diag()) if verts.shape[0] == ref.nv else float("inf")
    return exact, err


def parity_per_config(scheme, cfg4_spec, cfg4_kept, cfg5_reqs, cfg5_kept):
    """The GPU result of every configuration against the FP64 oracle: cfg1-3
    refined again here, the first meshes of the cfg4 batch and the kept cfg5
    requests from their legs."""
    from oracle import pyoracle
    from slak_amd.mesh import subdivide_host

    out = {}
    for w, (_, levels) in SINGLE.items():
        m = make_single(w)
        r, _ = subdivide_host(m, levels, scheme)
        exact, err = parity_check(pyoracle, m, levels, scheme, (r.ids, r.verts, r.ptr, r.vals))
        out[w] = {"topology_bit_exact": exact, "max_err_over_diag": err, "sample": "full mesh"}
    if cfg4_kept:
        errs, ok = [], True
        for i, got in enumerate(cfg4_kept):
            e, err = parity_check(pyoracle, cfg4_spec.meshes[i], cfg4_spec.levels[i], scheme, got)
            ok &= e
            errs.append(err)
        out["cfg4"] = {"topology_bit_exact": ok, "max_err_over_diag": max(errs),
                       "sample": f"the first {len(cfg4_kept)} meshes of the batch, from the graph replay"}
    if cfg5_kept:
        errs, ok = [], True
        for i, got in sorted(cfg5_kept.items()):
            m, lv = cfg5_reqs[i]
            e, err = parity_check(pyoracle, m, lv, scheme, got)
            ok &= e
            errs.append(err)
        out["cfg5"] = {"topology_bit_exact": ok, "max_err_over_diag": max(errs),
                       "sample": f"served requests {sorted(cfg5_kept)} of the Poisson run"}
    for v in out.values():
        v["max_err_over_diag"] = float(f"{v['max_err_over_diag']:.3e}")
    return out


# ---------------------------------------------------------------------------


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scheme", default="lak", choices=["lak", "slak"])
    ap.add_argument("--batch-meshes", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle parity check")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of single-thread CPU baseline work")
    ap.add_argument("--no-legs", action="store_true", help="skip the cfg2 / cfg3 single-mesh legs")
    ap.add_argument("--no-plain", action="store_true", help="skip the per-mesh plain form of the batch")
    ap.add_argument("--leg", choices=["cfg2", "cfg3"], help="run only this single-mesh leg (profiling runs)")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-qps-fraction", type=float, default=0.5)
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM counter runs")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args.scheme, args.batch_meshes)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (meshes are independent and would be "
              "spread over GPUs by projected output faces; one huge mesh would be cut into face blocks "
              "with a one-ring halo). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.leg:  # one single-mesh chain, for kernel traces
        from slak_amd import capi
        scheme = capi.SLAK_SCHEME_LAK if args.scheme == "lak" else capi.SLAK_SCHEME_SLAK
        m = make_single(args.leg)
        lv = SINGLE[args.leg][1]
        secs, faces, st, prof = time_single(m, lv, scheme, args.steps, args.warmup)
        print(json.dumps({"leg": args.leg, "scheme": args.scheme, "faces_per_s": round(faces * args.steps / secs, 1),
                          "ms_per_chain": round(secs / args.steps * 1e3, 4), "stages": stage_table(prof)}), flush=True)
        return

    # HBM counters first, in child processes, before this one touches the GPU
    traffic = None if args.no_traffic else hbm_traffic(args.scheme, args.batch_meshes)

    import ctypes

    from slak_amd import capi, meshgen

    scheme = capi.SLAK_SCHEME_LAK if args.scheme == "lak" else capi.SLAK_SCHEME_SLAK
    lak_form = scheme == capi.SLAK_SCHEME_LAK
    copy_gbps = ctypes.c_double()
    capi.check(capi.lib().slak_copy_probe(1 << 30, 5, ctypes.byref(copy_gbps)))

    # headline: the 4096-mesh batch, graph replays
    spec = meshgen.config4(n_meshes=args.batch_meshes)
    B = batched_leg(spec, scheme, args.steps, args.warmup, plain=not args.no_plain)
    secs = B["secs"]
    value = B["faces"] * args.steps / secs
    roof = add_traffic(stage_roofline(B["prof"], CFG4_STAGE), traffic, "cfg4", args.scheme, True)
    olr = output_level_roofline(B["prof"])
    if olr is not None:
        roof["output_level_with_lak_form"] = olr
    floor_b = sum(io_floor_bytes(h, lv, lak_form) for h, lv in zip(spec.meshes, spec.levels))
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "faces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, slak_amd.meshgen)",
        "config": {
            "workload": f"cfg4: {CFG4_DESC}",
            "scheme": args.scheme.upper(),
            "meshes": len(spec.meshes),
            "levels": "per mesh, U{1..4}",
            "input_faces": int(sum(h.nfaces for h in spec.meshes)),
            "output_faces": B["faces"],
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "io_floor": io_floor(floor_b, secs / args.steps),
        "copy_kernel_gbps": round(copy_gbps.value, 1),
        "batched_vs_plain": None if B["plain_s"] is None else {
            "plain_per_mesh_s": round(B["plain_s"], 4),
            "batched_graph_s": round(secs / args.steps, 5),
            "speedup_vs_plain": round(B["plain_s"] / (secs / args.steps), 2),
            "batched_create_s": round(B["create_s"], 4),
            "batched_first_run_s": round(B["first_s"], 4),
            "speedup_vs_plain_incl_create": round(B["plain_s"] / (B["create_s"] + B["first_s"]), 2),
            "faces_per_s_plain": round(B["faces"] / B["plain_s"], 1),
        },
        "stages": stage_table(B["prof"]),
        "device_bytes": B["device_bytes"],
        "h2d_control_cages": {"seconds": round(B["h2d_s"], 4), "bytes": B["h2d_bytes"],
                              "note": "upload of the 4096 control cages (allocation included), outside the timed region"},
    }

    if not args.no_legs:
        # the same batch in the other scheme (graph replays, no plain form)
        other = "slak" if args.scheme == "lak" else "lak"
        osch = capi.SLAK_SCHEME_SLAK if other == "slak" else capi.SLAK_SCHEME_LAK
        O = batched_leg(spec, osch, args.steps, args.warmup, keep_results=0, plain=False)
        line["batched_other_scheme"] = {
            "scheme": other.upper(),
            "value": round(O["faces"] * args.steps / O["secs"], 1),
            "unit": "faces/s",
            "ms_per_step": round(O["secs"] / args.steps * 1e3, 4),
            "output_faces": O["faces"],
            "roofline": stage_roofline(O["prof"], CFG4_STAGE),
        }
        legs = {}
        for w in ("cfg2", "cfg3"):
            m = make_single(w)
            lv = SINGLE[w][1]
            leg = {"workload": f"{w}: {SINGLE[w][0]}"}
            for sname, sch in (("lak", capi.SLAK_SCHEME_LAK), ("slak", capi.SLAK_SCHEME_SLAK)):
                steps = max(3, args.steps // 2)
                s, faces, st, prof = time_single(m, lv, sch, steps, 2)
                r = stage_roofline(prof)
                if sname == args.scheme:
                    r = add_traffic(r, traffic, w, sname, False)
                leg[sname] = {
                    "faces_per_s": round(faces * steps / s, 1),
                    "ms_per_chain": round(s / steps * 1e3, 4),
                    "output_faces": faces,
                    "roofline": r,
                    "io_floor": io_floor(io_floor_bytes(m, lv, sname == "lak"), s / steps),
                    "host_syncs": int(st.host_syncs),
                }
                if w == "cfg3":
                    # GB/s of every SpGEMM / SpMV stage of the chain
                    tab = stage_table(prof)
                    for row in tab:
                        t = stage_traffic(traffic, w, row["name"], sname, False) if sname == args.scheme else None
                        if t:
                            row["hbm_read"], row["hbm_write"] = t
                    leg[sname]["stages"] = tab
            legs[w] = leg
        line["legs"] = legs

    cfg5_reqs, cfg5_kept = None, {}
    if not args.no_serve:
        keep = (0, 1, 2, 4, 5, 6, 7, 8)
        line["serve"], cfg5_reqs, cfg5_kept = serve_leg(args.serve_requests, args.serve_qps_fraction, scheme, keep=keep)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(spec, scheme, args.cpu_budget)
        line["parity"] = parity_per_config(scheme, spec, B["kept"], cfg5_reqs, cfg5_kept)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
