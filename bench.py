#!/usr/bin/env python3
"""Refined faces/s of SLAK's Catmull-Clark chain on one MI355X.

Default workload: the 256 x 128 mixed-polygon torus (quads, triangles with
p = 0.25, one pentagon per grid column), 4 levels: one poly level then three
quad levels, ~9.4 M refined faces. A step is one full
LAKCatmullClark::subdivideIteratively call through the C-ABI on the
device-resident control cage (output allocation and free included). Prints
one JSON line with the throughput, the roofline of the dominant kernel
(HIP events inside the library on the launch stream) and the CPU baseline
(the same LAK algorithm, FP32, on the host's cores).
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

WORKLOADS = {
    "cfg1": ("32x16 closed quad torus, 3 levels", 3),
    "cfg2": ("256x128 mixed-polygon torus (p_tri=0.25, 1 pentagon per column), 4 levels", 4),
    "cfg3": ("1024x1024 closed quad torus, 2 levels", 2),
}


def make_mesh(name):
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


def time_gpu(mesh, levels, scheme, steps, warmup, profile_steps=3):
    from slak_amd import capi
    from slak_amd.mesh import MeshInfo, _subdivide

    import ctypes
    c = MeshInfo.from_host(mesh)
    faces = 0
    try:
        # one library call per step (output allocation and free included),
        # with the ctypes structures built once
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
        capi.check(capi.lib().slak_device_synchronize())
        t0 = time.perf_counter()
        for _ in range(steps):
            st = step()
        capi.check(capi.lib().slak_device_synchronize())
        t1 = time.perf_counter()
        faces = int(st.out_faces)
        # per-stage HIP-event profile over a few extra steps
        prof = None
        if profile_steps:
            capi.profile_enable(True)
            for _ in range(profile_steps):
                step()
            prof = capi.profile_report()
            capi.profile_enable(False)
        return (t1 - t0), faces, st, prof
    finally:
        c.freeAndReset()


def roofline(prof):
    stages = prof["stages"]
    total_ms = sum(s["total_ms"] for s in stages)
    total_b = sum(s["bytes"] for s in stages)
    dom = max(stages, key=lambda s: s["total_ms"])
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
        "chain_gbps": round(total_b / (total_ms * 1e-3) / 1e9, 1),
        "chain_frac": round(total_b / (total_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
    }


def cpu_baseline(workload, budget_s):
    """The oracle's FP32 LAK algorithm on the host (a bounded sample: the same
    mesh with one level fewer), single-threaded and OpenMP."""
    from oracle import pyoracle
    from slak_amd import meshgen

    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
    mesh = make_mesh(workload)
    levels = max(1, WORKLOADS[workload][1] - 1)
    threads = min(16, os.cpu_count() or 1)
    ms1, faces = pyoracle.time_cpu(mesh, levels, 1, warmup=0, reps=1)
    reps1 = max(1, min(5, int(budget_s * 0.5 / max(ms1 * 1e-3, 1e-3))))
    if reps1 > 1:
        ms1, faces = pyoracle.time_cpu(mesh, levels, 1, warmup=1, reps=reps1)
    msn, _ = pyoracle.time_cpu(mesh, levels, threads, warmup=1, reps=5)
    return {
        "value": round(faces / (msn * 1e-3), 1),
        "unit": "faces/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{workload} mesh refined {levels} levels ({faces} faces), FP32 LAK, median of runs after 1 warm-up",
        "value_1thread": round(faces / (ms1 * 1e-3), 1),
        "cpu_model": cpu_model(),
        "host_cores": os.cpu_count(),
    }


def batched_vs_plain(n_meshes, steps, scheme, seed=4):
    """Batched uneven workload: n_meshes seeded tori (dims uniform in [4,64]^2,
    20% of quads triangulated, depth uniform in {1..4}). Plain = one
    subdivideIteratively call per mesh; batched = one launch per kernel per
    level over the union, replayed as a hipGraph. Outputs are identical."""
    from slak_amd import capi, meshgen
    from slak_amd.batch import Batch
    from slak_amd.mesh import MeshInfo, _subdivide

    t_gen = time.perf_counter()
    spec = meshgen.config4(n_meshes=n_meshes, seed=seed)
    gen_s = time.perf_counter() - t_gen
    dev = [MeshInfo.from_host(h) for h in spec.meshes]
    try:
        # plain: per-mesh calls (one warm-up pass over a slice first)
        for d, lv in list(zip(dev, spec.levels))[:64]:
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
        # batched: create (validation + first-level analysis + allocation), graph replays
        t0 = time.perf_counter()
        b = Batch(dev, spec.levels, scheme)
        create_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        b.run(use_graph=True)  # first run captures the graph
        first_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run(use_graph=True)
        batched_s = (time.perf_counter() - t0) / steps
        inf = b.info()
        assert int(inf.out_faces) == faces
        capi.profile_enable(True)
        b.run(use_graph=False)
        prof = capi.profile_report()
        capi.profile_enable(False)
        b.close()
    finally:
        for d in dev:
            d.freeAndReset()
    rl = roofline(prof)
    return {
        "workload": f"{n_meshes} seeded tori, dims U[4,64]^2, 20% triangulated, depth U{{1..4}}",
        "out_faces": faces,
        "plain_per_mesh_s": round(plain_s, 4),
        "batched_graph_s": round(batched_s, 5),
        "batched_create_s": round(create_s, 4),
        "batched_first_run_s": round(first_s, 4),
        "speedup_vs_plain": round(plain_s / batched_s, 2),
        "speedup_vs_plain_incl_create": round(plain_s / (create_s + first_s), 2),
        "faces_per_s_batched": round(faces / batched_s, 1),
        "faces_per_s_plain": round(faces / plain_s, 1),
        "roofline": rl,
        "stages": prof["stages"],
        "generation_s": round(gen_s, 2),
    }


def serve_poisson(n_requests, qps_fraction, scheme, seed=5):
    """Served stream: Poisson arrivals of requests (torus, faces log-uniform in
    [100, 1e6], 10% triangles, level uniform in {1..3}); latency from submit
    until the refined mesh is resident. QPS is stated as a fraction of the
    saturation rate measured with every request submitted at once."""
    import numpy as np

    from slak_amd import meshgen
    from slak_amd.mesh import MeshInfo
    from slak_amd.serve import Server

    reqs = [meshgen.request_mesh(seed, i) for i in range(n_requests)]
    dev = [MeshInfo.from_host(m) for m, _ in reqs]
    try:
        def run(qps):
            rng = np.random.default_rng(seed)
            gaps = rng.exponential(1.0 / qps, size=n_requests) if qps else np.zeros(n_requests)
            lat, faces = [], 0
            with Server(scheme) as srv:
                t0 = time.perf_counter()
                ids = []
                t_next = t0
                for i, (d, (m, lv)) in enumerate(zip(dev, reqs)):
                    t_next += gaps[i]
                    now = time.perf_counter()
                    if t_next > now:
                        time.sleep(t_next - now)
                    ids.append(srv.submit(d, lv))
                for rid in ids:
                    r, st = srv.wait(rid)
                    lat.append(st.latency_us / 1e3)
                    faces += int(st.out_faces)
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

        sat = run(0.0)  # everything at once: saturation throughput
        sat_qps = n_requests / sat["wall_s"]
        qps = qps_fraction * sat_qps
        loaded = run(qps)
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
    }


# profiler stage -> rocprof kernel name prefix ({s}: scheme template argument)
# stage -> pattern of its kernel's name (re.match; {s}: 0 LAK, 1 SLAK)
STAGE_KERNELS = {
    "refine_vertex_final_quad": r"k_final<{s}, \w+, \w+, true>",
    "refine_vertex_pos_quad": r"k_final<{s}, \w+, \w+, false>",
    "refine_vertex_quad": r"k_level<true, {s},",
    "refine_vertex_poly": r"k_level<false, {s},",
    "facepoints_quad": r"k_faces<true, true, false>",
    "facepoints_poly": r"k_faces<false, true, false>",
}


def pmc_child(workload, scheme_name):
    """The profiled program of hbm_traffic(): the 1 GiB copy probe (the
    counters' calibration), then one warm-up and one timed chain."""
    import ctypes
    from slak_amd import capi
    gbps = ctypes.c_double()
    capi.check(capi.lib().slak_copy_probe(1 << 30, 1, ctypes.byref(gbps)))
    scheme = capi.SLAK_SCHEME_LAK if scheme_name == "lak" else capi.SLAK_SCHEME_SLAK
    time_gpu(make_mesh(workload), WORKLOADS[workload][1], scheme, 1, 1, profile_steps=0)


def hbm_traffic(workload, scheme_name):
    """HBM bytes per dispatch of every kernel of the chain, from two
    rocprofv3 --pmc runs of pmc_child (FETCH_SIZE, then WRITE_SIZE), each
    scaled by what the counter reads for the copy probe (1 GiB read, 1 GiB
    written). Runs before this process touches the GPU; None when rocprofv3
    is unavailable or a run fails."""
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
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", counter.lower(), "--",
               sys.executable, os.path.abspath(__file__), "--pmc-child", "--workload", workload,
               "--scheme", scheme_name]
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=180, env=dict(os.environ))
        except subprocess.TimeoutExpired:
            return None
        if r.returncode != 0:
            return None
        per = defaultdict(list)
        for f in glob.glob(os.path.join(out, "**", f"{counter.lower()}_counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") == counter:
                        name = row["Kernel_Name"]
                        for pre in ("void ", "slak::(anonymous namespace)::", "slak::"):
                            name = name.replace(pre, "")
                        per[name.split("(")[0]].append(float(row["Counter_Value"]) * 1024.0)
        counts[counter] = per
    shutil.rmtree(out, ignore_errors=True)
    fe, wr = counts["FETCH_SIZE"], counts["WRITE_SIZE"]
    if not fe.get("k_copy_probe") or not wr.get("k_copy_probe"):
        return None
    copy = float(1 << 30)
    cal_f = copy / (sum(fe["k_copy_probe"]) / len(fe["k_copy_probe"]))
    cal_w = copy / (sum(wr["k_copy_probe"]) / len(wr["k_copy_probe"]))
    kernels = {}
    for name in set(fe) | set(wr):
        f, w = fe.get(name, []), wr.get(name, [])
        kernels[name] = {"dispatches": max(len(f), len(w)),
                         "read": cal_f * sum(f) / max(len(f), 1), "write": cal_w * sum(w) / max(len(w), 1)}
    return {"calibration": {"fetch": round(cal_f, 4), "write": round(cal_w, 4)}, "kernels": kernels}


def add_traffic(roof, traffic, scheme_name):
    """roofline.traffic: average calibrated HBM bytes per dispatch of the
    dominant stage's kernels (over the same levels as bytes_per_launch)."""
    if not traffic or roof["kernel"] not in STAGE_KERNELS:
        return roof
    import re

    pattern = re.compile(STAGE_KERNELS[roof["kernel"]].format(s=0 if scheme_name == "lak" else 1))
    sel = [v for k, v in traffic["kernels"].items() if pattern.match(k)]
    n = sum(v["dispatches"] for v in sel)
    if not n:
        return roof
    rd = sum(v["read"] * v["dispatches"] for v in sel) / n
    wt = sum(v["write"] * v["dispatches"] for v in sel) / n
    roof["traffic"] = int(rd + wt)
    roof["traffic_read"] = int(rd)
    roof["traffic_write"] = int(wt)
    roof["traffic_calibration"] = traffic["calibration"]
    roof["traffic_source"] = "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, scaled by the 1 GiB copy probe"
    return roof


def io_floor_bytes(mesh, levels, lak_form):
    """Implementation-independent I/O floor of the chain: per level, read the
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
    he = set(zip(a.tolist(), b.tolist()))
    n_ext = sum(1 for (x, y) in he if (y, x) not in he)
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


def io_floor(mesh, levels, lak_form, secs_per_step):
    b = io_floor_bytes(mesh, levels, lak_form)
    gbps = b / secs_per_step / 1e9
    return {"bytes_per_chain": int(b), "gbps": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4),
            "note": "per level: read ids(+ptr+vals)+positions, write refined ids(+ptr+vals)+positions"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cfg2", choices=sorted(WORKLOADS))
    ap.add_argument("--scheme", default="lak", choices=["lak", "slak"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain leg")
    ap.add_argument("--batch-meshes", type=int, default=4096)
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-qps-fraction", type=float, default=0.5)
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM counter runs")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args.workload, args.scheme)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (meshes are independent and would be "
              "spread over GPUs by projected output faces; one huge mesh would be cut into face blocks "
              "with a one-ring halo). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # HBM counters first, in child processes, before this one touches the GPU
    traffic = None if args.no_traffic else hbm_traffic(args.workload, args.scheme)

    from slak_amd import capi

    desc, levels = WORKLOADS[args.workload]
    mesh = make_mesh(args.workload)
    scheme = capi.SLAK_SCHEME_LAK if args.scheme == "lak" else capi.SLAK_SCHEME_SLAK
    import ctypes
    copy_gbps = ctypes.c_double()
    capi.check(capi.lib().slak_copy_probe(1 << 30, 5, ctypes.byref(copy_gbps)))
    secs, faces, st, prof = time_gpu(mesh, levels, scheme, args.steps, args.warmup)
    other = capi.SLAK_SCHEME_SLAK if scheme == capi.SLAK_SCHEME_LAK else capi.SLAK_SCHEME_LAK
    secs_o, _, _, _ = time_gpu(mesh, levels, other, max(3, args.steps // 2), 1)
    value = faces * args.steps / secs
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
            "workload": f"{args.workload}: {desc}",
            "scheme": args.scheme.upper(),
            "levels": levels,
            "input_faces": mesh.nfaces,
            "input_verts": mesh.nverts,
            "output_faces": faces,
            "output_verts": int(st.out_verts),
            "mesh_digest": mesh.digest(),
            "parallelism": "single GPU",
        },
        "roofline": add_traffic(roofline(prof), traffic, args.scheme),
        "io_floor": io_floor(mesh, levels, args.scheme == "lak", secs / args.steps),
        "copy_kernel_gbps": round(copy_gbps.value, 1),
        "other_scheme_faces_per_s": round(faces * max(3, args.steps // 2) / secs_o, 1),
        "stages": prof["stages"],
        "host_syncs_per_step": int(st.host_syncs),
        "peak_device_bytes": int(st.peak_bytes),
    }
    if not args.no_batch:
        line["batched"] = batched_vs_plain(args.batch_meshes, max(3, args.steps // 4), scheme)
    if not args.no_serve:
        line["serve"] = serve_poisson(args.serve_requests, args.serve_qps_fraction, scheme)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
