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


def time_gpu(mesh, levels, scheme, steps, warmup):
    from slak_amd import capi
    from slak_amd.mesh import MeshInfo, _subdivide

    c = MeshInfo.from_host(mesh)
    faces = 0
    try:
        def step():
            r = MeshInfo()
            st = _subdivide(scheme, c, r, levels)
            r.freeAndReset()
            return st

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
        capi.profile_enable(True)
        for _ in range(3):
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cfg2", choices=sorted(WORKLOADS))
    ap.add_argument("--scheme", default="lak", choices=["lak", "slak"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (meshes are independent; "
              "see DESIGN.md for how the work would shard). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from slak_amd import capi

    desc, levels = WORKLOADS[args.workload]
    mesh = make_mesh(args.workload)
    scheme = capi.SLAK_SCHEME_LAK if args.scheme == "lak" else capi.SLAK_SCHEME_SLAK
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
        "roofline": roofline(prof),
        "other_scheme_faces_per_s": round(faces * max(3, args.steps // 2) / secs_o, 1),
        "stages": prof["stages"],
        "host_syncs_per_step": int(st.host_syncs),
        "peak_device_bytes": int(st.peak_bytes),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
