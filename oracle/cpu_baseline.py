"""TEST INFRASTRUCTURE ONLY: the CPU baseline of the benchmark.

Times the reference's own OpenMP ccs_Refine_Scatter (CatmullClark.h built by
oracle/Makefile into oracle/_ref/, so it exists only where the reference tree
was present at build time) on the torus 64x32 cage, the way
examples/subd_cpu.c:132-180 times it: clock_gettime-style wall time per call,
median/mean/min/max over the runs, per stage (ccs_RefineCreases,
ccs_RefineHalfedges, ccs_RefineVertexPoints_Scatter, ccs_RefineVertexUvs,
the order of subd_cpu.c:220-253) and for the whole ccs_Refine_Scatter.

Run as its own process so that the OpenMP placement variables take effect
before libgomp starts:

    OMP_NUM_THREADS=16 OMP_PROC_BIND=close OMP_PLACES=cores \\
        python -m oracle.cpu_baseline --depth 8 --runs 5

Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

STAGES = ("ccs_RefineCreases", "ccs_RefineHalfedges", "ccs_RefineVertexPoints_Scatter", "ccs_RefineVertexUvs")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def stats(ts: list[float]) -> dict:
    s = sorted(ts)
    return {"median_ms": s[len(s) // 2] * 1e3, "mean_ms": sum(s) / len(s) * 1e3, "min_ms": s[0] * 1e3,
            "max_ms": s[-1] * 1e3, "runs": len(s)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--stage-runs", type=int, default=1)
    ap.add_argument("--threads", type=int, default=0, help="0: OMP_NUM_THREADS, else every CPU this process may use")
    args = ap.parse_args()

    from halfedgecatmullclark_amd import cage as C
    from oracle import ref

    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = args.threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or affinity
    out = {"kind": "reference", "cpu_model": cpu_model(), "cpus_online": os.cpu_count(), "cpus_usable": affinity,
           "threads": threads, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
           "omp_places": os.environ.get("OMP_PLACES"), "depth": args.depth,
           "workload": f"torus 64x32 cage (8192 halfedges), ccs_Refine_Scatter to depth {args.depth}"}
    if not ref.available():
        out["error"] = "reference build (oracle/_ref) not present"
        print(json.dumps(out))
        return
    t0 = time.perf_counter()
    r = ref.RefSubd(C.torus(64, 32, seed=0), args.depth)
    r.run("ccs_Refine_Scatter", threads=threads)   # first touch of the level arrays
    out["setup_s"] = time.perf_counter() - t0
    per_stage = {}
    for name in STAGES:
        ts = []
        for _ in range(args.stage_runs):
            t = time.perf_counter()
            r.run(name, threads=threads)
            ts.append(time.perf_counter() - t)
        per_stage[name] = stats(ts)
    total = []
    for _ in range(args.runs):
        t = time.perf_counter()
        r.run("ccs_Refine_Scatter", threads=threads)
        total.append(time.perf_counter() - t)
    out["refined_halfedges"] = int(r.n["Hcum"][args.depth])
    r.close()
    out["stages"] = per_stage
    out["total"] = stats(total)
    out["refined_halfedges_per_s"] = out["refined_halfedges"] / (out["total"]["median_ms"] * 1e-3)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
