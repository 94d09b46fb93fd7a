"""Benchmark of the GPU refinement path (one MI355X).

Workload: the 64x32 torus cage (2,048 quads, 8,192 halfedges, R=1, r=0.3,
N(0, 0.01) jitter, one longitudinal crease ring of sharpness 1.5 and one of
4.0, grid uvs) refined to depth 8 with every step of ccs_Refine_Scatter
(halfedges, creases, uvs, vertex points).  One "step" is one full
refinement of all 8 levels from the cage resident in HBM; value is refined
halfedges (levels 1..8, 715,816,960) per second.

Before timing, ccd_Autotune picks each level's kernel configuration (the
fused kernel's workgroup size and tiles per workgroup, the vertex kernel's
form and waves per SIMD, or the patch-local kernel) by timing them; the
timed steps replay ccd_Refine's captured graph.

Also reported on the same line:
  - every kernel of one step (per-kernel HIP events, a second pass);
  - the gather point rules on the same workload;
  - the points-only forward and the adjoint of the points at depth 8;
  - the same refinement with 12-byte level records (CC_DISABLE_UV library);
  - halfedge normals of level 8;
  - the cage builder on a 131,072-face OBJ (ccm_LoadObj, ccm_CreateFromFaces),
    with the reference's converter timed on the same file in the CPU leg;
  - batched refinement: 256 cages of uneven size (halfedges log-uniform in
    [24, 60000], depths 4-6, 5% of edges creased U[0,4]) refined in one
    launch per kernel kind and level from a captured hipGraph, forward and
    adjoint, against the same 256 cages refined one after another through
    the single-cage device API (ccd_Refine / ccd_RefinePointsAdjoint);
  - forward + adjoint against forward alone on the 32x16 torus at depth 5
    (the cage-fitting workload);
  - roofline of the dominant kernel (per-kernel HIP events on the launch
    stream over the timed steps, algorithmic bytes per launch), also as a
    fraction of what torch's device copy reaches on the same GPU;
  - streaming ceilings: torch's device copy and fill of 2 GiB buffers;
  - the reference's own OpenMP CPU path (CatmullClark.h built by
    oracle/Makefile) timed on this host's cores on the same workload
    (torus 64x32 to depth 8), per stage and in total.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "refined halfedges/s + HBM GB/s at depth 8; fwd+adjoint and batched vs plain"


def level_bytes(n, d, D=None, owner_bits=False):
    """Algorithmic bytes of the fused level kernel d -> d+1: each input read
    once and each output written once.  Reads: the level-d halfedge records
    (16 B), level-d points (12 B), level-d creases (12 B); writes: the four
    16-B child records per parent halfedge and the level-(d+1) face and edge
    points (12 B).  With owner_bits (the masked vertex kernel follows) it
    also writes one owner bit per halfedge, reads the parent level's
    edge-representative bits and, from level 2 on, writes the vertex points
    of the face vertices born at level d (12 B each, one per parent-level
    face); on levels d + 1 <= D - 2 it reads level d's bits and writes level
    d+1's (one bit per halfedge each)."""
    b = 16 * n["H"][d] + 12 * n["V"][d] + 12 * n["C"][d] + 64 * n["H"][d] + 12 * (n["F"][d] + n["E"][d])
    if owner_bits and d >= 2:
        b += 12 * n["F"][d - 1]
    if D is not None and D >= 4:
        if owner_bits:
            b += n["H"][d] // 8 + (n["H"][d - 1] // 8 if d >= 2 else 0)
        if d + 1 <= D - 2:
            b += n["H"][d] // 8 + n["H"][d + 1] // 8
    return b


def vertex_kernel_bytes(n, d):
    """Algorithmic bytes of the masked vertex-point kernel d -> d+1: the owner
    bits (1 bit per level-d halfedge), the level-d points of its vertices and
    the creases (12 B), the level-(d+1) face and edge points it reads once,
    one parent record (16 B) per edge vertex born at level d (the other rings
    follow from the child numbering), and the level-(d+1) vertex points it
    writes.  From level 2 on the fused kernel finishes the face vertices born
    at level d (one per parent-level face), so they are left out."""
    born_edges = n["E"][d - 1] if d >= 2 else 0
    verts = n["V"][d] - (n["F"][d - 1] if d >= 2 else 0)
    return (n["H"][d] // 8 + 12 * verts + 12 * n["C"][d] + 12 * (n["F"][d] + n["E"][d]) + 16 * born_edges
            + 12 * verts)


def patch_bytes(n, d):
    """Algorithmic bytes of the patch-local kernel d -> d+1: the fused
    kernel's plus the level-(d+1) vertex points it writes."""
    return level_bytes(n, d) + 12 * n["V"][d]


def refinement_bytes(n, D, rec_bytes=16):
    """Algorithmic bytes of the whole refinement (all steps, levels 0..D-1):
    per step d -> d+1 read the level-d records (28 B per cage halfedge at
    d = 0), points and creases once, write the child records (rec_bytes
    each: 16 with uvIDs, 12 without), all level-(d+1) points and creases
    once."""
    total = 0
    for d in range(D):
        rec = 28 if d == 0 else rec_bytes
        total += (rec * n["H"][d] + 4 * rec_bytes * n["H"][d] + 12 * n["V"][d] + 12 * n["V"][d + 1] + 12 * n["C"][d]
                  + 24 * n["C"][d])
    return total


def adjoint_bytes(n, D):
    """Compulsory bytes of the adjoint of the points (all steps d+1 -> d,
    d = D-1 .. 0): per step the level-d topology (16 B per halfedge; 28 B per
    cage halfedge at d = 0) and creases (12 B) read once, the level-(d+1)
    gradient (12 B per point) read once and the level-d gradient written
    once."""
    total = 0
    for d in range(D):
        rec = 28 if d == 0 else 16
        total += rec * n["H"][d] + 12 * n["V"][d + 1] + 12 * n["V"][d] + 12 * n["C"][d]
    return total


def cpu_baseline(depth: int = 8, runs: int = 3):
    """The reference's OpenMP ccs_Refine_Scatter on this host's cores, on the
    same workload (torus 64x32 to depth 8), timed per stage and in total by
    oracle/cpu_baseline.py in a child process, so that OMP_PROC_BIND=close
    and OMP_PLACES=cores are in place before OpenMP starts.  Threads: every
    CPU this process may run on (its affinity mask), within the host's CPU
    share for this GPU when the environment states one in OMP_NUM_THREADS."""
    import subprocess
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(affinity, share) if share > 0 else affinity
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--depth", str(depth), "--runs", str(runs)]
    try:
        r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
        res = json.loads(r.stdout.strip().splitlines()[-1])
    except Exception as e:  # pragma: no cover
        return {"value": None, "unit": "refined halfedges/s", "cores": affinity, "kind": "reference",
                "sample": f"unavailable: {e}"}
    if "total" not in res:
        return {"value": None, "unit": "refined halfedges/s", "cores": affinity, "kind": "reference",
                "sample": res.get("error", "no result")}
    st = res["stages"]
    return {"value": res["refined_halfedges_per_s"], "unit": "refined halfedges/s", "cores": res["threads"],
            "kind": "reference", "cpu_model": res["cpu_model"], "cpus_online": res["cpus_online"],
            "total_ms": res["total"], "stages_ms": {k: v["median_ms"] for k, v in st.items()},
            "sample": f"CatmullClark.h ccs_Refine_Scatter (OpenMP, {res['threads']} threads, OMP_PROC_BIND=close, "
                      f"OMP_PLACES=cores, gcc -O3 default x86-64 target) on {res['cpu_model']} "
                      f"({res['cpus_online']} CPUs online, {res['cpus_usable']} usable): torus 64x32 to depth "
                      f"{depth} ({res['refined_halfedges']} refined halfedges), median of {res['total']['runs']} "
                      f"runs after one untimed run: {res['total']['median_ms']:.0f} ms; stages once each "
                      f"(creases, halfedges, points, uvs as examples/subd_cpu.c): "
                      + ", ".join(f"{v['median_ms']:.0f}" for v in st.values()) + " ms"}


def builder_cpu_baseline(obj_path: str, runs: int = 3):
    """The reference's converter (examples/obj_to_ccm.c built by
    oracle/Makefile, OpenMP on every core) on the builder workload's OBJ
    file: parse, build, write the .ccm; wall-clock median of `runs` after one
    untimed run."""
    import subprocess
    from oracle import ref
    exe = ref.obj_to_ccm_path()
    if exe is None:
        return {"ms": None, "sample": "the reference converter is not built"}
    cwd = os.path.dirname(obj_path)
    ts = []
    for i in range(runs + 1):
        t0 = time.perf_counter()
        r = subprocess.run([exe, obj_path], cwd=cwd, capture_output=True)
        if r.returncode != 0:
            return {"ms": None, "sample": f"obj_to_ccm failed ({r.returncode})"}
        if i:
            ts.append((time.perf_counter() - t0) * 1e3)
    return {"ms": sorted(ts)[runs // 2], "sample": f"examples/obj_to_ccm.c (gcc -O2 -fopenmp) on the same OBJ, "
                                                   f"median of {runs} runs"}


def batched(torch, capi, C, steps: int, stream):
    """Batched forward + adjoint of 256 uneven cages against the same cages
    refined one by one through the single-cage device API."""
    sptr = stream.cuda_stream
    t0 = time.perf_counter()
    cages, depths = C.batch_cages(256, seed=0)
    gen_s = time.perf_counter() - t0
    b = capi.Batch(cages, depths)
    total = b.count(capi.COUNT_HALFEDGES)
    nf = b.count(capi.COUNT_FINEST_POINTS)
    b.refine(capi.ALL, sptr)        # topology for the adjoint, and graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.autotune(capi.ALL, sptr)
    tune_s = time.perf_counter() - t0
    g = torch.randn((nf, 3), device="cuda", dtype=torch.float32)
    gc = torch.empty((b.count(capi.COUNT_CAGE_POINTS), 3), device="cuda", dtype=torch.float32)

    def run(f, k):
        f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(k):
            f()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    k = max(3, steps // 4)
    b_fwd = run(lambda: b.refine(capi.ALL, sptr), k)
    b_adj = run(lambda: b.refine_points_adjoint(g.data_ptr(), gc.data_ptr(), 0, sptr), k)
    offs = [b.offset(capi.COUNT_FINEST_POINTS, i) for i in range(len(cages))]
    b.close()
    torch.cuda.empty_cache()
    # the plain side: each cage refined by its own launch sequence
    # (ccd_Refine kernel by kernel); the same cages replaying a captured
    # graph each are reported beside it
    subs = [capi.DeviceSubd(c, d) for c, d in zip(cages, depths)]
    for s_ in subs:
        s_.set_option(capi.OPT_GRAPH, 0)
        s_.refine(capi.ALL, sptr)
    gslices = [g[o:] for o in offs]
    gcs = [torch.empty((c.V, 3), device="cuda", dtype=torch.float32) for c in cages]

    def plain_fwd():
        for s_ in subs:
            s_.refine(capi.ALL, sptr)

    def plain_adj():
        for s_, gi, go in zip(subs, gslices, gcs):
            s_.refine_points_adjoint(gi.data_ptr(), go.data_ptr(), 0, sptr)

    p_fwd = run(plain_fwd, k)
    p_adj = run(plain_adj, k)
    for s_ in subs:
        s_.set_option(capi.OPT_GRAPH, 1)
    p_fwd_graphs = run(plain_fwd, k)
    for s_ in subs:
        s_.close()
    return {"workload": "256 cages (cube/torus/irregular generators, halfedges log-uniform in [24, 60000], "
                        "depths uniform in {4,5,6}, 5% of edges creased U[0,4], seed 0)",
            "refined_halfedges": total, "cage_generation_s": gen_s, "autotune_s": tune_s,
            "batched_fwd_ms": b_fwd, "plain_fwd_ms": p_fwd, "fwd_speedup": p_fwd / b_fwd,
            "plain_fwd_graph_per_cage_ms": p_fwd_graphs,
            "batched_fwd_value": total / (b_fwd * 1e-3),
            "batched_adj_ms": b_adj, "plain_adj_ms": p_adj, "adj_speedup": p_adj / b_adj,
            "batched_fwd_adj_ms": b_fwd + b_adj, "plain_fwd_adj_ms": p_fwd + p_adj,
            "fwd_adj_speedup": (p_fwd + p_adj) / (b_fwd + b_adj)}


def host_boundary(capi, C, D, runs: int = 3):
    """ccs_Refine_Scatter through the host cc_Subd (CatmullClark.h:200): the
    cage read from host memory and all levels written back to the subd's
    host arrays before the call returns, so the PCIe copies are inside the
    timed region.  Best of `runs` calls after one untimed call."""
    s = capi.HostSubd(C.torus(64, 32, seed=0), D)
    try:
        s.run("ccs_Refine_Scatter")
        best = float("inf")
        for _ in range(runs):
            t0 = time.perf_counter()
            s.run("ccs_Refine_Scatter")
            best = min(best, time.perf_counter() - t0)
        n = s.n
        nbytes = 16 * n["Hcum"][D] + 12 * n["Ccum"][D] + 12 * n["Vcum"][D]
        return {"workload": f"torus 64x32 cage, depth {D}, host arrays", "ms": best * 1e3,
                "value": n["Hcum"][D] / best, "unit": "refined halfedges/s",
                "host_bytes_written": nbytes, "host_gbs": nbytes / best / 1e9}
    finally:
        s.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched comparison")
    ap.add_argument("--host-boundary", action="store_true",
                    help="also time ccs_Refine_Scatter on host arrays (the drop-in boundary: cage upload, "
                         "refinement, every level copied back over PCIe)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed region, copy 1 GiB between two device buffers twice (a kernel of "
                         "known traffic for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the refinement of one cage runs on one device; "
              "run it with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from halfedgecatmullclark_amd import capi
    from halfedgecatmullclark_amd import cage as C

    torch.cuda.init()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sptr = stream.cuda_stream

    D = 8
    cage = C.torus(64, 32, seed=0)
    ds = capi.DeviceSubd(cage, D)
    n = ds.n
    hcum = n["Hcum"][D]

    def timed(steps_flags, k, w, per_kernel=False):
        for _ in range(w):
            ds.refine(steps_flags, sptr)
        ds.set_timing(per_kernel)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        kern = {}
        e0.record(stream)
        for _ in range(k):
            ds.refine(steps_flags, sptr)
            if per_kernel:
                for i, (name, ms) in enumerate(ds.timings()):
                    kern.setdefault((i, name), []).append(ms)
        e1.record(stream)
        torch.cuda.synchronize()
        ds.set_timing(False)
        return e0.elapsed_time(e1) / k, kern

    ds.refine(capi.ALL, sptr)
    ds.autotune(capi.ALL, sptr)
    level_cfg = ds.level_options()
    # the step time with nothing between the launches; then the same steps
    # again with a HIP event pair around every launch (read back after each
    # step) for the per-kernel durations
    ms_scatter, _ = timed(capi.ALL, args.steps, args.warmup)
    ms_events, kern = timed(capi.ALL, args.steps, 1, per_kernel=True)
    ms_gather, _ = timed(capi.ALL | capi.GATHER, max(3, args.steps // 2), 1)

    # the adjoint on the same workload: points-only forward (the topology
    # stays), the adjoint of the points, and the two in a row
    ds.refine(capi.ALL, sptr)
    cage_pts = torch.from_numpy(cage.vertex_points.astype("float32")).to(dev)
    pts_out = torch.empty((n["Vcum"][D], 3), device=dev)
    grad_fin = torch.randn((n["V"][D], 3), device=dev)
    grad_cage = torch.empty((n["V"][0], 3), device=dev)

    def ev_time(f, k):
        for _ in range(2):
            f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(k):
            f()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    k1 = max(5, args.steps // 2)
    fwd_pts_ms = ev_time(lambda: ds.refine_points(cage_pts.data_ptr(), pts_out.data_ptr(), 0, sptr), k1)
    adj_ms = ev_time(lambda: ds.refine_points_adjoint(grad_fin.data_ptr(), grad_cage.data_ptr(), 0, sptr), k1)
    adj_bytes = adjoint_bytes(n, D)
    adjoint_d8 = {"workload": "torus 64x32 cage, depth 8: points-only forward (ccd_RefinePoints) and the adjoint "
                              "of the points (ccd_RefinePointsAdjoint), topology fixed",
                  "fwd_points_ms": fwd_pts_ms, "adj_ms": adj_ms, "fwd_adj_ms": fwd_pts_ms + adj_ms,
                  "ratio": (fwd_pts_ms + adj_ms) / fwd_pts_ms, "ratio_vs_full_refinement": (fwd_pts_ms + adj_ms) / ms_scatter,
                  "adj_algorithmic_bytes": adj_bytes, "adj_gbs": adj_bytes / (adj_ms * 1e-3) / 1e9,
                  "adj_roofline_frac": adj_bytes / (adj_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    del pts_out, grad_fin, grad_cage

    # halfedge normals of the finest level (537 M halfedges): corner
    # normals, and vertex normals summed around each ring
    nrm = torch.empty((n["H"][D], 3), device=dev)
    nrm_ms = {s_: ev_time(lambda: ds.halfedge_normals(nrm.data_ptr(), s_, sptr), 5) for s_ in (False, True)}
    del nrm
    HD, VD = n["H"][D], n["V"][D]
    nrm_bytes = {False: 16 * HD + 12 * VD + 12 * HD, True: 16 * HD + 12 * VD + 12 * HD + 16 * HD + 24 * HD}
    normals = {"workload": "halfedge normals of the torus 64x32 at depth 8 (ccd_HalfedgeNormals)",
               "corner_ms": nrm_ms[False], "vertex_ms": nrm_ms[True],
               "corner_gbs": nrm_bytes[False] / (nrm_ms[False] * 1e-3) / 1e9,
               "vertex_gbs": nrm_bytes[True] / (nrm_ms[True] * 1e-3) / 1e9}

    # the cage builder on a 512 x 256 quad torus with uvs and two crease
    # rings (131,072 faces, 524,288 halfedges): its OBJ text through
    # ccm_LoadObj (parsed on the host as the reference's converter parses,
    # built on the GPU, copied back into a cc_Mesh), and ccm_CreateFromFaces
    # on the flat face arrays (upload, GPU build, the cc_Mesh copied back);
    # wall-clock medians of 3 after one warm-up
    import tempfile
    mesh = C.torus_mesh(512, 256, seed=0)
    obj_dir = tempfile.mkdtemp()
    obj_path = os.path.join(obj_dir, "torus512x256.obj")
    with open(obj_path, "w") as f:
        f.write(C.to_obj(*mesh))

    def wall_ms(f, k=3):
        f()
        ts = []
        for _ in range(k):
            t0 = time.perf_counter()
            f()
            ts.append((time.perf_counter() - t0) * 1e3)
        return sorted(ts)[k // 2]

    built = capi.build_cage(*mesh)
    arrs = capi.face_arrays(*mesh)
    builder = {"workload": "OBJ of a 512x256 quad torus with uvs and two crease rings (131072 faces, "
                           f"{built.H} halfedges) to a cc_Mesh",
               "load_obj_ms": wall_ms(lambda: capi.release_mesh(capi.load_obj_mesh(obj_path))),
               "from_faces_ms": wall_ms(lambda: capi.release_mesh(capi.create_mesh(arrs))), "halfedges": built.H}
    del built, arrs

    # the same refinement with 12-byte level records (libcchip_nouv.so, the
    # CC_DISABLE_UV layout): no uvIDs, 48 instead of 64 bytes of children
    # per parent halfedge
    ds12 = capi.DeviceSubd(cage, D, variant="nouv")
    ds12.refine(capi.ALL, sptr)
    ds12.autotune(capi.ALL, sptr)
    ms12 = ev_time(lambda: ds12.refine(capi.ALL, sptr), max(5, args.steps // 2))
    ds12.close()
    nouv = {"workload": "torus 64x32 cage to depth 8, ccs_Refine_Scatter steps without uvs, 12-byte level records "
                        "(CC_DISABLE_UV library)",
            "ms_per_step": ms12, "value": hcum / (ms12 * 1e-3),
            "hbm_gbs_algorithmic": refinement_bytes(n, D, 12) / (ms12 * 1e-3) / 1e9}

    # dominant kernel: largest mean duration among the launches of one step;
    # the per-level kernels run in level order, each level starting with its
    # fused or patch-local kernel
    labelled, d = [], 0
    for (i, nm) in sorted(kern):
        if nm in ("level_fused", "level_points", "level_patch"):
            d += 1
        labelled.append(((i, nm), d))
    ((idx, name), level) = max(labelled, key=lambda kv: sum(kern[kv[0]]) / len(kern[kv[0]]))
    # every launch of one step: (name, level, mean ms), the level a kernel
    # reads (0: cage step and creases)
    step_kernels = [(nm, lv, round(sum(kern[(i, nm)]) / len(kern[(i, nm)]), 4)) for (i, nm), lv in labelled]
    durs = kern[(idx, name)]
    mean_ms = sum(durs) / len(durs)
    if name == "level_fused":
        nbytes = level_bytes(n, level, D, owner_bits=level_cfg[level - 1]["scan"] == 0)
    elif name == "level_vertices":
        nbytes = vertex_kernel_bytes(n, level)
    elif name == "level_patch":
        nbytes = patch_bytes(n, level)
    else:
        nbytes = None
    achieved = nbytes / (mean_ms * 1e-3) / 1e9 if nbytes else None
    # HBM bytes per launch of this kernel at this level, from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by a 1 GiB copy)
    traffic = None
    prof = os.path.join(ROOT, "profiles", "traffic_torus64x32_d8.json")
    if os.path.exists(prof):
        try:
            traffic = json.load(open(prof)).get("kernels", {}).get(f"{name}@{level}", {}).get("traffic_bytes")
        except Exception:
            traffic = None

    # forward + adjoint on the cage-fitting workload (32x16 torus, depth 5)
    from halfedgecatmullclark_amd.autograd import Subdivider
    sub = Subdivider(C.torus(32, 16, seed=0), 5)
    V = sub.cage_points().clone()
    out = torch.empty((sub.Vcum, 3), device=dev)
    gcage = torch.empty((sub.V0, 3), device=dev)
    target = torch.randn((sub.VD, 3), device=dev)
    K3 = max(20, args.steps)

    def fwd():
        sub.ds.refine_points(V.data_ptr(), out.data_ptr(), 0, sptr)

    def fwd_adj():
        fwd()
        r = out[sub.finest_offset:] - target
        sub.ds.refine_points_adjoint(r.data_ptr(), gcage.data_ptr(), 0, sptr)

    for f in (fwd, fwd_adj):
        for _ in range(3):
            f()
    res = {}
    for label, f in (("fwd", fwd), ("fwd_adj", fwd_adj)):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(K3):
            f()
        e1.record(stream)
        torch.cuda.synchronize()
        res[label] = e0.elapsed_time(e1) / K3
    sub.close()

    ds.close()
    # what plain streaming kernels reach on this GPU, for comparison with the
    # level kernels: torch's device copy (2 GiB read + 2 GiB written) and
    # fill (2 GiB written)
    ca = torch.ones(512 * 1024 * 1024, dtype=torch.float32, device=dev)
    cb = torch.empty_like(ca)
    copy_ms = ev_time(lambda: cb.copy_(ca), 10)
    fill_ms = ev_time(lambda: cb.fill_(1.0), 10)
    nb = ca.numel() * 4
    ceilings = {"copy_gbs": 2 * nb / (copy_ms * 1e-3) / 1e9, "fill_gbs": nb / (fill_ms * 1e-3) / 1e9,
                "bytes_per_buffer": nb}
    del ca, cb
    if args.copy_calibration:
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(2):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    host = host_boundary(capi, C, D) if args.host_boundary else None
    batch = None if args.no_batch else batched(torch, capi, C, args.steps, stream)
    cpu = None if args.no_cpu else cpu_baseline()
    if cpu is not None:
        cpu["builder"] = builder_cpu_baseline(obj_path)
    import shutil
    shutil.rmtree(obj_dir, ignore_errors=True)

    line = {
        "metric": METRIC,
        "value": hcum / (ms_scatter * 1e-3),
        "unit": "refined halfedges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_scatter,
        "ms_per_step_with_kernel_events": ms_events,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded torus cage)",
        "config": {"workload": "torus 64x32 cage (8192 halfedges, crease rings s=1.5 and s=4.0, uvs) "
                               "refined to depth 8 with ccs_Refine_Scatter semantics",
                   "depth": D, "refined_halfedges": hcum, "finest_halfedges": n["H"][D]},
        "hbm_gbs_algorithmic": refinement_bytes(n, D) / (ms_scatter * 1e-3) / 1e9,
        "gather": {"ms_per_step": ms_gather, "value": hcum / (ms_gather * 1e-3)},
        "adjoint_d8": adjoint_d8,
        "nouv_records": nouv,
        "normals": normals,
        "builder": builder,
        "fwd_adjoint": {"workload": "torus 32x16 cage, depth 5, loss 0.5*||S(V)-T||^2",
                        "fwd_ms": res["fwd"], "fwd_adj_ms": res["fwd_adj"],
                        "ratio": res["fwd_adj"] / res["fwd"]},
        "batched": batch,
        "host_boundary": host,
        "level_kernels": level_cfg,
        "step_kernels_ms": step_kernels,
        "roofline": {"bound": "hbm", "kernel": f"{name} (level {level} -> {level + 1})",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": traffic,
                     "kernel_ms": mean_ms, "algorithmic_bytes": nbytes,
                     "frac_of_copy": (achieved / ceilings["copy_gbs"]) if achieved else None},
        "streaming_ceilings": ceilings,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
