"""Compile libcchip.so (HIP kernels + C ABI) for gfx950, in-tree, and its
12-byte-record variant libcchip_nouv.so (CC_DISABLE_UV).

``python -m halfedgecatmullclark_amd.build`` rebuilds unconditionally.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libcchip.so")
ARCH = os.environ.get("CCHIP_ARCH", "gfx950")

SOURCES = ["cc_kernels.hip", "cc_fused.hip", "cc_patch.hip", "cc_batch.hip", "cc_adjoint.hip", "cc_runtime.hip",
           "cc_batch_runtime.hip", "cc_noalloc.hip", "cc_normals.hip", "cc_builder.hip",
           "cc_host.cpp"]
# -ffp-contract=off keeps the reference's rounding (no fused multiply-add);
# packed FP32 instructions are disabled on the device.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


# the two libraries built from the same sources: 16-byte level records with
# uvIDs, and 12-byte records without them (the reference's CC_DISABLE_UV)
VARIANTS = {"uv": (LIB, []), "nouv": (os.path.join(HERE, "libcchip_nouv.so"), ["-DCC_DISABLE_UV"])}


def _compile(job: tuple[str, str]) -> str:
    variant, src = job
    defines = VARIANTS[variant][1]
    obj = os.path.join(BUILD, variant, src + ".o")
    cmd = [hipcc(), f"--offload-arch={ARCH}", *FLAGS, *defines, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{err}")
    if err.strip():
        print(err, file=sys.stderr)
    return obj


def build(verbose: bool = False, variants=("uv", "nouv")) -> str:
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = [(v, s) for v in variants for s in srcs]
    for v in variants:
        os.makedirs(os.path.join(BUILD, v), exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(_compile, jobs))
    for v in variants:
        lib = VARIANTS[v][0]
        vobjs = [o for (jv, _), o in zip(jobs, objs) if jv == v]
        tmp = lib + ".tmp"
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *vobjs, "-o", tmp]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}{r.stderr}")
        os.replace(tmp, lib)
        if verbose:
            print(f"built {lib}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
