"""Writes tests/golden/nouv_digests.json from the reference's own
CatmullClark.h built with CC_DISABLE_UV (12-byte cc_Halfedge_SemiRegular
records, no uv refinement; oracle/Makefile builds it into
oracle/_ref/libccref32_nouv.so).

Run in a container that has the reference tree:
    python tests/golden/make_nouv_digests.py

The cases are the golden cages of make_golden.py and three of the
reference's example meshes; for each ccs_Refine_* driver the fixture keeps
the SHA-256 digests of the reference's halfedge, crease and point arrays
(one thread, so the scatter drivers are deterministic).
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from halfedgecatmullclark_amd import cage as C  # noqa: E402
from oracle import ref  # noqa: E402

CASES = [("cube", 5), ("torus16x8", 4), ("irregular_l2", 3), ("irregular_soft", 3), ("Rook", 3), ("Bishop", 3),
         ("Car", 3), ("Bigguy", 3), ("Monsterfrog", 3)]
DRIVERS = ["ccs_Refine_Scatter", "ccs_Refine_Gather", "ccs_Refine_NoCreases_Scatter", "ccs_Refine_NoCreases_Gather"]
OUT = os.path.join(HERE, "nouv_digests.json")


def digest(a) -> str:
    return hashlib.sha256(memoryview(a.reshape(-1).view("u1"))).hexdigest()


def main():
    if not ref.available(variant="nouv"):
        raise SystemExit("the reference's CC_DISABLE_UV build is missing (make -C oracle)")
    out = {}
    for name, depth in CASES:
        cg = C.load_ccm(os.path.join(HERE, "cages", name + ".ccm"))
        for drv in DRIVERS:
            r = ref.refine(cg, depth, drv, threads=1, variant="nouv")
            assert r["halfedges"].dtype.itemsize == 12
            out[f"{name}/{depth}/{drv}"] = {k: digest(r[k]) for k in ("halfedges", "creases", "points")}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(f"wrote {OUT} ({len(out)} cases)")


if __name__ == "__main__":
    main()
