"""Accuracy of the GPU points against the double-precision oracle, level by
level: the cube at depth 6, the 32x16 torus at depth 5 (its points) and an
irregular cage with a boundary and semi-sharp creases at depth 6, through
ccs_Refine_Scatter and ccs_Refine_Gather.  Per level d:

    max_abs = max |p_gpu - p_fp64|             (over every coordinate)
    max_rel = max_abs / max |p_fp64|           (relative to the level's scale)
    rms     = sqrt(mean (p_gpu - p_fp64)^2)

and every level must keep max_rel <= 1e-5.  Run as a script to print the
table:  python tests/test_gpu_accuracy.py
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from halfedgecatmullclark_amd import cage as C  # noqa: E402
from oracle import oracle as O  # noqa: E402

POINT_TOL = 1e-5

CASES = [("cube_d6", lambda: C.cube(0), 6),
         ("torus32x16_d5", lambda: C.torus(32, 16, seed=0), 5),
         ("irregular_l4_d6", lambda: C.irregular(4, seed=0), 6)]
DRIVERS = [("ccs_Refine_Scatter", O.ALL), ("ccs_Refine_Gather", O.ALL | O.GATHER)]


def level_errors(got: dict, want: dict, depth: int):
    """[(d, max_abs, max_rel, rms)] of the GPU points of levels 1..depth."""
    rows = []
    for d in range(1, depth + 1):
        ref = O.level(want, "points", d).astype(np.float64)
        lo = want["counts"]["Vcum"][d - 1]
        p = got["points"][lo:lo + len(ref)].astype(np.float64)
        diff = p - ref
        max_abs = float(np.abs(diff).max())
        rows.append((d, max_abs, max_abs / max(float(np.abs(ref).max()), 1e-30), float(np.sqrt(np.mean(diff * diff)))))
    return rows


def case_rows(capi, name, make, depth, driver, flags):
    cage = make()
    got = capi.refine_host(cage, depth, driver)
    want = O.refine(cage, depth, flags, real=np.float64)
    return level_errors(got, want, depth)


@pytest.mark.gpu
@pytest.mark.parametrize("driver,flags", DRIVERS, ids=[d[0] for d in DRIVERS])
@pytest.mark.parametrize("name,make,depth", CASES, ids=[c[0] for c in CASES])
def test_points_vs_fp64_per_level(name, make, depth, driver, flags):
    from halfedgecatmullclark_amd import capi
    capi.lib()
    for d, max_abs, max_rel, rms in case_rows(capi, name, make, depth, driver, flags):
        assert max_rel <= POINT_TOL, (name, driver, d, max_abs, max_rel, rms)
        assert rms <= max_abs


def main():
    from halfedgecatmullclark_amd import capi
    capi.lib()
    print(f"{'case':18s} {'driver':20s} {'level':>5s} {'max_abs':>10s} {'max_rel':>10s} {'rms':>10s}")
    for name, make, depth in CASES:
        for driver, flags in DRIVERS:
            for d, max_abs, max_rel, rms in case_rows(capi, name, make, depth, driver, flags):
                print(f"{name:18s} {driver:20s} {d:5d} {max_abs:10.3e} {max_rel:10.3e} {rms:10.3e}", flush=True)


if __name__ == "__main__":
    main()
