"""Cage construction (examples/obj_to_ccm.c conventions), .ccm files and the
workload generators."""
import os
import subprocess

import numpy as np
import pytest

from halfedgecatmullclark_amd import cage as C
from oracle import ref


@pytest.mark.skipif(ref.obj_to_ccm_path() is None, reason="reference obj_to_ccm not built")
@pytest.mark.parametrize("mesh", [C.cube_mesh(0), C.torus_mesh(8, 6, seed=2), C.irregular_mesh(2, seed=5),
                                  C.irregular_mesh(3, seed=8)], ids=["cube", "torus", "irregular2", "irregular3"])
def test_builder_matches_obj_to_ccm(tmp_path, mesh):
    pts, faces, uvs, fuv, cr = mesh
    mine = C.from_faces(pts, faces, uvs, fuv, cr)
    obj = tmp_path / "m.obj"
    obj.write_text(C.to_obj(pts, faces, uvs, fuv, cr))
    subprocess.run([ref.obj_to_ccm_path(), str(obj)], cwd=tmp_path, check=True, capture_output=True)
    theirs = C.load_ccm(str(tmp_path / "m.ccm"))
    for k in ("vertex_points", "uvs", "halfedges", "v2h", "e2h", "f2h"):
        assert np.array_equal(getattr(mine, k), getattr(theirs, k)), k
    assert mine.creases.tobytes() == theirs.creases.tobytes()


def test_ccm_roundtrip(tmp_path):
    cg = C.torus(8, 4, seed=3)
    p = str(tmp_path / "t.ccm")
    C.save_ccm(cg, p)
    back = C.load_ccm(p)
    for k in ("vertex_points", "uvs", "halfedges", "v2h", "e2h", "f2h"):
        assert np.array_equal(getattr(cg, k), getattr(back, k))
    with pytest.raises(ValueError):
        open(p + ".bad", "wb").write(b"x" * 40)
        C.load_ccm(p + ".bad")


@pytest.mark.parametrize("make", [lambda: C.cube(1), lambda: C.torus(64, 32), lambda: C.irregular(4, seed=0),
                                  lambda: C.irregular(2, seed=7)], ids=["cube", "torus64x32", "irregular4", "irregular2"])
def test_generators_are_manifold(make):
    cg = make()
    C.check_manifold(cg)
    assert cg.V - cg.E + cg.F in (2, 0, 1)  # sphere, torus, disk


def test_workload_shapes():
    t = C.torus(64, 32)
    assert (t.V, t.H, t.E, t.F) == (2048, 8192, 4096, 2048)
    assert np.count_nonzero(t.creases["sharpness"] == 1.5) == 64
    assert np.count_nonzero(t.creases["sharpness"] == 4.0) == 64
    ir = C.irregular(4, seed=0)
    sizes = np.bincount(np.diff(np.append(ir.f2h, ir.H)))
    assert sizes[3] > 0 and sizes[4] > 0          # mixed triangles and quads
    assert np.any(ir.halfedges[:, 0] < 0)          # a boundary loop
    assert np.all(ir.creases["sharpness"][ir.halfedges[ir.e2h, 0] < 0] == 16.0)


def test_batch_generator():
    cages, depths = C.batch_cages(12, seed=3)
    assert len(cages) == 12 and set(depths) <= {4, 5, 6}
    for cg in cages:
        C.check_manifold(cg)
        assert 20 <= cg.H <= 70000
