"""The reference's own example cages (examples/meshes/*.ccm, copied
unchanged into tests/golden/cages/) against digests the reference itself
wrote (tests/golden/make_mesh_digests.py: CatmullClark.h compiled here, one
thread, every ccs_Refine_* driver).

These are production meshes: pentagons and hexagons among quads and
triangles, valences up to 24 (Bishop), boundaries with sharpness 16 and
semi-sharp creases (Car s = 6, Rook s = 7.9, ArmorGuy s in 1..4).

- CPU: the library's ccm_Load/ccm_Save read and rewrite each file byte for
  byte (the 4 padding bytes of the header aside), and the oracle's
  refinement equals the reference's digests.
- GPU: ccm_Load, ccs_Create and each ccs_Refine_* driver through the C ABI
  equal the reference's digests bit for bit (topology, creases, points).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from halfedgecatmullclark_amd import cage as C
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
CAGES = os.path.join(HERE, "golden", "cages")
GOLD = json.load(open(os.path.join(HERE, "golden", "meshes_digests.json")))
CASES = sorted({(k.split("/")[0], int(k.split("/")[1])) for k in GOLD})
FLAGS = {
    "ccs_Refine_Scatter": O.ALL,
    "ccs_Refine_Gather": O.ALL | O.GATHER,
    "ccs_Refine_NoCreases_Scatter": O.ALL | O.NOCREASES,
    "ccs_Refine_NoCreases_Gather": O.ALL | O.GATHER | O.NOCREASES,
}


def digest(a) -> str:
    return hashlib.sha256(memoryview(np.ascontiguousarray(a).reshape(-1).view("u1"))).hexdigest()


def test_all_eight_meshes_covered():
    assert [c[0] for c in CASES] == ["ArmorGuy", "Bigguy", "Bishop", "Car", "Imrod", "Monsterfrog", "Rook", "T-Rex"]
    assert len(GOLD) == 8 * len(FLAGS)


@pytest.mark.parametrize("name,depth", CASES, ids=[c[0] for c in CASES])
def test_library_ccm_roundtrip(name, depth, tmp_path):
    from halfedgecatmullclark_amd import capi
    L = capi.lib()
    path = os.path.join(CAGES, name + ".ccm")
    m = L.ccm_Load(path.encode())
    assert m, capi.last_error()
    try:
        out = str(tmp_path / "out.ccm")
        assert L.ccm_Save(m, out.encode())
    finally:
        L.ccm_Release(m)
    a, b = open(path, "rb").read(), open(out, "rb").read()
    assert len(a) == len(b)
    assert a[:28] == b[:28] and a[32:] == b[32:]


@pytest.mark.parametrize("drv", list(FLAGS))
@pytest.mark.parametrize("name,depth", CASES, ids=[c[0] for c in CASES])
def test_oracle_matches_reference_digests(name, depth, drv):
    cg = C.load_ccm(os.path.join(CAGES, name + ".ccm"))
    want = GOLD[f"{name}/{depth}/{drv}"]
    got = O.refine(cg, depth, FLAGS[drv])
    for k in ("halfedges", "creases", "points"):
        assert digest(got[k]) == want[k], f"{name} depth {depth} {drv} {k}"


@pytest.mark.gpu
@pytest.mark.parametrize("drv", list(FLAGS))
@pytest.mark.parametrize("name,depth", CASES, ids=[c[0] for c in CASES])
def test_gpu_matches_reference_digests(name, depth, drv):
    from halfedgecatmullclark_amd import capi
    want = GOLD[f"{name}/{depth}/{drv}"]
    got = capi.refine_file(os.path.join(CAGES, name + ".ccm"), depth, drv)
    for k in ("halfedges", "creases", "points"):
        assert digest(got[k]) == want[k], f"{name} depth {depth} {drv} {k}"
