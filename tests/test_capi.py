"""The C-ABI library without a GPU: it loads, exports every symbol the
public header declares, and its host-side functions (counts, accessors,
maps, cage I/O, refusals) agree with the oracle.  No kernel is launched."""
import ctypes
import os
import re

import numpy as np
import pytest

from halfedgecatmullclark_amd import capi
from halfedgecatmullclark_amd import cage as C
from halfedgecatmullclark_amd.abi import MeshView, cc_Mesh, cc_Subd
from oracle import oracle as O
from tests.helpers import small_cages

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b((?:ccs|ccm|ccd|ccb)_[A-Za-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    assert len(declared_symbols()) > 100


def test_struct_layouts():
    assert ctypes.sizeof(cc_Mesh) == 5 * 4 + 4 + 7 * 8
    assert ctypes.sizeof(cc_Subd) == 4 * 8 + 8


@pytest.mark.parametrize("name,cage,depth", small_cages()[:5], ids=[c[0] for c in small_cages()[:5]])
def test_counts_and_maps(name, cage, depth):
    L = capi.lib()
    m = MeshView(cage)
    n = O.refine(cage, 1, O.HALFEDGES)["counts"]
    for d in range(0, depth + 1):
        assert L.ccs_CumulativeHalfedgeCountAtDepth(m.pointer, d) == O.count(cage, "Hcum", d)
        assert L.ccs_CumulativeVertexCountAtDepth(m.pointer, d) == O.count(cage, "Vcum", d)
        assert L.ccs_CumulativeCreaseCountAtDepth(m.pointer, d) == O.count(cage, "Ccum", d)
        assert L.ccm_VertexCountAtDepth(m.pointer, d) == O.count(cage, "V", d)
        assert L.ccm_EdgeCountAtDepth(m.pointer, d) == O.count(cage, "E", d)
        assert L.ccm_FaceCountAtDepth(m.pointer, d) == O.count(cage, "F", d)
    s = L.ccs_Create(m.pointer, depth)
    assert s
    for d in range(1, min(depth, 3) + 1):
        for e in range(0, O.count(cage, "E", d), 7):
            assert L.ccs_EdgeToHalfedgeID(s, e, d) == O.edge_to_halfedge(cage, e, d)
        for v in range(0, O.count(cage, "V", d), 5):
            assert L.ccs_VertexPointToHalfedgeID(s, v, d) == O.vertex_to_halfedge(cage, v, d)
            assert L.ccs_VertexToHalfedgeID(s, v, d) == O.vertex_to_halfedge(cage, v, d)
    L.ccs_Release(s)


def test_create_refusals():
    L = capi.lib()
    torus = MeshView(C.torus(64, 32))
    s = L.ccs_Create(torus.pointer, 8)  # the largest depth the int32 formulas allow
    assert s
    L.ccs_Release(s)
    assert not L.ccs_Create(torus.pointer, 9)
    assert "overflows int32" in capi.last_error()[1]
    assert not L.ccs_Create(torus.pointer, 0)
    assert capi.last_error()[0] != 0
    assert not L.ccs_Create(None, 2)


def test_ccm_save_load_roundtrip(tmp_path):
    L = capi.lib()
    cage = C.irregular(2, seed=3)
    py_path = str(tmp_path / "a.ccm")
    c_path = str(tmp_path / "b.ccm")
    C.save_ccm(cage, py_path)
    m = L.ccm_Load(py_path.encode())
    assert m
    assert m.contents.halfedgeCount == cage.H and m.contents.edgeCount == cage.E
    assert L.ccm_Save(m, c_path.encode())
    L.ccm_Release(m)
    assert open(py_path, "rb").read() == open(c_path, "rb").read()
    back = C.load_ccm(c_path)
    assert np.array_equal(back.halfedges, cage.halfedges)
    assert back.creases.tobytes() == cage.creases.tobytes()


def test_ccm_load_errors(tmp_path):
    L = capi.lib()
    assert not L.ccm_Load(str(tmp_path / "missing.ccm").encode())
    assert capi.last_error()[0] != 0
    bad = tmp_path / "bad.ccm"
    bad.write_bytes(b"not a mesh at all, not at all....")
    assert not L.ccm_Load(str(bad).encode())
    good = tmp_path / "good.ccm"
    C.save_ccm(C.cube(0), str(good))
    data = good.read_bytes()
    (tmp_path / "short.ccm").write_bytes(data[:-10])
    assert not L.ccm_Load(str(tmp_path / "short.ccm").encode())
    assert "data reading failed" in capi.last_error()[1]


def test_refine_without_gpu_fails_loudly():
    """No CPU fallback: without a device the drop-in records an error."""
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(capi.CCError):
        capi.refine_host(C.cube(0), 2)
