"""The 12-byte-record variant (CC_DISABLE_UV): libcchip_nouv.so, built from the
same sources with cc_Halfedge_SemiRegular = {twinID, edgeID, vertexID}, as
the reference's CC_DISABLE_UV build (CatmullClark.h:47-55) and its GL demo
(examples/subd_gpu.c:11) lay it out.

Pins, from tests/golden/nouv_digests.json (written by the reference built
with CC_DISABLE_UV, one thread, every ccs_Refine_* driver):
- CPU: the oracle's 16-byte records with the uvID dropped, its creases and
  points equal the reference's 12-byte digests (the variant changes the
  record layout and nothing else);
- CPU: the library exports no uv entry point, as the reference declares
  none under CC_DISABLE_UV;
- GPU: the drop-in drivers of libcchip_nouv.so equal the digests bit for
  bit, and its device-resident, batched and adjoint paths equal the 16-byte
  library's with the uvID dropped.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from halfedgecatmullclark_amd import cage as C
from oracle import oracle as O
from tests.helpers import assert_bytes_equal

HERE = os.path.dirname(os.path.abspath(__file__))
CAGES = os.path.join(HERE, "golden", "cages")
GOLD = json.load(open(os.path.join(HERE, "golden", "nouv_digests.json")))
CASES = sorted({(k.split("/")[0], int(k.split("/")[1])) for k in GOLD})
FLAGS = {
    "ccs_Refine_Scatter": O.ALL,
    "ccs_Refine_Gather": O.ALL | O.GATHER,
    "ccs_Refine_NoCreases_Scatter": O.ALL | O.NOCREASES,
    "ccs_Refine_NoCreases_Gather": O.ALL | O.GATHER | O.NOCREASES,
}


def digest(a) -> str:
    return hashlib.sha256(memoryview(np.ascontiguousarray(a).reshape(-1).view("u1"))).hexdigest()


def drop_uv(he: np.ndarray) -> np.ndarray:
    """16-byte records as the 12-byte layout (twin, edge, vertex)"""
    from halfedgecatmullclark_amd.abi import HALFEDGE12_DTYPE
    out = np.empty(len(he), HALFEDGE12_DTYPE)
    for k in ("twinID", "edgeID", "vertexID"):
        out[k] = he[k]
    return out


@pytest.mark.parametrize("drv", list(FLAGS))
@pytest.mark.parametrize("name,depth", CASES, ids=[c[0] for c in CASES])
def test_oracle_without_uv_matches_reference_digests(name, depth, drv):
    cg = C.load_ccm(os.path.join(CAGES, name + ".ccm"))
    got = O.refine(cg, depth, FLAGS[drv] & ~O.UVS)
    want = GOLD[f"{name}/{depth}/{drv}"]
    assert digest(drop_uv(got["halfedges"])) == want["halfedges"]
    assert digest(got["creases"]) == want["creases"]
    assert digest(got["points"]) == want["points"]


def test_nouv_library_exports():
    from halfedgecatmullclark_amd import capi
    L16, L12 = capi.lib("uv"), capi.lib("nouv")
    for name in ("ccs_RefineVertexUvs", "ccs_RefineCageVertexUvs", "ccs_HalfedgeVertexUv"):
        assert hasattr(L16, name) and not hasattr(L12, name), name
    for name in ("ccs_Refine_Scatter", "ccs_Refine_Gather", "ccs_RefineHalfedges", "ccd_Refine", "ccb_Refine",
                 "ccd_RefinePointsAdjoint", "ccd_RefineNoAlloc", "ccm_Load"):
        assert hasattr(L12, name), name


@pytest.mark.gpu
@pytest.mark.parametrize("drv", list(FLAGS))
@pytest.mark.parametrize("name,depth", CASES, ids=[c[0] for c in CASES])
def test_gpu_nouv_matches_reference_digests(name, depth, drv):
    from halfedgecatmullclark_amd import capi
    got = capi.refine_file(os.path.join(CAGES, name + ".ccm"), depth, drv, variant="nouv")
    assert got["halfedges"].dtype.itemsize == 12
    want = GOLD[f"{name}/{depth}/{drv}"]
    for k in ("halfedges", "creases", "points"):
        assert digest(got[k]) == want[k], f"{name} depth {depth} {drv} {k}"


DEVICE_CASES = [("torus16x8_d5", C.torus(16, 8, seed=1), 5), ("irregular_l3_d4", C.irregular(3, seed=11), 4),
                ("cube_d6", C.cube(0), 6)]


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [1, 2], ids=["fused", "patch"])
@pytest.mark.parametrize("name,cage,depth", DEVICE_CASES, ids=[c[0] for c in DEVICE_CASES])
def test_gpu_nouv_device_paths_equal_uv_library(name, cage, depth, fused):
    from halfedgecatmullclark_amd import capi
    for steps in (capi.ALL, capi.ALL | capi.GATHER, capi.HALFEDGES | capi.CREASES):
        outs = {}
        for variant in ("uv", "nouv"):
            ds = capi.DeviceSubd(cage, depth, variant=variant)
            ds.set_option(capi.OPT_FUSED, fused)
            ds.refine(steps)
            outs[variant] = ds.download()
            ds.close()
        assert_bytes_equal(outs["nouv"]["halfedges"], drop_uv(outs["uv"]["halfedges"]), f"{name} {steps} halfedges")
        assert_bytes_equal(outs["nouv"]["creases"], outs["uv"]["creases"], f"{name} {steps} creases")
        if steps & capi.POINTS:
            assert_bytes_equal(outs["nouv"]["points"], outs["uv"]["points"], f"{name} {steps} points")


@pytest.mark.gpu
def test_gpu_nouv_batch_and_adjoint_equal_uv_library():
    import torch
    from halfedgecatmullclark_amd import capi
    cages = [c for _, c, _ in DEVICE_CASES] + [C.torus(6, 5, seed=2)]
    depths = [d for _, _, d in DEVICE_CASES] + [2]
    got = {}
    for variant in ("uv", "nouv"):
        b = capi.Batch(cages, depths, variant=variant)
        b.refine(capi.ALL)
        nf = b.count(capi.COUNT_FINEST_POINTS)
        g = torch.randn((nf, 3), generator=torch.Generator(device="cuda").manual_seed(3), device="cuda")
        out = torch.empty((b.count(capi.COUNT_CAGE_POINTS), 3), device="cuda")
        b.refine_points_adjoint(g.data_ptr(), out.data_ptr(), 0)
        b.synchronize()
        got[variant] = ([b.download(i) for i in range(len(cages))], out.cpu().numpy())
        b.close()
    for i in range(len(cages)):
        u, n = got["uv"][0][i], got["nouv"][0][i]
        assert_bytes_equal(n["halfedges"], drop_uv(u["halfedges"]), f"cage {i} halfedges")
        assert_bytes_equal(n["creases"], u["creases"], f"cage {i} creases")
        assert_bytes_equal(n["points"], u["points"], f"cage {i} points")
    assert_bytes_equal(got["nouv"][1], got["uv"][1], "batched adjoint")


def test_nouv_library_exports_every_symbol_its_header_declares():
    """The public header preprocessed with CC_DISABLE_UV (the declarations a
    caller of the 12-byte build sees) against libcchip_nouv.so's exports."""
    import re
    import subprocess
    from halfedgecatmullclark_amd import capi
    hdr = os.path.join(os.path.dirname(HERE), "include", "CatmullClarkHIP.h")
    text = subprocess.run(["gcc", "-E", "-P", "-DCC_DISABLE_UV", "-x", "c", hdr], capture_output=True, text=True,
                          check=True).stdout
    names = set(re.findall(r"\b((?:ccs|ccm|ccd|ccb)_[A-Za-z0-9_]+)\s*\(", text))
    assert len(names) > 100 and "ccs_RefineVertexUvs" not in names
    L = capi.lib("nouv")
    missing = sorted(n for n in names if not hasattr(L, n))
    assert not missing, missing
