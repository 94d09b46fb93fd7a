"""Every kernel configuration the autotuner can pick gives the oracle's bytes.

The autotuner (ccd_Autotune, ccb_Autotune) picks per level among:
  - the fused kernel's workgroup size (256, 512, 1024 threads) and tiles of
    that size processed in turn per workgroup (1, 2, 4);
  - the masked vertex kernel compiled for 4, 6 or 8 waves per SIMD, or the
    compacted vertex kernel testing 2, 4 or 8 halfedges per thread;
  - the patch-local kernel (tested in test_gpu_parity and test_gpu_batch).
Each candidate is set for every level with the option keys and the whole
refinement is compared with the float oracle, bit for bit (halfedges,
creases, points), for the scatter and gather rules, on one cage at a time
and on a batch.
"""
import itertools

import pytest

from halfedgecatmullclark_amd import cage as C
from oracle import oracle as O
from tests.helpers import assert_bytes_equal, small_cages

pytestmark = pytest.mark.gpu

GEOMETRY = list(itertools.product((256, 512, 1024), (1, 2, 4)))
VERTEX = [(0, 4), (0, 6), (0, 8), (2, 8), (4, 8), (8, 8)]   # (scan, waves)
RULES = [("scatter", 15, O.ALL), ("gather", 15 | 16, O.ALL | O.GATHER)]


@pytest.fixture(scope="module")
def capi():
    from halfedgecatmullclark_amd import capi as api
    api.lib()
    return api


def cases():
    by_name = {n: (c, d) for n, c, d in small_cages()}
    return [("torus16x8_d5", C.torus(16, 8, seed=1), 5), ("torus_r1.3",) + by_name["torus_r1.3"],
            ("irregular_l3",) + by_name["irregular_l3"], ("cube_s7",) + by_name["cube_s7"]]


def set_all(obj, capi, fb, ft, scan, waves):
    obj.set_option(capi.OPT_FUSED, 1)
    obj.set_option(capi.OPT_FUSED_BLOCK, fb)
    obj.set_option(capi.OPT_FUSED_TILES, ft)
    obj.set_option(capi.OPT_SCAN, scan)
    obj.set_option(capi.OPT_VERTEX_WAVES, waves)


@pytest.mark.parametrize("name,cage,depth", cases(), ids=[c[0] for c in cases()])
def test_single_cage_candidates_bitexact(capi, name, cage, depth):
    want = {label: O.refine(cage, depth, flags) for label, _, flags in RULES}
    ds = capi.DeviceSubd(cage, depth)
    for (fb, ft), (scan, waves) in itertools.product(GEOMETRY, VERTEX):
        set_all(ds, capi, fb, ft, scan, waves)
        for o in ds.level_options():
            assert (o["block"], o["tiles"], o["scan"], o["waves"]) == (fb, ft, scan, waves)
        for label, steps, _ in RULES:
            ds.refine(steps)
            got = ds.download()
            what = f"{name} fb={fb} ft={ft} scan={scan} waves={waves} {label}"
            for k in ("halfedges", "creases", "points"):
                assert_bytes_equal(got[k], want[label][k], f"{what} {k}")
    ds.close()


def test_batch_candidates_bitexact(capi):
    cs = [(c, d) for _, c, d in cases()] + [(C.cube(0), 1), (C.torus(6, 5, seed=2), 2)]
    want = {label: [O.refine(c, d, flags) for c, d in cs] for label, _, flags in RULES}
    b = capi.Batch([c for c, _ in cs], [d for _, d in cs])
    for (fb, ft), (scan, waves) in itertools.product(GEOMETRY, VERTEX):
        set_all(b, capi, fb, ft, scan, waves)
        for label, steps, _ in RULES:
            b.refine(steps)
            b.synchronize()
            for i in range(len(cs)):
                got = b.download(i)
                what = f"cage {i} fb={fb} ft={ft} scan={scan} waves={waves} {label}"
                for k in ("halfedges", "creases", "points"):
                    assert_bytes_equal(got[k], want[label][i][k], f"{what} {k}")
    b.close()


def test_autotune_picks_are_candidates(capi):
    cage, depth = C.torus(16, 8, seed=1), 6
    ds = capi.DeviceSubd(cage, depth)
    ds.refine(capi.ALL)
    ds.autotune(capi.ALL)
    opts = ds.level_options()
    assert len(opts) == depth - 1
    for o in opts:
        assert o["kind"] in (1, 2)
        if o["kind"] == 1:
            assert (o["block"], o["tiles"]) in GEOMETRY and (o["scan"], o["waves"]) in VERTEX, o
        else:
            assert o["tile"] in (128, 256, 512), o
    ds.refine(capi.ALL)
    got = ds.download()
    want = O.refine(cage, depth, O.ALL)
    for k in ("halfedges", "creases", "points"):
        assert_bytes_equal(got[k], want[k], f"autotuned {k}")
    ds.close()


def test_option_values_refused(capi):
    cage, depth = C.torus(6, 5, seed=2), 3
    ds = capi.DeviceSubd(cage, depth)
    b = capi.Batch([cage], [depth])
    for obj in (ds, b):
        for key, bad in ((capi.OPT_FUSED_BLOCK, 128), (capi.OPT_FUSED_TILES, 3), (capi.OPT_VERTEX_WAVES, 5)):
            with pytest.raises(capi.CCError):
                obj.set_option(key, bad)
    ds.close()
    b.close()
