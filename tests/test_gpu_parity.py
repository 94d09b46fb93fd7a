"""GPU parity: the HIP path against the CPU oracle on the same cages.

Topology (twin/edge/vertex IDs), uvIDs and creases must match bit for bit,
and so must the points of every rule on the default paths: gather rules keep
the reference's operation order, and scatter sums add the per-halfedge
contributions from zero in ascending halfedge order, as the reference's
sequential scatter does.  The optional per-element path keeps the
reference's float atomics and the optional ring-walk vertex kernel sums in
ring order; their scatter points are held to a tolerance relative to the
mesh scale against the double-precision oracle:

    max |p_gpu - p_fp64| / max|p_fp64| <= 1e-5
"""
import numpy as np
import pytest

from halfedgecatmullclark_amd import cage as C
from oracle import oracle as O
from tests.helpers import assert_bytes_equal, point_error, small_cages

pytestmark = pytest.mark.gpu

POINT_TOL = 1e-5

DRIVERS = [
    ("ccs_Refine_Scatter", O.ALL),
    ("ccs_Refine_Gather", O.ALL | O.GATHER),
    ("ccs_Refine_NoCreases_Scatter", O.ALL | O.NOCREASES),
    ("ccs_Refine_NoCreases_Gather", O.ALL | O.GATHER | O.NOCREASES),
]


@pytest.fixture(scope="module")
def capi():
    from halfedgecatmullclark_amd import capi as api
    api.lib()
    return api


@pytest.mark.parametrize("name,cage,depth", small_cages(), ids=[c[0] for c in small_cages()])
def test_topology_bitexact(capi, name, cage, depth):
    got = capi.refine_host(cage, depth, "ccs_RefineHalfedges", "ccs_RefineCreases", "ccs_RefineVertexUvs")
    want = O.refine(cage, depth, O.HALFEDGES | O.CREASES | O.UVS)
    assert_bytes_equal(got["halfedges"], want["halfedges"], f"{name} halfedges")
    assert_bytes_equal(got["creases"], want["creases"], f"{name} creases")


@pytest.mark.parametrize("driver,flags", DRIVERS, ids=[d[0] for d in DRIVERS])
@pytest.mark.parametrize("name,cage,depth", small_cages(), ids=[c[0] for c in small_cages()])
def test_refine_drivers(capi, name, cage, depth, driver, flags):
    got = capi.refine_host(cage, depth, driver)
    want = O.refine(cage, depth, flags)
    assert_bytes_equal(got["halfedges"], want["halfedges"], f"{name} halfedges")
    assert_bytes_equal(got["creases"], want["creases"], f"{name} creases")
    assert_bytes_equal(got["points"], want["points"], f"{name} points")
    if not flags & O.GATHER:
        ref64 = O.refine(cage, depth, flags, real=np.float64)["points"]
        err_max, err_rms = point_error(got["points"], ref64)
        assert err_max <= POINT_TOL, (err_max, err_rms)
        # and the float oracle (sequential scatter) is as close
        err32, _ = point_error(got["points"], want["points"])
        assert err32 <= POINT_TOL, err32


PATHS = [  # (fused, store, xcd remap, vertex rings); vertex kernel: the default (owner bits)
    (1, 1, 1, 3),
    (2, 1, 1, 2),
    (2, 0, 0, 2),
    (1, 1, 1, 2),
    (1, 1, 1, 1),
    (1, 0, 0, 0),
    (0, 1, 1, 1),
]
STEPS = [
    ("scatter", 15, O.ALL),
    ("gather", 15 | 16, O.ALL | O.GATHER),
    ("nocreases_scatter", 15 | 32, O.ALL | O.NOCREASES),
    ("nocreases_gather", 15 | 16 | 32, O.ALL | O.GATHER | O.NOCREASES),
]


@pytest.mark.parametrize("fused,store,remap,fan", PATHS, ids=["fused_vertexdirect", "patch_lds", "patch_direct", "fused_compact", "fused_fan", "fused_direct_walk", "per_element"])
@pytest.mark.parametrize("name,cage,depth", small_cages(), ids=[c[0] for c in small_cages()])
def test_device_paths(capi, name, cage, depth, fused, store, remap, fan):
    ds = capi.DeviceSubd(cage, depth)
    ds.set_option(capi.OPT_FUSED, fused)
    ds.set_option(capi.OPT_STORE, store)
    ds.set_option(capi.OPT_XCD_REMAP, remap)
    ds.set_option(capi.OPT_FAN, fan)
    for label, steps, flags in STEPS:
        ds.refine(steps)
        got = ds.download()
        want = O.refine(cage, depth, flags)
        assert_bytes_equal(got["halfedges"], want["halfedges"], f"{name}/{label} halfedges")
        assert_bytes_equal(got["creases"], want["creases"], f"{name}/{label} creases")
        if flags & O.GATHER:
            assert_bytes_equal(got["points"], want["points"], f"{name}/{label} points")
        else:
            ref64 = O.refine(cage, depth, flags, real=np.float64)["points"]
            err, _ = point_error(got["points"], ref64)
            assert err <= POINT_TOL, (label, err)
    ds.close()


DETERMINISTIC = [  # (fused, fan, scan): paths whose scatter sums run in ascending halfedge order
    (1, 2, 0),
    (1, 2, 4),
    (1, 1, 0),
    (1, 3, 0),
    (2, 2, 0),
]


@pytest.mark.parametrize("fused,fan,scan", DETERMINISTIC,
                         ids=["fused_masked", "fused_compact", "fused_fan", "fused_vertexdirect", "patch"])
@pytest.mark.parametrize("name,cage,depth", small_cages(), ids=[c[0] for c in small_cages()])
def test_scatter_points_bitexact(capi, name, cage, depth, fused, fan, scan):
    """Scatter vertex points are sums of per-halfedge contributions; the fused
    and patch paths add them from zero in ascending halfedge order, as the
    reference's sequential scatter does, so they equal the float oracle bit
    for bit (the per-element path and the ring-walk vertex kernel are held to
    the tolerance above)."""
    ds = capi.DeviceSubd(cage, depth)
    ds.set_option(capi.OPT_FUSED, fused)
    ds.set_option(capi.OPT_FAN, fan)
    ds.set_option(capi.OPT_SCAN, scan)
    for label, steps, flags in (STEPS[0], STEPS[2]):
        ds.refine(steps)
        got = ds.download()
        want = O.refine(cage, depth, flags)
        assert_bytes_equal(got["points"], want["points"], f"{name}/{label} scatter points")
    ds.close()
