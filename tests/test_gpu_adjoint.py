"""The adjoint of the point refinement on the GPU.

Refined points are linear in the cage points at fixed topology and
sharpness, S(V) = J V (per coordinate).  The adjoint must apply J^T:
  - dot-product identity <u, S(v)> = <S^T u, v> (forward from the FP64 oracle);
  - the reference's own Jacobian, built by running it on basis vectors
    (golden fixture, cube at depth 3): S^T u = J^T u to 1e-12;
  - FP64 central differences (h = 1e-6) of L(V) = 0.5 ||S(V) - T||^2.
FP64 tolerances are 1e-11 relative; the float adjoint is held to 1e-5.
"""
import os

import numpy as np
import pytest

from halfedgecatmullclark_amd import cage as C
from oracle import oracle as O
from tests.helpers import small_cages

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def capi():
    from halfedgecatmullclark_amd import capi as api
    api.lib()
    return api


def finest(out):
    n = out["counts"]
    D = len(n["V"]) - 1
    return O.level(out, "points", D)


def forward64(cage, depth, points, flags=O.ALL):
    cg = cage.copy()
    cg.vertex_points = np.asarray(points, np.float64)
    return finest(O.refine(cg, depth, flags, real=np.float64))


@pytest.mark.parametrize("name,cage,depth", small_cages(), ids=[c[0] for c in small_cages()])
def test_dot_product_identity(capi, name, cage, depth):
    depth = min(depth, 4)
    s = capi.HostSubd(cage, depth).run("ccs_RefineHalfedges", "ccs_RefineCreases")
    rng = np.random.default_rng(0)
    v = rng.normal(size=(cage.V, 3))
    Sv = forward64(cage, depth, v)
    u = rng.normal(size=Sv.shape)
    g64 = s.adjoint(u, np.float64)
    lhs, rhs = float(np.sum(u * Sv)), float(np.sum(g64 * v))
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + np.abs(u).sum()), (lhs, rhs)
    g32 = s.adjoint(u.astype(np.float32), np.float32)
    assert float(np.abs(g32 - g64).max()) <= 1e-5 * float(np.abs(g64).max())
    s.close()


@pytest.mark.parametrize("depth", [1, 2, 3])
@pytest.mark.parametrize("name,cage", [("irregular_l2", C.irregular(2, seed=3)), ("torus_r1.3",
                                        C.torus(6, 5, seed=4, ring_sharpness=((0, 1.3), (2, 0.4))))],
                         ids=["irregular_l2", "torus_r1.3"])
def test_dot_product_identity_shallow(capi, name, cage, depth):
    """Depths 1-3: the adjoint that copies the finest gradient (a cage
    refined once: its cage step updates the gradient in place) and the ones
    that read it where it lies, on a cage with a boundary and polygons other
    than quads, and on creased rings."""
    s = capi.HostSubd(cage, depth).run("ccs_RefineHalfedges", "ccs_RefineCreases")
    rng = np.random.default_rng(depth)
    v = rng.normal(size=(cage.V, 3))
    Sv = forward64(cage, depth, v)
    u = rng.normal(size=Sv.shape)
    g64 = s.adjoint(u, np.float64)
    lhs, rhs = float(np.sum(u * Sv)), float(np.sum(g64 * v))
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + np.abs(u).sum()), (lhs, rhs)
    g32 = s.adjoint(u.astype(np.float32), np.float32)
    assert float(np.abs(g32 - g64).max()) <= 1e-5 * float(np.abs(g64).max())
    s.close()


def test_dot_product_identity_cage_fitting_workload(capi):
    """The same identity on the cage-fitting workload itself: torus 32x16
    (2,048 halfedges, crease rings of sharpness 1.5 and 4.0) at depth 5."""
    cage, depth = C.torus(32, 16), 5
    s = capi.HostSubd(cage, depth).run("ccs_RefineHalfedges", "ccs_RefineCreases")
    rng = np.random.default_rng(1)
    v = rng.normal(size=(cage.V, 3))
    Sv = forward64(cage, depth, v)
    u = rng.normal(size=Sv.shape)
    g64 = s.adjoint(u, np.float64)
    lhs, rhs = float(np.sum(u * Sv)), float(np.sum(g64 * v))
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + np.abs(u).sum()), (lhs, rhs)
    g32 = s.adjoint(u.astype(np.float32), np.float32)
    assert float(np.abs(g32 - g64).max()) <= 1e-5 * float(np.abs(g64).max())
    s.close()


def test_against_reference_jacobian(capi):
    gold = np.load(os.path.join(HERE, "golden", "golden.npz"))
    J = gold["jacobian/cube_d3_scatter"]          # V_3 x V_0 (per coordinate)
    cube = C.load_ccm(os.path.join(HERE, "golden", "cages", "cube.ccm"))
    s = capi.HostSubd(cube, 3).run("ccs_RefineHalfedges", "ccs_RefineCreases")
    rng = np.random.default_rng(1)
    u = rng.normal(size=(J.shape[0], 3))
    got = s.adjoint(u, np.float64)
    np.testing.assert_allclose(got, J.T @ u, rtol=0, atol=1e-12 * np.abs(J.T @ u).max())
    s.close()


def test_central_differences(capi):
    cube = C.cube(0)
    D = 3
    rng = np.random.default_rng(2)
    V0 = cube.vertex_points.astype(np.float64)
    T = forward64(cube, D, V0 + rng.normal(0, 0.02, V0.shape))

    def loss(V):
        r = forward64(cube, D, V) - T
        return 0.5 * float(np.sum(r * r))

    s = capi.HostSubd(cube, D).run("ccs_RefineHalfedges", "ccs_RefineCreases")
    grad = s.adjoint(forward64(cube, D, V0) - T, np.float64)
    h = 1e-6
    fd = np.zeros_like(V0)
    for i in range(V0.shape[0]):
        for k in range(3):
            e = np.zeros_like(V0)
            e[i, k] = h
            fd[i, k] = (loss(V0 + e) - loss(V0 - e)) / (2 * h)
    np.testing.assert_allclose(grad, fd, rtol=0, atol=1e-7 * np.abs(fd).max())
    s.close()


@pytest.mark.parametrize("label,cage,depth", [("torus16x8_d5", C.torus(16, 8, seed=3), 5),
                                               ("irregular3_d4", C.irregular(3, seed=2), 4),
                                               ("cube_d4", C.cube(1), 4)])
@pytest.mark.parametrize("steps", [0, 32], ids=["creased", "nocreases"])
def test_vertex_pass_orders_agree(capi, label, cage, depth, steps):
    """The adjoint's vertex passes over the owners in halfedge order (with
    the rare-vertex kernel) and one thread per vertex visit the same rings
    in the same order, and every pass adds in a fixed order (no atomics): the
    two give the same bits, and a second run repeats them.  The device
    gradient, read where it lies, is left unchanged."""
    import torch
    ds = capi.DeviceSubd(cage, depth)
    ds.refine(capi.ALL)
    n = ds.download()["counts"]
    gen = torch.Generator(device="cuda").manual_seed(5)
    g = torch.randn((n["V"][depth], 3), generator=gen, device="cuda", dtype=torch.float32)
    g0 = g.clone()
    outs = []
    for per_vertex_max in (0, 1 << 30, 0):
        ds.set_option(capi.OPT_ADJ_VERTEX, per_vertex_max)
        out = torch.empty((cage.V, 3), device="cuda", dtype=torch.float32)
        ds.refine_points_adjoint(g.data_ptr(), out.data_ptr(), steps)
        ds.synchronize()
        outs.append(out.cpu().numpy())
    assert outs[0].tobytes() == outs[1].tobytes(), label
    assert outs[0].tobytes() == outs[2].tobytes(), label
    assert torch.equal(g, g0), label
    ds.close()


def test_parameter_cache_follows_rules_and_topology(capi):
    """The float adjoint keeps the vertex rules' parameters of the current
    topology and creases between calls: a cached run, a run after the rules
    change (creased / uncreased) and a run after the topology and creases
    are refined again give the bits of a fresh subd's first run."""
    import torch
    cage, depth = C.torus(6, 5, seed=4, ring_sharpness=((0, 1.3), (2, 0.4))), 5

    def first_run(steps):
        ds = capi.DeviceSubd(cage, depth)
        ds.refine(capi.ALL)
        out = run(ds, steps)
        ds.close()
        return out

    n = capi.DeviceSubd(cage, depth)
    n.refine(capi.ALL)
    V = n.download()["counts"]["V"]
    gen = torch.Generator(device="cuda").manual_seed(11)
    g = torch.randn((V[depth], 3), generator=gen, device="cuda", dtype=torch.float32)

    def run(ds, steps):
        out = torch.empty((cage.V, 3), device="cuda", dtype=torch.float32)
        ds.refine_points_adjoint(g.data_ptr(), out.data_ptr(), steps)
        ds.synchronize()
        return out.cpu().numpy().tobytes()

    want = {0: first_run(0), 32: first_run(32)}
    assert want[0] != want[32]
    got = [run(n, 0), run(n, 0), run(n, 32), run(n, 32), run(n, 0)]
    n.refine(capi.ALL)
    got += [run(n, 0), run(n, 0)]
    n.refine(capi.POINTS)      # points only: the cached parameters stay
    got.append(run(n, 0))
    for i, (steps, b) in enumerate(zip([0, 0, 32, 32, 0, 0, 0, 0], got)):
        assert b == want[steps], (i, steps)
    n.close()
