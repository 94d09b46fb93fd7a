"""The step after refinement: halfedge normals and OBJ export.

OBJ export (ccs_ExportToObj) against the reference example's own exporter
(examples/subd_cpu.c:22-115, compiled by oracle/Makefile from the reference
tree): both write every level of the same refined arrays, and the files must
be equal byte for byte, with and without uvs (CC_DISABLE_UV).

Halfedge normals (ccs_HalfedgeNormals[_Fast], ccd_HalfedgeNormals) follow the
reference's shaders (glsl/CatmullClark_Scatter.glsl:931-960); no CPU code of
the reference computes them, so the oracle's restatement is pinned by what
the geometry fixes:
- every normal has unit length;
- a flat cage in the plane z = 0 gives (0, 0, +-1) everywhere, one sign;
- a rotation of the cage rotates the normals (double build, 1e-12);
- on the closed cube every normal points away from the centroid;
- the vertex normal of a halfedge is the normalized sum of the corner
  normals of all halfedges leaving its vertex, grouped here by vertex ID
  (no ring walk), and equals the corner normal on a boundary.
The GPU normals are held to the float oracle at 1e-6 per component (unit
vectors: a few ulp of single-precision rounding in sqrt and division).
"""
import os

import numpy as np
import pytest

from halfedgecatmullclark_amd import cage as C
from oracle import oracle as O
from oracle import ref

HERE = os.path.dirname(os.path.abspath(__file__))
CAGES = os.path.join(HERE, "golden", "cages")
NORMAL_TOL = 1e-6


def flat_grid(nx=5, ny=4, seed=0):
    rng = np.random.default_rng(seed)
    pts = np.array([[x + rng.uniform(-0.2, 0.2), y + rng.uniform(-0.2, 0.2), 0.0]
                    for y in range(ny + 1) for x in range(nx + 1)], np.float32)
    faces = [[y * (nx + 1) + x, y * (nx + 1) + x + 1, (y + 1) * (nx + 1) + x + 1, (y + 1) * (nx + 1) + x]
             for y in range(ny) for x in range(nx)]
    return C.from_faces(pts, faces)


def rotation(seed=3):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def refined(cage, depth, real=np.float32):
    c = cage.copy()
    c.vertex_points = np.asarray(c.vertex_points, real)
    return O.refine(c, depth, O.ALL, real=real)


def group_by_vertex(he):
    order = np.argsort(he["vertexID"], kind="stable")
    v = he["vertexID"][order]
    starts = np.r_[0, np.nonzero(np.diff(v))[0] + 1]
    return order, starts


def boundary_vertices(he):
    """vertices with a boundary edge among the edges of their ring"""
    H = len(he)
    prev = (np.arange(H) & ~3) | ((np.arange(H) - 1) & 3)
    b = np.zeros(int(he["vertexID"].max()) + 1, bool)
    b[he["vertexID"][he["twinID"] < 0]] = True          # own edge on the boundary
    b[he["vertexID"][he["twinID"][prev] < 0]] = True    # previous edge on the boundary
    return b


# ---- oracle pins ----------------------------------------------------------------

@pytest.mark.parametrize("real", [np.float32, np.float64])
def test_flat_cage_normals_are_axis(real):
    r = refined(flat_grid(), 3, real)
    for smooth in (False, True):
        n = O.halfedge_normals(r, 3, smooth)
        assert np.all(n[:, :2] == 0)
        assert np.all(np.sign(n[:, 2]) == np.sign(n[0, 2]))
        assert np.abs(np.abs(n[:, 2]) - 1).max() <= 4 * np.finfo(real).eps


@pytest.mark.parametrize("name,cage,depth", [("cube", C.cube(0), 4), ("torus", C.torus(8, 6, seed=2), 3),
                                             ("irregular", C.irregular(2, seed=3), 3)])
def test_unit_length_and_rotation(name, cage, depth):
    r = refined(cage, depth, np.float64)
    R = rotation()
    rc = cage.copy()
    rc.vertex_points = np.asarray(cage.vertex_points, np.float64) @ R.T
    r2 = refined(rc, depth, np.float64)
    for smooth in (False, True):
        n = O.halfedge_normals(r, depth, smooth)
        assert np.abs(np.linalg.norm(n, axis=1) - 1).max() <= 1e-14
        n2 = O.halfedge_normals(r2, depth, smooth)
        np.testing.assert_allclose(n2, n @ R.T, rtol=0, atol=1e-12)
    n32 = O.halfedge_normals(refined(cage, depth), depth, True)
    assert np.abs(np.linalg.norm(n32, axis=1) - 1).max() <= 1e-6


def test_cube_normals_point_outward():
    r = refined(C.cube(0), 4)
    he, p = O.level(r, "halfedges", 4), O.level(r, "points", 4)
    c = p.mean(axis=0)
    for smooth in (False, True):
        n = O.halfedge_normals(r, 4, smooth)
        assert np.all(np.sum(n * (p[he["vertexID"]] - c), axis=1) > 0)


@pytest.mark.parametrize("name,cage,depth", [("torus", C.torus(8, 6, seed=2), 3),
                                             ("irregular", C.irregular(2, seed=3), 3),
                                             ("grid", flat_grid(4, 3, seed=1), 2)])
def test_vertex_normal_is_sum_over_vertex(name, cage, depth):
    r = refined(cage, depth, np.float64)
    he = O.level(r, "halfedges", depth)
    fast = O.halfedge_normals(r, depth, False)
    smooth = O.halfedge_normals(r, depth, True)
    order, starts = group_by_vertex(he)
    sums = np.add.reduceat(fast[order], starts, axis=0)
    verts = he["vertexID"][order][starts]
    want = np.zeros((int(verts.max()) + 1, 3))
    want[verts] = sums / np.linalg.norm(sums, axis=1, keepdims=True)
    bnd = boundary_vertices(he)
    inner = ~bnd[he["vertexID"]]
    np.testing.assert_allclose(smooth[inner], want[he["vertexID"]][inner], rtol=0, atol=1e-12)
    assert np.array_equal(smooth[~inner], fast[~inner])
    if name == "grid":
        assert (~inner).any() and inner.any()


# ---- OBJ export against the reference's exporter ---------------------------------

EXPORT_CASES = [("cube", 3), ("torus16x8", 3), ("irregular_l2", 2), ("Bishop", 2), ("irregular_soft", 2)]


@pytest.mark.skipif(not ref.export_available(), reason="the reference exporter is not built")
@pytest.mark.parametrize("variant", ["uv", "nouv"])
@pytest.mark.parametrize("name,depth", EXPORT_CASES, ids=[c[0] for c in EXPORT_CASES])
def test_export_matches_reference_exporter(name, depth, variant, tmp_path):
    """Both exporters write every level of the same arrays (the reference's
    own refinement, in a subd its library created)."""
    from halfedgecatmullclark_amd import capi
    cg = C.load_ccm(os.path.join(CAGES, name + ".ccm"))
    r = ref.RefSubd(cg, depth, variant=variant).run("ccs_Refine_Scatter")
    L = capi.lib(variant)
    for d in range(depth + 1):
        a, b = str(tmp_path / f"ref{d}.obj"), str(tmp_path / f"ours{d}.obj")
        ref.export_obj(r.subd, d, a, variant)
        assert L.ccs_ExportToObj(r.subd, d, b.encode()), capi.last_error(variant)
        ta, tb = open(a, "rb").read(), open(b, "rb").read()
        assert len(ta) > 0 and ta == tb, (name, d)
    r.close()


# ---- GPU ------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("name,depth", EXPORT_CASES, ids=[c[0] for c in EXPORT_CASES])
def test_gpu_refine_then_export_matches_reference(name, depth, tmp_path):
    """GPU refinement written by ccs_ExportToObj equals the reference's
    refinement written by its own exporter, level by level."""
    from halfedgecatmullclark_amd import capi
    if not ref.export_available():
        pytest.skip("the reference exporter is not built")
    cg = C.load_ccm(os.path.join(CAGES, name + ".ccm"))
    r = ref.RefSubd(cg, depth).run("ccs_Refine_Scatter")
    s = capi.HostSubd(cg, depth).run("ccs_Refine_Scatter")
    for d in range(depth + 1):
        a, b = str(tmp_path / f"ref{d}.obj"), str(tmp_path / f"gpu{d}.obj")
        ref.export_obj(r.subd, d, a)
        s.export_obj(d, b)
        assert open(a, "rb").read() == open(b, "rb").read(), (name, d)
    r.close()
    s.close()


NORMAL_CASES = [("cube_d5", C.cube(0), 5), ("torus16x8_d4", C.torus(16, 8, seed=1), 4),
                ("irregular_l3_d3", C.irregular(3, seed=11), 3), ("grid_d3", flat_grid(), 3)]


@pytest.mark.gpu
@pytest.mark.parametrize("name,cage,depth", NORMAL_CASES, ids=[c[0] for c in NORMAL_CASES])
def test_gpu_normals_match_oracle(name, cage, depth):
    from halfedgecatmullclark_amd import capi
    s = capi.HostSubd(cage, depth).run("ccs_Refine_Scatter")
    arrays = s.arrays()
    for smooth in (False, True):
        got = s.normals(smooth)
        want = O.halfedge_normals(arrays, depth, smooth)
        assert float(np.abs(got - want).max()) <= NORMAL_TOL, (name, smooth)
    s.close()


@pytest.mark.gpu
def test_gpu_device_normals_torus64x32_d8():
    """ccd_HalfedgeNormals on the bandwidth workload (537 M halfedges at
    depth 8) against the oracle on the same (bit-exact) refined arrays."""
    import torch
    from halfedgecatmullclark_amd import capi
    cage, D = C.torus(64, 32), 8
    ds = capi.DeviceSubd(cage, D)
    ds.refine(capi.ALL)
    arrays = ds.download()
    H = ds.n["H"][D]
    out = torch.empty((H, 3), device="cuda", dtype=torch.float32)
    for smooth in (True, False):
        ds.halfedge_normals(out.data_ptr(), smooth)
        ds.synchronize()
        got = out.cpu().numpy()
        want = O.halfedge_normals(arrays, D, smooth)
        assert float(np.abs(got - want).max()) <= NORMAL_TOL, smooth
        del got, want
    ds.close()
