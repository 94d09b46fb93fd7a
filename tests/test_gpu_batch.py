"""Batched refinement (ccb_*): many cages of uneven size and depth in one
launch per kernel kind and level, replayed as a hipGraph.

Every cage of a batch must come out exactly as the same cage refined alone
through ccd_* (same kernels, same operation order: bytes equal), and equal
the oracle bit for bit (topology, uvIDs, creases and points of every rule;
scatter points are also checked within 1e-5 of the mesh scale of the
double-precision oracle).  The adjoint adds in a fixed order everywhere
(no atomics), so the batched adjoint equals the single-cage one bit for bit.
"""
import numpy as np
import pytest

from halfedgecatmullclark_amd import cage as C
from oracle import oracle as O
from tests.helpers import assert_bytes_equal, point_error, small_cages

pytestmark = pytest.mark.gpu

POINT_TOL = 1e-5


def tiny_cages():
    """One quad and one triangle: levels much shorter than a workgroup."""
    quad = C.from_faces(np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0.2], [0, 1, 0]], np.float32), [[0, 1, 2, 3]])
    tri = C.from_faces(np.array([[0, 0, 0], [1, 0, 0.1], [0, 1, 0]], np.float32), [[0, 1, 2]])
    return [(quad, 2), (tri, 2), (quad, 3)]


def batch_cases():
    cases = [(c, d) for _, c, d in small_cages()]
    cases += [(C.cube(0), 1), (C.torus(6, 5, seed=2), 2), (C.irregular(2, seed=8), 1)]
    return cases + tiny_cages()


@pytest.mark.parametrize("k", range(3))
def test_tiny_cages_single(capi, k):
    cage, depth = tiny_cages()[k]
    for steps, flags in ((15, O.ALL), (15 | 16, O.ALL | O.GATHER)):
        ds = capi.DeviceSubd(cage, depth)
        ds.refine(steps)
        got = ds.download()
        want = O.refine(cage, depth, flags)
        assert_bytes_equal(got["halfedges"], want["halfedges"], "halfedges")
        assert_bytes_equal(got["creases"], want["creases"], "creases")
        assert_bytes_equal(got["points"], want["points"], "points")
        if not flags & O.GATHER:
            ref64 = O.refine(cage, depth, flags, real=np.float64)["points"]
            assert point_error(got["points"], ref64)[0] <= POINT_TOL
        ds.close()


@pytest.fixture(scope="module")
def capi():
    from halfedgecatmullclark_amd import capi as api
    api.lib()
    return api


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


STEPS = [
    ("scatter", 15, O.ALL),
    ("gather", 15 | 16, O.ALL | O.GATHER),
    ("nocreases_scatter", 15 | 32, O.ALL | O.NOCREASES),
    ("nocreases_gather", 15 | 16 | 32, O.ALL | O.GATHER | O.NOCREASES),
]


def download_all(b):
    return [b.download(i) for i in range(len(b))]


@pytest.mark.parametrize("label,steps,flags", STEPS, ids=[s[0] for s in STEPS])
def test_batch_matches_oracle(capi, label, steps, flags):
    cases = batch_cases()
    b = capi.Batch([c for c, _ in cases], [d for _, d in cases])
    b.refine(steps)
    b.synchronize()
    for i, (cage, depth) in enumerate(cases):
        got = b.download(i)
        want = O.refine(cage, depth, flags)
        assert_bytes_equal(got["halfedges"], want["halfedges"], f"cage {i}/{label} halfedges")
        assert_bytes_equal(got["creases"], want["creases"], f"cage {i}/{label} creases")
        assert_bytes_equal(got["points"], want["points"], f"cage {i}/{label} points")
        if not flags & O.GATHER:
            ref64 = O.refine(cage, depth, flags, real=np.float64)["points"]
            assert point_error(got["points"], ref64)[0] <= POINT_TOL, (i, label)
    b.close()


@pytest.mark.parametrize("label,steps,flags", STEPS[:2], ids=[s[0] for s in STEPS[:2]])
def test_batch_equals_single_cage(capi, label, steps, flags):
    cases = batch_cases()
    b = capi.Batch([c for c, _ in cases], [d for _, d in cases])
    b.refine(steps)
    got = download_all(b)
    for i, (cage, depth) in enumerate(cases):
        ds = capi.DeviceSubd(cage, depth)
        ds.refine(steps)
        want = ds.download()
        for k in ("halfedges", "creases", "points"):
            assert_bytes_equal(got[i][k], want[k], f"cage {i}/{label} {k}")
        ds.close()
    b.close()


@pytest.mark.parametrize("key,value", [("kind", 2), ("scan", 2), ("scan", 4), ("scan", 8), ("tile", 128),
                                       ("tile", 512), ("remap", 0)])
def test_batch_kernel_options_agree(capi, key, value):
    cases = batch_cases()
    b = capi.Batch([c for c, _ in cases], [d for _, d in cases])
    b.refine(capi.ALL)
    want = download_all(b)
    if key == "kind":
        b.set_option(capi.OPT_FUSED, value)
    elif key == "scan":
        b.set_option(capi.OPT_SCAN, value)
    elif key == "tile":
        b.set_option(capi.OPT_FUSED, 2)
        b.set_option(capi.OPT_PATCH_TILE, value)
    else:
        b.set_option(capi.OPT_XCD_REMAP, value)
    b.refine(capi.ALL)
    got = download_all(b)
    for i in range(len(cases)):
        for k in ("halfedges", "creases", "points"):
            assert_bytes_equal(got[i][k], want[i][k], f"{key}={value} cage {i} {k}")
    b.close()


def test_graph_replays_match_eager(capi, torch):
    cases = batch_cases()
    cages = [c for c, _ in cases]
    depths = [d for _, d in cases]
    eager = capi.Batch(cages, depths)
    eager.set_option(capi.OPT_GRAPH, 0)
    eager.refine(capi.ALL)
    want = download_all(eager)
    eager.close()
    b = capi.Batch(cages, depths)
    other = capi.DeviceSubd(C.torus(8, 4), 3)
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(4):
        # unrelated eager work and synchronizations between the replays
        x = torch.randn(1 << 16, device="cuda")
        _ = (x * 2).sum().item()
        other.refine(capi.ALL, stream)
        b.refine(capi.ALL, stream)
        torch.cuda.synchronize()
        got = download_all(b)
        for i in range(len(cases)):
            for k in ("halfedges", "creases", "points"):
                assert_bytes_equal(got[i][k], want[i][k], f"replay {rep} cage {i} {k}")
    other.close()
    b.close()


@pytest.mark.parametrize("label,steps,flags", STEPS[:2], ids=[s[0] for s in STEPS[:2]])
def test_single_cage_graph_replays_match_oracle(capi, torch, label, steps, flags):
    """ccd_Refine replays a captured graph: every replay, with unrelated
    torch work between them, new cage points uploaded before some, and the
    graphs dropped by an option change halfway, equals the oracle."""
    cage, depth = C.torus(6, 5, seed=4, ring_sharpness=((0, 1.3), (2, 0.4))), 5
    rng = np.random.default_rng(3)
    ds = capi.DeviceSubd(cage, depth)
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(6):
        x = torch.randn(1 << 16, device="cuda")
        _ = (x * 2).sum().item()
        if rep in (2, 5):
            cage.vertex_points = (cage.vertex_points + rng.normal(0, 0.05, cage.vertex_points.shape)).astype(np.float32)
            ds.upload_cage_points(cage.vertex_points)
        if rep == 3:
            ds.set_option(capi.OPT_SCAN, 4)
        ds.refine(steps, stream)
        torch.cuda.synchronize()
        got = ds.download()
        want = O.refine(cage, depth, flags)
        for k in ("halfedges", "creases", "points"):
            assert_bytes_equal(got[k], want[k], f"replay {rep} {k}")
    ds.close()


def test_refine_inside_a_caller_graph_capture(capi, torch):
    """ccd_Refine and ccb_Refine on a stream the caller is capturing put their
    launches into the caller's graph (no capture of their own): the torch
    graph, replayed with eager work and new cage points between replays,
    equals the oracle each time."""
    cage, depth = C.torus(6, 5, seed=4, ring_sharpness=((0, 1.3), (2, 0.4))), 4
    rng = np.random.default_rng(21)
    ds = capi.DeviceSubd(cage, depth)
    b = capi.Batch([cage], [depth])
    ds.refine(capi.ALL)
    b.refine(capi.ALL)               # warm-up outside the capture (tables, first launches)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        ds.refine(capi.ALL, s.cuda_stream)
        b.refine(capi.ALL, s.cuda_stream)
    for rep in range(3):
        x = torch.randn(1 << 16, device="cuda")
        _ = (x * 2).sum().item()
        if rep:
            cage.vertex_points = (cage.vertex_points + rng.normal(0, 0.05, cage.vertex_points.shape)).astype(np.float32)
            ds.upload_cage_points(cage.vertex_points)
        g.replay()
        torch.cuda.synchronize()
        want = O.refine(cage, depth, O.ALL)
        got = ds.download()
        for k in ("halfedges", "creases", "points"):
            assert_bytes_equal(got[k], want[k], f"ccd replay {rep} {k}")
        if rep == 0:   # the batch's cage points were not moved
            gb = b.download(0)
            for k in ("halfedges", "creases", "points"):
                assert_bytes_equal(gb[k], want[k], f"ccb replay {rep} {k}")
    del g
    ds.close()
    b.close()


def test_batch_refine_points_and_finest(capi, torch):
    cases = batch_cases()
    cages = [c for c, _ in cases]
    depths = [d for _, d in cases]
    b = capi.Batch(cages, depths)
    b.refine(capi.ALL)
    rng = np.random.default_rng(7)
    moved = [c.vertex_points + rng.normal(0.0, 0.02, c.vertex_points.shape).astype(np.float32) for c in cages]
    dev = torch.from_numpy(np.concatenate(moved).astype(np.float32)).cuda()
    assert dev.shape[0] == b.count(capi.COUNT_CAGE_POINTS)
    b.refine_points(dev.data_ptr(), 0)
    fin = torch.empty((b.count(capi.COUNT_FINEST_POINTS), 3), dtype=torch.float32, device="cuda")
    b.finest_points(fin.data_ptr())
    b.synchronize()
    fin = fin.cpu().numpy()
    for i, (cage, depth) in enumerate(cases):
        ds = capi.DeviceSubd(cage, depth)
        ds.refine(capi.ALL)
        ds.upload_cage_points(moved[i])
        ds.refine(capi.POINTS)
        want = ds.download()
        got = b.download(i)
        assert_bytes_equal(got["points"], want["points"], f"cage {i} points")
        n = want["counts"]
        o = b.offset(capi.COUNT_FINEST_POINTS, i)
        nf = n["V"][depth]
        assert_bytes_equal(fin[o:o + nf], want["points"][-nf:], f"cage {i} finest points")
        ds.close()
    b.close()


def test_batch_adjoint_matches_single_cage(capi, torch):
    cases = batch_cases()
    cages = [c for c, _ in cases]
    depths = [d for _, d in cases]
    b = capi.Batch(cages, depths)
    b.refine(capi.ALL)
    nf = b.count(capi.COUNT_FINEST_POINTS)
    gen = torch.Generator(device="cuda").manual_seed(3)
    g = torch.randn((nf, 3), generator=gen, device="cuda", dtype=torch.float32)
    out = torch.empty((b.count(capi.COUNT_CAGE_POINTS), 3), device="cuda", dtype=torch.float32)
    b.refine_points_adjoint(g.data_ptr(), out.data_ptr(), 0)
    b.synchronize()
    out = out.cpu().numpy()
    g_host = g.cpu().numpy()
    for i, (cage, depth) in enumerate(cases):
        ds = capi.DeviceSubd(cage, depth)
        ds.refine(capi.ALL)
        o = b.offset(capi.COUNT_FINEST_POINTS, i)
        n = b.count(capi.COUNT_FINEST_POINTS, i)
        gi = torch.from_numpy(np.ascontiguousarray(g_host[o:o + n])).cuda()
        oi = torch.empty((cage.V, 3), device="cuda", dtype=torch.float32)
        ds.refine_points_adjoint(gi.data_ptr(), oi.data_ptr(), 0)
        ds.synchronize()
        want = oi.cpu().numpy()
        oc = b.offset(capi.COUNT_CAGE_POINTS, i)
        got = out[oc:oc + cage.V]
        assert_bytes_equal(np.ascontiguousarray(got), want, f"cage {i} adjoint")
        ds.close()
    b.close()


def test_batch_adjoint_reads_callers_gradient(capi, torch):
    """With every cage refined at least twice, the batched adjoint's finest
    steps read the caller's gradient where it lies (its address is written
    before each run; the captured graph replays): two calls with gradients
    in different buffers each equal the single-cage adjoints bit for bit,
    and leave the callers' gradients unchanged."""
    cases = [(c, d) for c, d in batch_cases() if d >= 2]
    cages = [c for c, _ in cases]
    depths = [d for _, d in cases]
    b = capi.Batch(cages, depths)
    b.refine(capi.ALL)
    nf = b.count(capi.COUNT_FINEST_POINTS)
    singles = []
    for cage, depth in cases:
        ds = capi.DeviceSubd(cage, depth)
        ds.refine(capi.ALL)
        singles.append(ds)
    for seed in (5, 6):
        gen = torch.Generator(device="cuda").manual_seed(seed)
        g = torch.randn((nf, 3), generator=gen, device="cuda", dtype=torch.float32)
        g0 = g.clone()
        out = torch.empty((b.count(capi.COUNT_CAGE_POINTS), 3), device="cuda", dtype=torch.float32)
        b.refine_points_adjoint(g.data_ptr(), out.data_ptr(), 0)
        b.synchronize()
        assert torch.equal(g, g0), "the batched adjoint wrote into its input"
        got_all = out.cpu().numpy()
        g_host = g.cpu().numpy()
        for i, (cage, depth) in enumerate(cases):
            o = b.offset(capi.COUNT_FINEST_POINTS, i)
            n = b.count(capi.COUNT_FINEST_POINTS, i)
            gi = torch.from_numpy(np.ascontiguousarray(g_host[o:o + n])).cuda()
            gi0 = gi.clone()
            oi = torch.empty((cage.V, 3), device="cuda", dtype=torch.float32)
            singles[i].refine_points_adjoint(gi.data_ptr(), oi.data_ptr(), 0)
            singles[i].synchronize()
            assert torch.equal(gi, gi0), f"cage {i}: the adjoint wrote into its input"
            oc = b.offset(capi.COUNT_CAGE_POINTS, i)
            assert_bytes_equal(np.ascontiguousarray(got_all[oc:oc + cage.V]), oi.cpu().numpy(),
                               f"seed {seed} cage {i} adjoint")
    for ds in singles:
        ds.close()
    b.close()


def test_adjoint_with_and_without_kept_bits(capi, torch):
    """The adjoint's owner scan reads the edge-representative bits only once
    a forward pass has written them.  A forward pass through the patch
    kernel writes none: the adjoint that follows, with the fused kernel
    selected again, must fall back to the descent, and equal (bit for bit)
    the adjoint after a fused forward pass and that of a fresh batch."""
    cases = batch_cases()
    cages = [c for c, _ in cases]
    depths = [d for _, d in cases]
    ref = capi.Batch(cages, depths)
    ref.refine(capi.ALL)
    nf = ref.count(capi.COUNT_FINEST_POINTS)
    gen = torch.Generator(device="cuda").manual_seed(5)
    g = torch.randn((nf, 3), generator=gen, device="cuda", dtype=torch.float32)
    nc = ref.count(capi.COUNT_CAGE_POINTS)
    want = torch.empty((nc, 3), device="cuda", dtype=torch.float32)
    ref.refine_points_adjoint(g.data_ptr(), want.data_ptr(), 0)
    ref.synchronize()
    ref.close()
    b = capi.Batch(cages, depths)
    b.set_option(capi.OPT_FUSED, 2)
    b.refine(capi.ALL)
    b.set_option(capi.OPT_FUSED, 1)
    outs = []
    for fwd in (False, True):
        if fwd:
            b.refine(capi.ALL)
        out = torch.empty((nc, 3), device="cuda", dtype=torch.float32)
        b.refine_points_adjoint(g.data_ptr(), out.data_ptr(), 0)
        b.synchronize()
        outs.append(out.cpu().numpy())
    b.close()
    w = want.cpu().numpy()
    assert_bytes_equal(outs[0], w, "adjoint after a patch forward")
    assert_bytes_equal(outs[1], w, "adjoint after a fused forward")
    # the same on one cage through the single-cage path
    cage, depth = C.torus(16, 8, seed=1), 5
    grads = []
    for kind in (1, 2):
        ds = capi.DeviceSubd(cage, depth)
        ds.set_option(capi.OPT_FUSED, kind)
        ds.refine(capi.ALL)
        ds.set_option(capi.OPT_FUSED, 1)
        gf = torch.randn((ds.n["V"][depth], 3), generator=torch.Generator(device="cuda").manual_seed(7),
                         device="cuda", dtype=torch.float32)
        oc = torch.empty((cage.V, 3), device="cuda", dtype=torch.float32)
        ds.refine_points_adjoint(gf.data_ptr(), oc.data_ptr(), 0)
        ds.synchronize()
        grads.append(oc.cpu().numpy())
        ds.close()
    assert_bytes_equal(grads[1], grads[0], "single-cage adjoint after a patch forward")


def test_batch_autotune_keeps_results(capi):
    cases = batch_cases()
    b = capi.Batch([c for c, _ in cases], [d for _, d in cases])
    b.refine(capi.ALL)
    want = download_all(b)
    b.autotune(capi.ALL)
    opts = b.level_options()
    assert len(opts) == max(d for _, d in cases) - 1
    for o in opts:
        assert o["kind"] in (1, 2) and o["scan"] in (0, 2, 4, 8) and o["tile"] in (128, 256, 512)
    b.refine(capi.ALL)
    got = download_all(b)
    for i in range(len(cases)):
        for k in ("halfedges", "creases", "points"):
            assert_bytes_equal(got[i][k], want[i][k], f"autotuned cage {i} {k}")
    b.close()


def test_single_cage_autotune_keeps_results(capi):
    cage, depth = C.torus(16, 8, seed=1), 5
    ds = capi.DeviceSubd(cage, depth)
    ds.refine(capi.ALL)
    want = ds.download()
    ds.autotune(capi.ALL)
    assert len(ds.level_options()) == depth - 1
    ds.refine(capi.ALL)
    got = ds.download()
    for k in ("halfedges", "creases", "points"):
        assert_bytes_equal(got[k], want[k], f"autotuned {k}")
    ds.close()


def test_batch_errors(capi):
    cases = batch_cases()[:2]
    b = capi.Batch([c for c, _ in cases], [d for _, d in cases])
    with pytest.raises(capi.CCError):
        b.refine(capi.HALFEDGES)            # points are part of every batched refinement
    with pytest.raises(capi.CCError):
        b.refine_points(None, 0)            # topology not refined yet
    with pytest.raises(capi.CCError):
        b.set_option(capi.OPT_PATCH_TILE, 100)
    b.close()


def test_batch_timings_are_the_last_run(capi):
    cases = batch_cases()
    b = capi.Batch([c for c, _ in cases], [d for _, d in cases])
    b.refine(capi.ALL)
    b.set_timing(True)
    b.refine(capi.ALL)
    b.synchronize()
    first = b.timings()
    b.refine(capi.ALL)
    b.synchronize()
    second = b.timings()
    b.set_timing(False)
    assert len(first) == len(second) > 0
    assert [(n, lv) for n, lv, _ in first] == [(n, lv) for n, lv, _ in second]
    assert all(ms >= 0.0 for _, _, ms in second)
    b.close()


def oracle_jacobian(cage, depth, flags=O.ALL):
    """J (V_D x V_0) of the double-precision oracle, one basis vector per
    cage point (the points are linear in the cage points, per coordinate)."""
    cols = []
    for i in range(cage.V):
        basis = cage.copy()
        basis.vertex_points = np.zeros((cage.V, 3), np.float64)
        basis.vertex_points[i, 0] = 1.0
        cols.append(O.level(O.refine(basis, depth, flags, real=np.float64), "points", depth)[:, 0])
    return np.stack(cols, axis=1)


def test_batch_adjoint_parameter_cache(capi, torch):
    """The batched adjoint keeps its vertex-rule parameters after the first
    call for a rule set: later calls (graph replays without the parameter
    pass), and switching between the creased and uncreased rules, give the
    bits of a fresh batch's first call."""
    cases = batch_cases()
    cages = [c for c, _ in cases]
    depths = [d for _, d in cases]

    def fresh(steps):
        b = capi.Batch(cages, depths)
        b.refine(capi.ALL)
        r = adj(b, steps)
        b.close()
        return r

    b = capi.Batch(cages, depths)
    b.refine(capi.ALL)
    gen = torch.Generator(device="cuda").manual_seed(13)
    g = torch.randn((b.count(capi.COUNT_FINEST_POINTS), 3), generator=gen, device="cuda", dtype=torch.float32)

    def adj(bb, steps):
        out = torch.empty((bb.count(capi.COUNT_CAGE_POINTS), 3), device="cuda", dtype=torch.float32)
        bb.refine_points_adjoint(g.data_ptr(), out.data_ptr(), steps)
        bb.synchronize()
        return out.cpu().numpy().tobytes()

    want = {0: fresh(0), 32: fresh(32)}
    for i, steps in enumerate([0, 0, 0, 32, 32, 0]):
        assert adj(b, steps) == want[steps], (i, steps)
    b.close()


def test_batch_adjoint_against_oracle_jacobian(capi, torch):
    """The batched float adjoint of a few small cages against J^T g from the
    double oracle's basis-vector Jacobian, to 1e-5 of the gradient's scale
    (float sums of up to a few hundred terms)."""
    cases = [(C.cube(0), 3), (C.torus(6, 5, seed=2), 3), (C.irregular(2, seed=8), 2), tiny_cages()[1]]
    cages = [c for c, _ in cases]
    depths = [d for _, d in cases]
    b = capi.Batch(cages, depths)
    b.refine(capi.ALL)
    nf = b.count(capi.COUNT_FINEST_POINTS)
    rng = np.random.default_rng(9)
    g = rng.normal(size=(nf, 3)).astype(np.float32)
    gd = torch.from_numpy(g).cuda()
    out = torch.empty((b.count(capi.COUNT_CAGE_POINTS), 3), device="cuda", dtype=torch.float32)
    b.refine_points_adjoint(gd.data_ptr(), out.data_ptr(), 0)
    b.synchronize()
    out = out.cpu().numpy()
    for i, (cage, depth) in enumerate(cases):
        J = oracle_jacobian(cage, depth)
        o, n = b.offset(capi.COUNT_FINEST_POINTS, i), b.count(capi.COUNT_FINEST_POINTS, i)
        assert J.shape == (n, cage.V)
        want = J.T @ g[o:o + n].astype(np.float64)
        oc = b.offset(capi.COUNT_CAGE_POINTS, i)
        got = out[oc:oc + cage.V]
        err = float(np.abs(got - want).max()) / float(np.abs(want).max())
        assert err <= 1e-5, (i, err)
    b.close()
