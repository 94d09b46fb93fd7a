"""GPU parity at the workloads' full sizes, bit for bit against the oracle.

- torus 64x32 (8,192 halfedges, two crease rings of sharpness 1.5 and 4.0,
  grid UVs) refined to depth 8: 716 M halfedges over levels 1..8, 537 M of
  them at the finest level;
- the irregular cage (icosphere level 4, seeded triangle pairs merged into
  quads, a cap deleted to open a boundary loop, 10% of the edges creased)
  refined to depth 7: 247 M halfedges;
- the batch of 256 uneven cages (24..60,000 halfedges, depths 4..6, 5.26 G
  halfedges in one arena, so the later cages sit past 2^32 halfedges):
  every cage against the oracle's SHA-256 digests (about 100 GB of output
  in all, too much for the oracle to hold next to it, so the digests are
  committed), and the batched adjoint of every eighth of the cages of at
  most 2 M refined halfedges and of the last cage against the single-cage
  adjoint.

The oracle runs its gather loops on all host threads (up to 16); its scatter
sums stay sequential, so its points equal the sequential reference's.  Sizes
like these are where 32-bit index and byte-offset mistakes show, which the
small cages of the other parity tests cannot reach.
"""
import concurrent.futures
import hashlib
import json
import os

import numpy as np
import pytest

from halfedgecatmullclark_amd import cage as C
from halfedgecatmullclark_amd.abi import level_counts
from oracle import oracle as O
from tests.helpers import assert_bits_equal

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
HERE = os.path.dirname(os.path.abspath(__file__))

CASES = [
    ("torus64x32_d8_scatter", lambda: C.torus(64, 32), 8, 15, O.ALL),
    ("torus64x32_d8_gather", lambda: C.torus(64, 32), 8, 15 | 16, O.ALL | O.GATHER),
    ("irregular4_d7_scatter", lambda: C.irregular(4), 7, 15, O.ALL),
    ("irregular4_d7_nocreases_gather", lambda: C.irregular(4), 7, 15 | 16 | 32, O.ALL | O.GATHER | O.NOCREASES),
]


@pytest.fixture(scope="module")
def capi():
    from halfedgecatmullclark_amd import capi as api
    api.lib()
    return api


@pytest.mark.parametrize("label,make,depth,steps,flags", CASES, ids=[c[0] for c in CASES])
def test_full_size_bitexact(capi, label, make, depth, steps, flags):
    """The full refinement, then a points-only pass over the same topology
    (the owner bits of the first pass kept, each halfedge's side of its edge
    through LDS), both against the oracle."""
    import torch
    cage = make()
    ds = capi.DeviceSubd(cage, depth)
    ds.refine(steps)
    got = ds.download()
    out = torch.empty((ds.n["Vcum"][depth], 3), device="cuda", dtype=torch.float32)
    ds.refine_points(None, out.data_ptr(), steps & (16 | 32))   # the gather and no-crease bits
    ds.synchronize()
    points_only = out.cpu().numpy()
    del out
    ds.close()
    want = O.refine(cage, depth, flags, threads=THREADS)
    assert_bits_equal(points_only, want["points"], f"{label} points-only pass")
    del points_only
    for k in ("halfedges", "creases", "points"):
        assert_bits_equal(got[k], want[k], f"{label} {k}")
        got[k] = want[k] = None


def _digest(a) -> str:
    return hashlib.sha256(memoryview(np.ascontiguousarray(a).reshape(-1).view("u1"))).hexdigest()


def test_batch_workload_bitexact(capi):
    """Every one of the 256 cages (5.26 G halfedges in one arena, the later
    cages past 2^32 halfedges) against the oracle's SHA-256 digests, written
    by tests/golden/make_batch_digests.py."""
    gold = json.load(open(os.path.join(HERE, "golden", "batch256_digests.json")))["cages"]
    cages, depths = C.batch_cages(256, seed=0)
    assert len(gold) == len(cages) == 256
    b = capi.Batch(cages, depths)
    b.refine(capi.ALL)
    b.synchronize()
    hcum = [level_counts(c.V, c.H, c.E, c.F, d)["Hcum"][d] for c, d in zip(cages, depths)]
    assert sum(hcum[:-1]) > 2 ** 32
    bad = []

    def check(i, got):
        row = gold[i]
        assert (row["H"], row["depth"]) == (cages[i].H, depths[i])
        for k in ("halfedges", "creases", "points"):
            if _digest(got[k]) != row[k]:
                bad.append(f"cage {i} (H={cages[i].H}, depth {depths[i]}) {k}")

    # downloads in order, digests on a few threads (hashlib releases the GIL);
    # at most 4 cages (up to 6 GB each) wait in host memory
    with concurrent.futures.ThreadPoolExecutor(4) as pool:
        pending = []
        for i in range(len(cages)):
            pending.append(pool.submit(check, i, b.download(i)))
            if len(pending) >= 4:
                pending.pop(0).result()
        for f in pending:
            f.result()
    b.close()
    assert not bad, bad[:10]


def test_batch_workload_adjoint_matches_single_cage(capi):
    """The batched adjoint of the 256-cage workload, checked bit for bit
    against the single-cage adjoint for the cages of at most 2 M refined
    halfedges and the last cage (past 2^32 halfedges in the arena): both add
    in the same fixed order, with the same kept edge-representative bits."""
    import torch
    cages, depths = C.batch_cages(256, seed=0)
    b = capi.Batch(cages, depths)
    b.refine(capi.ALL)
    nf = b.count(capi.COUNT_FINEST_POINTS)
    gen = torch.Generator(device="cuda").manual_seed(11)
    g = torch.randn((nf, 3), generator=gen, device="cuda", dtype=torch.float32)
    out = torch.empty((b.count(capi.COUNT_CAGE_POINTS), 3), device="cuda", dtype=torch.float32)
    b.refine_points_adjoint(g.data_ptr(), out.data_ptr(), 0)
    b.synchronize()
    hcum = [level_counts(c.V, c.H, c.E, c.F, d)["Hcum"][d] for c, d in zip(cages, depths)]
    pick = [i for i, h in enumerate(hcum) if h <= 2_000_000][::8] + [len(cages) - 1]
    offs = {i: (b.offset(capi.COUNT_FINEST_POINTS, i), b.count(capi.COUNT_FINEST_POINTS, i),
                b.offset(capi.COUNT_CAGE_POINTS, i)) for i in pick}
    got_all = out.cpu().numpy()
    b.close()
    del out
    for i in pick:
        o, n, oc = offs[i]
        ds = capi.DeviceSubd(cages[i], depths[i])
        ds.refine(capi.ALL)
        gi = g[o:o + n].contiguous()
        oi = torch.empty((cages[i].V, 3), device="cuda", dtype=torch.float32)
        ds.refine_points_adjoint(gi.data_ptr(), oi.data_ptr(), 0)
        ds.synchronize()
        assert_bits_equal(np.ascontiguousarray(got_all[oc:oc + cages[i].V]), oi.cpu().numpy(),
                          f"cage {i} (H={cages[i].H}, depth {depths[i]}) adjoint")
        ds.close()
