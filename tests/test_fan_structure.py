"""The structural vertex rings used by the GPU vertex-point kernel.

The kernel never walks twin pointers to find the halfedges around a vertex:
it derives them from the level at which the vertex appeared (children 0 keep
their parent's vertex; face and edge vertices have fixed child patterns,
CatmullClark.h:2770-2793).  This restates that derivation in Python and
checks, on oracle-refined topology, that for every vertex exactly one
halfedge claims to be its representative (the reference's
ccs_VertexPointToHalfedgeID, CatmullClark.h:1336), that the claimed fan is
exactly the ring found by walking twins, in forward order, and that every
fan halfedge is in range.
"""
import numpy as np
import pytest

from halfedgecatmullclark_amd import cage as C
from oracle import oracle as O
from tests.helpers import small_cages


def qnext(h):
    return (h & ~3) | ((h + 1) & 3)


def qprev(h):
    return (h & ~3) | ((h - 1) & 3)


def owner_fan(h, vert, d, cage, counts, levels):
    """Mirror of owner_fan() in cc_fused.hip; returns None or (fan, ri, boundary)."""
    he = cage.halfedges
    k = 0
    while k < d - 1 and ((h >> (2 * k)) & 3) == 0:
        k += 1
    q = h >> (2 * k)
    lvl = d - k
    digit = q & 3
    p = q >> 2
    shift = 2 * k
    if digit == 0:
        if cage.v2h[vert] != p:
            return None
        ring, g = [p], he[he[p, 2], 0]
        while g >= 0 and g != p:
            ring.append(int(g))
            g = he[he[g, 2], 0]
        return [(4 * x) << shift for x in ring], 0, bool(g < 0)
    if digit == 3:
        return None
    if digit == 2:
        if lvl - 1 >= 1:
            if p & 3:
                return None
            return [((4 * p + 2) + 4 * i) << shift for i in range(4)], 0, False
        face = he[p, 3]
        if cage.f2h[face] != p:
            return None
        ring, it = [p], he[p, 1]
        while it != p:
            ring.append(int(it))
            it = he[it, 1]
        return [(4 * x + 2) << shift for x in ring], 0, False
    e = vert - counts["V"][lvl - 1] - counts["F"][lvl - 1]
    if lvl - 1 >= 1:
        if O.edge_to_halfedge(cage, int(e), lvl - 1) != p:
            return None
        tw = int(levels[lvl - 1]["twinID"][p])
        np_, nt = qnext(p), qnext(tw)
    else:
        if cage.e2h[e] != p:
            return None
        tw = int(he[p, 0])
        np_ = int(he[p, 1])
        nt = int(he[tw, 1]) if tw >= 0 else -1
    if tw >= 0:
        return [x << shift for x in (4 * p + 1, 4 * nt + 3, 4 * tw + 1, 4 * np_ + 3)], 0, False
    return [x << shift for x in (4 * np_ + 3, 4 * p + 1)], 1, True


def walked_fan(first_candidates, lv):
    """Ring of the vertex by walking twins: from the halfedge whose twin is
    negative (boundary) or from the given start, forward via twin(prev(x))."""
    twin = lv["twinID"]
    h = first_candidates
    it, start = h, h
    # back up to the boundary, if any
    for _ in range(10000):
        t = twin[it]
        if t < 0:
            start = it
            break
        nx = qnext(int(t))
        if nx == h:
            start = h
            break
        it = nx
    ring, x = [start], start
    for _ in range(10000):
        nxt = int(twin[qprev(x)])
        if nxt < 0 or nxt == start:
            return ring, nxt < 0
        ring.append(nxt)
        x = nxt


@pytest.mark.parametrize("name,cage,depth", small_cages()[:6], ids=[c[0] for c in small_cages()[:6]])
def test_structural_fans(name, cage, depth):
    depth = min(depth, 4)
    out = O.refine(cage, depth, O.HALFEDGES)
    counts = out["counts"]
    levels = {d: O.level(out, "halfedges", d) for d in range(1, depth + 1)}
    for d in range(1, depth + 1):
        lv = levels[d]
        H = counts["H"][d]
        owners = {}
        for h in range(H):
            v = int(lv["vertexID"][h])
            r = owner_fan(h, v, d, cage, counts, levels)
            if r is None:
                continue
            assert v not in owners, f"level {d}: vertex {v} has two owners"
            owners[v] = h
            fan, ri, boundary = r
            assert fan[ri] == h
            assert all(0 <= x < H for x in fan), (d, h, fan)
            ring, is_boundary = walked_fan(h, lv)
            assert boundary == is_boundary, (d, h)
            if boundary:
                assert fan == ring, (d, h, fan, ring)
            else:  # same cycle, starting at the representative
                i = ring.index(h)
                assert fan == ring[i:] + ring[:i], (d, h, fan, ring)
            assert all(int(lv["vertexID"][x]) == v for x in fan)
            assert O.vertex_to_halfedge(cage, v, d) == h
        assert len(owners) == counts["V"][d], (d, len(owners), counts["V"][d])


def birth(h, d):
    """(k, lvl, digit, p): h = ((4 p + digit) << 2k), the vertex born at lvl = d - k
    (k from the count of trailing zero base-4 digits, capped at d - 1)."""
    k = min(((h & -h).bit_length() - 1) >> 1, d - 1) if h else d - 1
    q = h >> (2 * k)
    return k, d - k, q & 3, q >> 2


def closed_ring(h, d, cage, counts, levels):
    """Mirror of marked_ring_fan + structural_ring in cc_fan.h: the ring
    (halfedges and their edges) of the vertex represented by level-d
    halfedge h, for face vertices born at level >= 2, face vertices of cage
    quads whose halfedges are p .. p+3 in next order, and interior edge
    vertices, and cage vertices with a closed ring of four, with the child-0
    descent d - lvl levels down in closed form; None for any other vertex."""
    k, lvl, digit, p = birth(h, d)
    Ep2 = 2 * counts["E"][lvl - 1]
    if digit == 0:
        # a cage vertex: its closed ring of four (forward walk from p), each
        # R with a twin whose next is not R itself
        he = cage.halfedges
        ring, g = [p], int(he[he[p, 2], 0])
        while g >= 0 and g != p and len(ring) <= 4:
            ring.append(g)
            g = int(he[he[g, 2], 0])
        if g != p or len(ring) != 4:
            return None
        tw0 = [int(he[R, 0]) for R in ring]
        nts = [int(he[w, 1]) for w in tw0]
        if any(nt == R for nt, R in zip(nts, ring)):
            return None
        v = int(he[p, 5])
        hs = [4 * R for R in ring]
        es = [2 * int(he[R, 4]) + (0 if R > w else 1) for R, w in zip(ring, tw0)]
        tws = [4 * nt + 3 for nt in nts]
    elif digit == 2 and lvl == 1:
        he = cage.halfedges
        if p + 3 >= len(he) or [int(he[p + q, 1]) for q in range(4)] != [p + 1, p + 2, p + 3, p]:
            return None
        v = counts["V"][0] + int(he[p, 3])
        hs = [4 * (p + q) + 2 for q in range(4)]
        es = [Ep2 + p + ((q + 3) & 3) for q in range(4)]
        tws = [4 * (p + ((q + 3) & 3)) + 1 for q in range(4)]
    elif digit == 2 and lvl >= 2:
        v = counts["V"][lvl - 1] + (p >> 2)
        hs = [4 * (p + q) + 2 for q in range(4)]
        es = [Ep2 + p + ((q + 3) & 3) for q in range(4)]
        tws = [4 * (p + ((q + 3) & 3)) + 1 for q in range(4)]
    elif digit == 1:
        if lvl >= 2:
            rec = levels[lvl - 1][p]
            tw, e = int(rec["twinID"]), int(rec["edgeID"])
            np_, nw = qnext(p), (qnext(tw) if tw >= 0 else -1)
        else:
            he = cage.halfedges
            tw, e, np_ = int(he[p, 0]), int(he[p, 4]), int(he[p, 1])
            nw = int(he[tw, 1]) if tw >= 0 else -1
        if tw < 0:
            return None
        v = counts["V"][lvl - 1] + counts["F"][lvl - 1] + e
        hs = [4 * p + 1, 4 * nw + 3, 4 * tw + 1, 4 * np_ + 3]
        es = [Ep2 + p, 2 * e + (1 if tw > p else 0), Ep2 + tw, 2 * e + (1 if p > tw else 0)]
        tws = [4 * np_ + 2, 4 * p, 4 * nw + 2, 4 * tw]
    else:
        return None
    n = d - lvl
    # h and its twin lie in different quads, so children 0 of both keep their
    # order at every level below: the edge bit is the same n times
    es = [(e << n) + (((1 << n) - 1) if x < t else 0) for e, x, t in zip(es, hs, tws)]
    hs = [x << (2 * n) for x in hs]
    return v, hs, es


@pytest.mark.parametrize("name,cage,depth", small_cages()[:6], ids=[c[0] for c in small_cages()[:6]])
def test_closed_form_rings(name, cage, depth):
    """Every ring the closed form gives is the vertex's ring in the oracle's
    topology: each halfedge carries the vertex, its edge is e[q], and the
    edge of its previous halfedge is e[q + 1]."""
    depth = min(depth, 4)
    out = O.refine(cage, depth, O.HALFEDGES)
    counts = out["counts"]
    levels = {d: O.level(out, "halfedges", d) for d in range(1, depth + 1)}
    cage_quads = cage_verts = 0
    for d in range(1, depth + 1):
        lv = levels[d]
        seen = 0
        for h in range(counts["H"][d]):
            v = int(lv["vertexID"][h])
            if O.vertex_to_halfedge(cage, v, d) != h:
                continue
            r = closed_ring(h, d, cage, counts, levels)
            if r is None:
                continue
            cv, hs, es = r
            assert cv == v and hs[0] == h, (d, h, cv, v, hs)
            for q in range(4):
                x = hs[q]
                assert int(lv["vertexID"][x]) == v, (d, h, q)
                assert int(lv["edgeID"][x]) == es[q], (d, h, q, int(lv["edgeID"][x]), es[q])
                assert int(lv["edgeID"][qprev(x)]) == es[(q + 1) & 3], (d, h, q)
            seen += 1
            cage_quads += 1 if counts["V"][0] <= v < counts["V"][0] + counts["F"][0] else 0
            cage_verts += 1 if v < counts["V"][0] else 0
        assert seen > 0
    # every cage here but the irregular ones is built from quads with
    # consecutive halfedges: their face vertices take the structural ring;
    # every torus vertex has a closed ring of four
    if not name.startswith("irregular"):
        assert cage_quads == depth * counts["F"][0], (cage_quads, counts["F"][0])
    if name.startswith("torus"):
        assert cage_verts == depth * counts["V"][0], (cage_verts, counts["V"][0])


def edge_rep_bits(cage, counts, levels, depth):
    """Mirror of the edge-representative bits the fused kernel carries from
    level to level: R_L[x] is true when level-L halfedge x is the
    reference's representative of its edge (ccs_EdgeToHalfedgeID,
    CatmullClark.h:1206-1262).  Level 1 from the cage (e2h, next, prev,
    twin); level L + 1 from level L: child 0 of p when p > twin(p) and
    R_L[p]; child 1 when p is the quad's last halfedge; child 2 unless it is
    the first; child 3 when prev(p) > twin(prev(p)) and R_L[prev(p)]."""
    he = cage.halfedges
    twin_c, next_c, prev_c, edge_c = he[:, 0], he[:, 1], he[:, 2], he[:, 4]
    p = np.arange(cage.H)
    r0 = cage.e2h[edge_c] == p
    R = {1: np.zeros(4 * cage.H, bool)}
    R[1][0::4] = (p > twin_c) & r0
    R[1][1::4] = p > next_c
    R[1][2::4] = p > prev_c
    R[1][3::4] = (prev_c > twin_c[prev_c]) & r0[prev_c]
    for L in range(1, depth):
        H = counts["H"][L]
        p = np.arange(H)
        tw = levels[L]["twinID"]
        pp = (p & ~3) | ((p - 1) & 3)
        nxt = np.zeros(4 * H, bool)
        nxt[0::4] = (p > tw) & R[L]
        nxt[1::4] = (p & 3) == 3
        nxt[2::4] = (p & 3) != 0
        nxt[3::4] = (pp > tw[pp]) & R[L][pp]
        R[L + 1] = nxt
    return R


@pytest.mark.parametrize("name,cage,depth", small_cages(), ids=[c[0] for c in small_cages()])
def test_edge_rep_bits(name, cage, depth):
    """The carried bits mark exactly the reference's edge representatives
    at every level (one per edge, and it is a halfedge of that edge)."""
    depth = min(depth, 4)
    out = O.refine(cage, depth, O.HALFEDGES)
    counts = out["counts"]
    levels = {d: O.level(out, "halfedges", d) for d in range(1, depth + 1)}
    R = edge_rep_bits(cage, counts, levels, depth)
    for L in range(1, depth + 1):
        marked = np.nonzero(R[L])[0]
        edges = levels[L]["edgeID"][marked]
        assert marked.size == counts["E"][L], (L, marked.size, counts["E"][L])
        want = np.array([O.edge_to_halfedge(cage, int(e), L) for e in edges])
        assert np.array_equal(marked, want), (L, marked[marked != want][:5])
