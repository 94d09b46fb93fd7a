"""Halfedge cages (``cc_Mesh``) on the host: construction, ``.ccm`` I/O, generators.

A cage is the level-0 mesh the refinement path reads.  Its arrays use the
exact layouts of ``cc_Mesh`` in CatmullClark.h:17-71, so they can be handed to
the C ABI without conversion:

* ``halfedges``  int32 (H, 7): twinID, nextID, prevID, faceID, edgeID, vertexID, uvID
* ``creases``    structured (E,): nextID int32, prevID int32, sharpness float32
* ``vertex_points`` float32 (V, 3), ``uvs`` float32 (U, 2)
* ``v2h``/``e2h``/``f2h`` int32 vertex/edge/face -> halfedge maps

``from_faces`` turns an indexed polygon mesh into a cage with the conventions
of the reference's cage builder, examples/obj_to_ccm.c:
twins by matching (v1, v0) against (v0, v1) keys (:151-190), edge IDs as the
rank of max(h, twin) among halfedges with h > twin (:364-402), one halfedge
per vertex (:417-450), crease tags then boundary sharpness 16 (:265-279,
:538-561) and crease neighbours (:197-253).
"""
from __future__ import annotations

import dataclasses
import struct
from typing import Iterable, Sequence

import numpy as np

CREASE_DTYPE = np.dtype([("nextID", "<i4"), ("prevID", "<i4"), ("sharpness", "<f4")])
CCM_MAGIC = b"cc_Mesh1"


@dataclasses.dataclass
class Cage:
    vertex_points: np.ndarray  # (V, 3) float32
    uvs: np.ndarray            # (U, 2) float32
    halfedges: np.ndarray      # (H, 7) int32
    creases: np.ndarray        # (E,) CREASE_DTYPE
    v2h: np.ndarray            # (V,) int32
    e2h: np.ndarray            # (E,) int32
    f2h: np.ndarray            # (F,) int32

    @property
    def V(self) -> int:
        return int(self.vertex_points.shape[0])

    @property
    def U(self) -> int:
        return int(self.uvs.shape[0])

    @property
    def H(self) -> int:
        return int(self.halfedges.shape[0])

    @property
    def E(self) -> int:
        return int(self.creases.shape[0])

    @property
    def F(self) -> int:
        return int(self.f2h.shape[0])

    def counts(self) -> tuple[int, int, int, int, int]:
        return self.V, self.U, self.H, self.E, self.F

    def is_quad_only(self) -> bool:
        return self.H == 4 * self.F and bool(np.all(self.f2h == 4 * np.arange(self.F)))

    def copy(self) -> "Cage":
        return Cage(*(np.array(getattr(self, f.name), copy=True) for f in dataclasses.fields(self)))

    def contiguous(self) -> "Cage":
        """Return a cage whose arrays are C-contiguous with the ABI dtypes."""
        return Cage(
            np.ascontiguousarray(self.vertex_points, dtype=np.float32).reshape(-1, 3),
            np.ascontiguousarray(self.uvs, dtype=np.float32).reshape(-1, 2),
            np.ascontiguousarray(self.halfedges, dtype=np.int32).reshape(-1, 7),
            np.ascontiguousarray(self.creases.astype(CREASE_DTYPE, copy=False)),
            np.ascontiguousarray(self.v2h, dtype=np.int32),
            np.ascontiguousarray(self.e2h, dtype=np.int32),
            np.ascontiguousarray(self.f2h, dtype=np.int32),
        )


# --------------------------------------------------------------------------
# construction from an indexed polygon mesh (examples/obj_to_ccm.c conventions)
# --------------------------------------------------------------------------
def from_faces(points: np.ndarray,
               faces: Sequence[Sequence[int]],
               uvs: np.ndarray | None = None,
               face_uvs: Sequence[Sequence[int]] | None = None,
               creases: Iterable[tuple[int, int, float]] = (),
               sharpen_boundaries: bool = True) -> Cage:
    """Build a cage from polygons given as 0-based vertex index lists.

    ``creases`` holds (v0, v1, sharpness) tags, like ``t crease 2/1/0 v0 v1 s``
    lines of an OBJ file; every edge whose endpoints are {v0, v1} takes the
    sharpness (obj_to_ccm.c:538-561).  Boundary edges then get sharpness 16
    (obj_to_ccm.c:265-279) unless ``sharpen_boundaries`` is False.
    """
    points = np.asarray(points, dtype=np.float32).reshape(-1, 3)
    V = points.shape[0]
    sizes = np.fromiter((len(f) for f in faces), dtype=np.int64, count=len(faces))
    if np.any(sizes < 3):
        raise ValueError("faces need at least 3 vertices")
    F = len(faces)
    H = int(sizes.sum())
    begin = np.zeros(F + 1, dtype=np.int64)
    np.cumsum(sizes, out=begin[1:])
    vert = np.fromiter((v for f in faces for v in f), dtype=np.int64, count=H)
    if vert.size and (vert.min() < 0 or vert.max() >= V):
        raise ValueError("face vertex index out of range")
    face_of = np.repeat(np.arange(F, dtype=np.int64), sizes)
    local = np.arange(H, dtype=np.int64) - begin[face_of]
    k = sizes[face_of]
    nxt = begin[face_of] + (local + 1) % k
    prv = begin[face_of] + (local - 1) % k
    if uvs is None or face_uvs is None:
        uv_arr = np.zeros((0, 2), dtype=np.float32) if uvs is None else np.asarray(uvs, np.float32).reshape(-1, 2)
        uvid = np.zeros(H, dtype=np.int64)
    else:
        uv_arr = np.asarray(uvs, dtype=np.float32).reshape(-1, 2)
        uvid = np.fromiter((t for f in face_uvs for t in f), dtype=np.int64, count=H)

    # twins (obj_to_ccm.c:151-190): key(v0, v1) = v0 + V*v1, matched against key(v1, v0)
    v0 = vert
    v1 = vert[nxt]
    key = v0.astype(np.uint64) + np.uint64(V) * v1.astype(np.uint64)
    order = np.argsort(key, kind="stable")
    skey = key[order]
    want = v1.astype(np.uint64) + np.uint64(V) * v0.astype(np.uint64)
    pos = np.searchsorted(skey, want, side="left")
    posc = np.minimum(pos, H - 1)
    found = (pos < H) & (skey[posc] == want)
    twin = np.where(found, order[posc], -1)

    # edges (obj_to_ccm.c:364-402): one bit per halfedge with h > twin
    hid = np.arange(H, dtype=np.int64)
    bit = (hid > twin).astype(np.int64)
    rank = np.concatenate([[0], np.cumsum(bit)])
    E = int(rank[-1])
    edge = rank[np.maximum(hid, twin)]
    e2h = np.nonzero(bit)[0]

    # vertex -> halfedge (obj_to_ccm.c:417-450): boundary vertices keep their
    # outgoing boundary halfedge, interior vertices their largest halfedge
    v2h = np.full(V, -1, dtype=np.int64)
    np.maximum.at(v2h, vert, hid)
    bnd = twin < 0
    v2h[vert[bnd]] = hid[bnd]

    he = np.stack([twin, nxt, prv, face_of, edge, vert, uvid], axis=1).astype(np.int32)
    cr = np.zeros(E, dtype=CREASE_DTYPE)
    cr["nextID"] = np.arange(E)
    cr["prevID"] = np.arange(E)

    # crease tags, then boundaries
    tags = list(creases)
    if tags:
        ev0 = vert[e2h]
        ev1 = vert[nxt[e2h]]
        lo = np.minimum(ev0, ev1)
        hi = np.maximum(ev0, ev1)
        lookup = {}
        for e, (a, b) in enumerate(zip(lo.tolist(), hi.tolist())):
            lookup.setdefault((a, b), []).append(e)
        for a, b, s in tags:
            for e in lookup.get((min(a, b), max(a, b)), ()):
                cr["sharpness"][e] = np.float32(s)
    if sharpen_boundaries:
        cr["sharpness"][twin[e2h] < 0] = np.float32(16.0)

    cage = Cage(points.copy(), uv_arr.copy(), he, cr,
                v2h.astype(np.int32), e2h.astype(np.int32), begin[:-1].astype(np.int32))
    compute_crease_neighbors(cage)
    return cage


def compute_crease_neighbors(cage: Cage) -> None:
    """Link each sharp edge to the unique sharp edge continuing it across an
    interior vertex (examples/obj_to_ccm.c:197-253); others point at themselves."""
    he = cage.halfedges
    twin = he[:, 0]
    nxt = he[:, 1]
    edge = he[:, 4]
    sharp = cage.creases["sharpness"]

    def next_vertex_halfedge(h: int) -> int:
        t = int(twin[h])
        return int(nxt[t]) if t >= 0 else -1

    for e in np.nonzero(sharp > 0.0)[0].tolist():
        h = int(cage.e2h[e])
        n = int(nxt[h])
        count, cand = 0, -1
        it = next_vertex_halfedge(h)
        while it != h and it >= 0:
            if sharp[edge[it]] > 0.0:
                cand = int(edge[it])
                count += 1
            it = next_vertex_halfedge(it)
        if count == 1 and it == h:
            cage.creases["prevID"][e] = cand
        count, cand = 0, -1
        if sharp[edge[n]] > 0.0:
            cand = int(edge[n])
            count += 1
        it = next_vertex_halfedge(n)
        while it != n and it >= 0:
            if sharp[edge[it]] > 0.0 and int(twin[it]) != h:
                cand = int(edge[it])
                count += 1
            it = next_vertex_halfedge(it)
        if count == 1 and it == n:
            cage.creases["nextID"][e] = cand


# --------------------------------------------------------------------------
# .ccm files (CatmullClark.h:3085-3269)
# --------------------------------------------------------------------------
def save_ccm(cage: Cage, path: str) -> None:
    """Write the binary cage format: 8-byte magic, 5 int32 counts, 4 padding
    bytes, then v2h, e2h, f2h, points, uvs, creases, halfedges."""
    c = cage.contiguous()
    with open(path, "wb") as f:
        f.write(CCM_MAGIC + struct.pack("<5i", c.V, c.U, c.H, c.E, c.F) + b"\0" * 4)
        for arr in (c.v2h, c.e2h, c.f2h, c.vertex_points, c.uvs, c.creases, c.halfedges):
            f.write(arr.tobytes())


def load_ccm(path: str) -> Cage:
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < 32 or data[:8] != CCM_MAGIC:
        raise ValueError(f"{path}: not a cc_Mesh file")
    V, U, H, E, F = struct.unpack_from("<5i", data, 8)
    off = 32
    out = []
    for dtype, n, shape in ((np.int32, V, (V,)), (np.int32, E, (E,)), (np.int32, F, (F,)),
                            (np.float32, 3 * V, (V, 3)), (np.float32, 2 * U, (U, 2)),
                            (CREASE_DTYPE, E, (E,)), (np.int32, 7 * H, (H, 7))):
        nbytes = np.dtype(dtype).itemsize * n
        if off + nbytes > len(data):
            raise ValueError(f"{path}: truncated")
        out.append(np.frombuffer(data, dtype=dtype, count=n, offset=off).reshape(shape).copy())
        off += nbytes
    v2h, e2h, f2h, pts, uvs, cr, he = out
    return Cage(pts, uvs, he, cr, v2h, e2h, f2h)


def to_obj(points, faces, uvs=None, face_uvs=None, creases=()) -> str:
    """Indexed mesh as OBJ text with ``t crease 2/1/0`` tags (1-based)."""
    lines = [f"v {p[0]:.9g} {p[1]:.9g} {p[2]:.9g}" for p in np.asarray(points, np.float64)]
    if uvs is not None:
        lines += [f"vt {t[0]:.9g} {t[1]:.9g}" for t in np.asarray(uvs, np.float64)]
    for i, f in enumerate(faces):
        if uvs is not None and face_uvs is not None:
            lines.append("f " + " ".join(f"{v + 1}/{t + 1}" for v, t in zip(f, face_uvs[i])))
        else:
            lines.append("f " + " ".join(str(v + 1) for v in f))
    for a, b, s in creases:
        # nine significant digits of the float32 value read back exactly
        lines.append(f"t crease 2/1/0 {a} {b} {float(np.float32(s)):.9g}")
    return "\n".join(lines) + "\n"


def cage_polygons(cage: Cage):
    """The polygons of a cage (from f2h by next), their per-corner uv IDs and
    one (v0, v1, sharpness) tag per sharp edge: the arguments of from_faces
    and to_obj that describe the cage again."""
    he = cage.halfedges
    faces, fuv = [], []
    for f in range(cage.F):
        h0 = int(cage.f2h[f])
        vs, ts, h = [], [], h0
        while True:
            vs.append(int(he[h, 5]))
            ts.append(int(he[h, 6]))
            h = int(he[h, 1])
            if h == h0:
                break
        faces.append(vs)
        fuv.append(ts)
    sharp = np.nonzero(cage.creases["sharpness"] > 0)[0]
    creases = [(int(he[cage.e2h[e], 5]), int(he[he[cage.e2h[e], 1], 5]), float(cage.creases["sharpness"][e]))
               for e in sharp.tolist()]
    return faces, fuv, creases


# --------------------------------------------------------------------------
# seeded generators for the benchmark and test workloads
# --------------------------------------------------------------------------
def cube_mesh(seed: int = 0, jitter: float = 0.05, n_creases: int = 4, max_sharpness: float = 3.0):
    """Cube: corners at +-1 with N(0, jitter) noise, 6 quads, unit-square
    per-corner UVs, ``n_creases`` random edges with sharpness U[0, max]."""
    rng = np.random.default_rng(seed)
    corners = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64)
    pts = corners + rng.normal(0.0, jitter, size=corners.shape)
    # outward-oriented quads over the corner index 4x+2y+z
    faces = [[0, 1, 3, 2], [4, 6, 7, 5], [0, 4, 5, 1], [2, 3, 7, 6], [0, 2, 6, 4], [1, 5, 7, 3]]
    uvs = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    face_uvs = [[0, 1, 2, 3]] * 6
    edges = sorted({tuple(sorted((f[i], f[(i + 1) % 4]))) for f in faces for i in range(4)})
    pick = rng.choice(len(edges), size=min(n_creases, len(edges)), replace=False)
    creases = [(edges[i][0], edges[i][1], float(rng.uniform(0.0, max_sharpness))) for i in pick]
    return pts.astype(np.float32), faces, uvs, face_uvs, creases


def torus_mesh(nu: int = 64, nv: int = 32, R: float = 1.0, r: float = 0.3, seed: int = 0,
               jitter: float = 0.01, ring_sharpness: Sequence[tuple[int, float]] = ((0, 1.5), (None, 4.0))):
    """Torus quad grid of nu x nv faces (no boundary) with grid UVs.

    ``ring_sharpness`` lists (j, s): the longitudinal edge ring at minor index
    j (all nu edges around the major circle) gets sharpness s; j=None means
    the ring opposite the first one (nv // 2)."""
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    u = 2 * np.pi * i / nu
    v = 2 * np.pi * j / nv
    pts = np.stack([(R + r * np.cos(v)) * np.cos(u), (R + r * np.cos(v)) * np.sin(u), r * np.sin(v)], -1)
    pts = pts.reshape(-1, 3) + rng.normal(0.0, jitter, size=(nu * nv, 3))
    vid = lambda a, b: (a % nu) * nv + (b % nv)
    uid = lambda a, b: a * (nv + 1) + b
    faces, face_uvs = [], []
    for a in range(nu):
        for b in range(nv):
            faces.append([vid(a, b), vid(a + 1, b), vid(a + 1, b + 1), vid(a, b + 1)])
            face_uvs.append([uid(a, b), uid(a + 1, b), uid(a + 1, b + 1), uid(a, b + 1)])
    ua, ub = np.meshgrid(np.arange(nu + 1), np.arange(nv + 1), indexing="ij")
    uvs = np.stack([ua / nu, ub / nv], -1).reshape(-1, 2).astype(np.float32)
    creases = []
    for jj, s in ring_sharpness:
        jr = nv // 2 if jj is None else jj
        creases += [(vid(a, jr), vid(a + 1, jr), s) for a in range(nu)]
    return pts.astype(np.float32), faces, uvs, face_uvs, creases


def _icosphere(level: int):
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
             (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in verts]
    tris = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
            (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
            (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]

        new = []
        for a, b, c in tris:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        tris = new
    return np.array(verts), [list(t) for t in tris]


def irregular_mesh(level: int = 4, seed: int = 0, merge_p: float = 0.5, cap_angle: float = 0.4,
                   jitter: float = 0.005, crease_frac: float = 0.1, max_sharpness: float = 5.0):
    """Icosphere of ``level`` with seeded triangle-pair -> quad merges
    (probability ``merge_p``), a deleted spherical cap of ``cap_angle`` radians
    (a boundary loop) and ``crease_frac`` of the edges creased U[0, max]."""
    rng = np.random.default_rng(seed)
    pts, tris = _icosphere(level)
    # delete the faces whose centroid lies inside the cap around a random axis
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    keep = [t for t in tris if np.dot(pts[t].mean(0) / np.linalg.norm(pts[t].mean(0)), axis) < np.cos(cap_angle)]
    # merge random adjacent triangle pairs into quads
    edge_faces = {}
    for fi, t in enumerate(keep):
        for k in range(3):
            edge_faces.setdefault((min(t[k], t[(k + 1) % 3]), max(t[k], t[(k + 1) % 3])), []).append(fi)
    used = np.zeros(len(keep), bool)
    faces = []
    keys = list(edge_faces.keys())
    for ei in rng.permutation(len(keys)):
        fs = edge_faces[keys[ei]]
        if len(fs) != 2 or used[fs[0]] or used[fs[1]] or rng.random() >= merge_p:
            continue
        a, b = keys[ei]
        t0, t1 = keep[fs[0]], keep[fs[1]]
        # t0 contains a->b or b->a; build the quad by inserting t1's apex
        apex = [v for v in t1 if v not in (a, b)][0]
        # t0 = (x, p, q) with p->q the shared edge; t1 runs q->p->apex, so the
        # merged quad keeps the orientation as (x, p, apex, q)
        x = [v for v in t0 if v not in (a, b)][0]
        kx = t0.index(x)
        seq = [t0[kx], t0[(kx + 1) % 3], t0[(kx + 2) % 3]]
        quad = [seq[0], seq[1], apex, seq[2]]
        used[fs[0]] = used[fs[1]] = True
        faces.append(quad)
    faces += [list(t) for fi, t in enumerate(keep) if not used[fi]]
    # compact vertices
    usedv = sorted({v for f in faces for v in f})
    remap = {v: i for i, v in enumerate(usedv)}
    faces = [[remap[v] for v in f] for f in faces]
    P = pts[usedv] + rng.normal(0.0, jitter, size=(len(usedv), 3))
    edges = sorted({(min(f[k], f[(k + 1) % len(f)]), max(f[k], f[(k + 1) % len(f)])) for f in faces for k in range(len(f))})
    n_cr = int(round(crease_frac * len(edges)))
    pick = rng.choice(len(edges), size=n_cr, replace=False)
    creases = [(edges[i][0], edges[i][1], float(rng.uniform(0.0, max_sharpness))) for i in pick]
    # planar UVs from the direction
    d = P / np.linalg.norm(P, axis=1, keepdims=True)
    uvs = np.stack([0.5 + np.arctan2(d[:, 1], d[:, 0]) / (2 * np.pi), 0.5 + np.arcsin(np.clip(d[:, 2], -1, 1)) / np.pi], -1)
    uvs = np.clip(uvs, 0.0, 1.0).astype(np.float32)
    return P.astype(np.float32), faces, uvs, faces, creases


def cube(seed: int = 0) -> Cage:
    return from_faces(*_split(cube_mesh(seed)))


def torus(nu: int = 64, nv: int = 32, seed: int = 0, **kw) -> Cage:
    return from_faces(*_split(torus_mesh(nu, nv, seed=seed, **kw)))


def irregular(level: int = 4, seed: int = 0, **kw) -> Cage:
    return from_faces(*_split(irregular_mesh(level, seed=seed, **kw)))


def _split(mesh):
    pts, faces, uvs, face_uvs, creases = mesh
    return pts, faces, uvs, face_uvs, creases


def check_manifold(cage: Cage) -> None:
    """Raise if a vertex's halfedges do not form one fan (a cycle, or a chain
    that starts at the vertex's boundary halfedge)."""
    he = cage.halfedges
    twin, prv, vert = he[:, 0], he[:, 2], he[:, 5]
    per_vertex = np.bincount(vert, minlength=cage.V)
    for v in range(cage.V):
        h = int(cage.v2h[v])
        if h < 0 or vert[h] != v:
            raise ValueError(f"vertex {v}: bad vertex->halfedge map")
        n, it = 1, int(twin[prv[h]])
        while it >= 0 and it != h:
            n += 1
            it = int(twin[prv[it]])
            if n > per_vertex[v]:
                break
        if n != per_vertex[v]:
            raise ValueError(f"vertex {v}: non-manifold fan ({n} of {per_vertex[v]} halfedges reached)")


def batch_cages(count: int = 256, seed: int = 0, h_min: int = 24, h_max: int = 60000,
                depths: Sequence[int] = (4, 5, 6), crease_frac: float = 0.05, max_sharpness: float = 4.0):
    """Cages of uneven size for batched refinement: halfedge counts
    log-uniform in [h_min, h_max] drawn from the cube / torus / irregular
    generators, depths uniform in ``depths``, ``crease_frac`` of the edges
    creased U[0, max_sharpness].  Returns (cages, depths)."""
    rng = np.random.default_rng(seed)
    cages, ds = [], []
    for k in range(count):
        target = float(np.exp(rng.uniform(np.log(h_min), np.log(h_max))))
        kind = rng.integers(0, 3)
        s = int(rng.integers(0, 2 ** 31 - 1))
        if target < 60 or kind == 0:
            pts, faces, uvs, fuv, _ = cube_mesh(s, n_creases=0)
            if target >= 60:  # a torus stands in when a cube is far too small
                n = max(2, int(round((target / 4) ** 0.5)))
                pts, faces, uvs, fuv, _ = torus_mesh(max(3, n), max(3, n), seed=s, ring_sharpness=())
        elif kind == 1:
            n = max(3, int(round((target / 8) ** 0.5)))
            pts, faces, uvs, fuv, _ = torus_mesh(2 * n, n, seed=s, ring_sharpness=())
        else:
            level = int(np.clip(np.round(np.log(max(target, 60) / 45.0) / np.log(4.0)), 0, 5))
            pts, faces, uvs, fuv, _ = irregular_mesh(level, seed=s, crease_frac=0.0)
        edges = sorted({(min(f[i], f[(i + 1) % len(f)]), max(f[i], f[(i + 1) % len(f)]))
                        for f in faces for i in range(len(f))})
        n_cr = int(round(crease_frac * len(edges)))
        pick = rng.choice(len(edges), size=n_cr, replace=False) if n_cr else []
        creases = [(edges[i][0], edges[i][1], float(rng.uniform(0.0, max_sharpness))) for i in pick]
        cages.append(from_faces(pts, faces, uvs, fuv, creases))
        ds.append(int(rng.choice(list(depths))))
    return cages, ds
