"""Independent double-precision Catmull-Clark step on face lists (test only).

Textbook rules (Catmull & Clark 1978; sharp rules of DeRose et al. 1998 for
sharpness >= 1): face point = mean of the face's vertices; smooth edge point
= (a + b + F0 + F1) / 4, sharp or boundary edge point = (a + b) / 2; smooth
vertex = (Q + 2R + (n - 3) V) / n with Q the mean of the adjacent face
points and R the mean of the incident edge midpoints (evaluated through the
new edge points, see below); a vertex with exactly
two sharp edges moves to (A + 6V + B) / 8; with more (or on an unsharpened
boundary) it stays.  Nothing here follows the reference's code: it is the
mathematics the reference's rules must agree with.
"""
from __future__ import annotations

import numpy as np


def cc_step(points, faces, sharp=None, pin_boundary=True):
    """One step.  ``sharp`` maps an undirected vertex pair to True when the
    edge is sharp (sharpness >= 1).  Returns dicts keyed by vertex index,
    face index and undirected edge."""
    P = np.asarray(points, np.float64)
    sharp = sharp or {}
    fpts = {fi: P[list(f)].mean(axis=0) for fi, f in enumerate(faces)}
    edge_faces = {}
    for fi, f in enumerate(faces):
        for k in range(len(f)):
            e = tuple(sorted((f[k], f[(k + 1) % len(f)])))
            edge_faces.setdefault(e, []).append(fi)
    epts = {}
    for e, fs in edge_faces.items():
        a, b = P[e[0]], P[e[1]]
        if len(fs) == 1 or sharp.get(e, False):
            epts[e] = (a + b) / 2.0
        else:
            epts[e] = (a + b + fpts[fs[0]] + fpts[fs[1]]) / 4.0
    vfaces, vedges = {}, {}
    for fi, f in enumerate(faces):
        for v in f:
            vfaces.setdefault(v, []).append(fi)
    for e in edge_faces:
        for v in e:
            vedges.setdefault(v, []).append(e)
    vpts = {}
    for v in range(len(P)):
        es = vedges.get(v, [])
        boundary = any(len(edge_faces[e]) == 1 for e in es)
        sharp_es = [e for e in es if sharp.get(e, False) or len(edge_faces[e]) == 1]
        n = len(es)
        if len(sharp_es) == 2 and (sharp or not pin_boundary):
            others = [e[0] if e[1] == v else e[1] for e in sharp_es]
            vpts[v] = (P[others[0]] + 6.0 * P[v] + P[others[1]]) / 8.0
        elif len(sharp_es) >= 2 or (boundary and pin_boundary):
            vpts[v] = P[v].copy()
        else:
            # (Q + 2R + (n-3)V)/n, written with the new edge points:
            # 2R = 4 mean(E') - Q - V on smooth edges.  At a dart (one sharp
            # edge) the sharp edge's new point is its midpoint, which is what
            # the smooth rule of CatmullClark.h (:2389-2400) uses as well.
            Q = np.mean([fpts[fi] for fi in vfaces[v]], axis=0)
            Em = np.mean([epts[e] for e in es], axis=0)
            vpts[v] = (4.0 * Em - Q + (n - 3.0) * P[v]) / n
    return vpts, fpts, epts


def level_faces(halfedges_level, H):
    """Quads of a refined level from its halfedge records (4f..4f+3)."""
    vert = halfedges_level["vertexID"].reshape(-1, 4)
    return [list(map(int, q)) for q in vert]


def level_edge_ids(halfedges_level):
    """Undirected vertex pair -> edge ID of a refined level."""
    vert = halfedges_level["vertexID"]
    out = {}
    H = len(vert)
    for h in range(H):
        nxt = (h & ~3) | ((h + 1) & 3)
        out[tuple(sorted((int(vert[h]), int(vert[nxt]))))] = int(halfedges_level["edgeID"][h])
    return out
