// Level-0 (cage) rules and crease refinement as per-element device
// functions, shared by the per-cage kernels and the batched kernels.  Each
// cites the CatmullClark.h rule it implements; the arithmetic keeps the
// reference's operation order (built with -ffp-contract=off).
#pragma once

#include "cc_common.h"

namespace cch {

// Children are built as full 16-byte records (uv code included) and stored
// in the library's record layout.
using HeFull = HeRec<16>;

__device__ __forceinline__ void store_children(He *o, const HeFull c[4], bool topo, bool uvs) {
    if (topo && (uvs || !kHasUv)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) he_store(o, k, make_int4(c[k].twin, c[k].edge, c[k].vert, c[k].uv));
    } else if (topo) {
#pragma unroll
        for (int k = 0; k < 4; ++k) he_store_ids(o, k, make_int4(c[k].twin, c[k].edge, c[k].vert, 0));
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) he_store_uv(o, k, c[k].uv);
    }
}

// children twin/edge/vertex IDs (CatmullClark.h:2721-2737, :2777-2793)
__device__ __forceinline__ void child_topology(HeFull c[4], int32_t h, int32_t twin, int32_t prev, int32_t next,
                                               int32_t face, int32_t edge, int32_t vert, int32_t prevEdge,
                                               int32_t prevTwin, int32_t twinNext, int32_t V, int32_t F,
                                               int32_t E) {
    c[0].twin = i4(twinNext) + 3;
    c[1].twin = i4(next) + 2;
    c[2].twin = i4(prev) + 1;
    c[3].twin = i4(prevTwin) + 0;
    c[0].edge = 2 * edge + (h > twin ? 0 : 1);
    c[1].edge = 2 * E + h;
    c[2].edge = 2 * E + prev;
    c[3].edge = 2 * prevEdge + (prev > prevTwin ? 1 : 0);
    c[0].vert = vert;
    c[1].vert = V + F + edge;
    c[2].vert = V + face;
    c[3].vert = V + F + prevEdge;
}

// ---- cage halfedges + uvs (CatmullClark.h:2693-2740, :2825-2869) ----------
// uvs: 0 leaves uvIDs untouched, 1 refines them, 2 writes 0 (cages without uvs)
__device__ inline void cage_halfedges_at(DCage c, He *out, int topo, int uvs, int32_t h) {
    if (h >= c.H) return;
    const CageHe x = c.he[h];
    HeFull ch[4];
    if (topo) {
        const CageHe p = c.he[x.prev];
        int32_t twinNext = x.twin >= 0 ? c.he[x.twin].next : -1;
        child_topology(ch, h, x.twin, x.prev, x.next, x.face, x.edge, x.vert, p.edge, p.twin, twinNext,
                       c.V, c.F, c.E);
    }
    if (uvs == 1) {
        UV uv = c.uvs[x.uv];
        UV nuv = c.uvs[c.he[x.next].uv];
        UV puv = c.uvs[c.he[x.prev].uv];
        UV fuv = uv;
        int32_t m = 1;
        for (int32_t it = x.next, g = 0; it != h && g < kWalkCap; it = c.he[it].next, ++g) {
            UV t = c.uvs[c.he[it].uv];
            fuv.u += t.u;
            fuv.v += t.v;
            ++m;
        }
        fuv.u /= (float)m;
        fuv.v /= (float)m;
        ch[0].uv = uv_encode(uv);
        ch[1].uv = uv_encode(uv_mid(uv, nuv));
        ch[2].uv = uv_encode(fuv);
        ch[3].uv = uv_encode(uv_mid(uv, puv));
    } else if (uvs == 2) {
        ch[0].uv = ch[1].uv = ch[2].uv = ch[3].uv = 0;
    }
    store_children(out + 4 * (int64_t)h, ch, topo, uvs);
}

// ---- creases (CatmullClark.h:2950-2983, :2992-3026) -----------------------
__device__ inline void creases_at(const Crease *__restrict__ in, int32_t C, Crease *__restrict__ out, int32_t e) {
    if (e >= C) return;
    const Crease x = in[e];
    const int32_t n = x.next, p = x.prev;
    const Crease xn = in[n], xp = in[p];
    const bool t1 = xn.prev == e && n != e;
    const bool t2 = xp.next == e && p != e;
    const float s3 = 3.0f * x.sharp;
    Crease c0, c1;
    c0.next = 2 * e + 1;
    c1.next = 2 * n + (t1 ? 0 : 1);
    c0.prev = 2 * p + (t2 ? 1 : 0);
    c1.prev = 2 * e + 0;
    c0.sharp = fmaxr(0.0f, (xp.sharp + s3) / 4.0f - 1.0f);
    c1.sharp = fmaxr(0.0f, (s3 + xn.sharp) / 4.0f - 1.0f);
    out[2 * (int64_t)e + 0] = c0;
    out[2 * (int64_t)e + 1] = c1;
}

// ==== cage points (level 0 -> 1) ===========================================
// Faces and vertex fans of at most kListFan halfedges take their halfedges
// from the cage's ring lists and load everything at once; longer ones walk.
constexpr int kListFan = 8;

// rank of entry i among the first m entries of r (distinct halfedge IDs):
// its position in ascending order
__device__ __forceinline__ int list_rank(const int32_t (&r)[kListFan], int m, int i) {
    int k = 0;
#pragma unroll
    for (int j = 0; j < kListFan; ++j) k += (j < m && r[j] < r[i]) ? 1 : 0;
    return k;
}

// face points: gather CatmullClark.h:1408-1435; scatter :1437-1463 summed per
// face in ascending halfedge order (the order of a sequential scatter)
__device__ inline void cage_faces_at(DCage c, P3 *fp, int gather, int32_t f) {
    if (f >= c.F) return;
    const int32_t h0 = c.f2h[f];
    if (gather) {
        P3 acc = c.pts[c.he[h0].vert];
        float n = 1.0f;
        for (int32_t it = c.he[h0].next, g = 0; it != h0 && g < kWalkCap; it = c.he[it].next, ++g) {
            acc = add(acc, c.pts[c.he[it].vert]);
            n = n + 1.0f;
        }
        fp[f] = mul(acc, 1.0f / n);
        return;
    }
    if (c.fring) {
        // faces of up to kListFan halfedges from the ring list: every point
        // loaded at once, then added by the rank of its halfedge
        const int32_t o = c.fring_off[f], m = c.fring_off[f + 1] - o;
        if (m >= 1 && m <= kListFan) {
            int32_t r[kListFan];
            P3 pt[kListFan];
#pragma unroll
            for (int i = 0; i < kListFan; ++i) r[i] = c.fring[o + (i < m ? i : m - 1)];
#pragma unroll
            for (int i = 0; i < kListFan; ++i) pt[i] = c.pts[c.he[r[i]].vert];
            const float n = (float)m;
            P3 acc = zero3();
#pragma unroll
            for (int k = 0; k < kListFan; ++k) {
                if (k >= m) break;
                P3 v = pt[0];
#pragma unroll
                for (int i = 0; i < kListFan; ++i)
                    if (i < m && list_rank(r, m, i) == k) v = pt[i];
                acc.x += v.x / n;
                acc.y += v.y / n;
                acc.z += v.z / n;
            }
            fp[f] = acc;
            return;
        }
    }
    float n = 1.0f;
    int32_t lo = h0;
    for (int32_t it = c.he[h0].next, g = 0; it != h0 && g < kWalkCap; it = c.he[it].next, ++g) {
        n = n + 1.0f;
        lo = it < lo ? it : lo;
    }
    // visit the face's halfedges in ascending ID order
    P3 acc = zero3();
    int32_t cur = lo;
    for (int g = 0; g < kWalkCap; ++g) {
        P3 v = c.pts[c.he[cur].vert];
        acc.x += v.x / n;
        acc.y += v.y / n;
        acc.z += v.z / n;
        int32_t nxt = INT32_MAX;
        for (int32_t it = c.he[h0].next, g = 0; g < kWalkCap; it = c.he[it].next, ++g) {
            if (it > cur && it < nxt) nxt = it;
            if (it == h0) break;
        }
        if (nxt == INT32_MAX) break;
        cur = nxt;
    }
    fp[f] = acc;
}

// scatter-form contribution of cage halfedge h to its edge point
// (CatmullClark.h:1527-1545 smooth, :1612-1645 creased)
__device__ __forceinline__ P3 cage_edge_contrib(const DCage &c, const P3 *fp, int32_t h, bool creased) {
    const CageHe x = c.he[h];
    const P3 v0 = c.pts[x.vert], v1 = c.pts[c.he[x.next].vert], f = fp[x.face];
    if (creased) {
        const float w = sat(c.cr[x.edge].sharp);
        P3 sharp = mul(lerp(v0, v1, 0.5f), x.twin < 0 ? 1.0f : 0.5f);
        P3 smooth = mul(lerp(v0, f, 0.5f), 0.5f);
        return lerp(smooth, sharp, w);
    }
    const float w = x.twin >= 0 ? 0.5f : 1.0f;
    return lerp(mul(f, 0.5f), lerp(mul(v0, w), mul(v1, w), 0.5f), w);
}

__device__ __forceinline__ P3 edge_gather_value(P3 v0, P3 v1, P3 f0, P3 f1, float w, bool creased) {
    P3 t1 = add(v0, v1), t2 = add(f0, f1);
    P3 sharp = mul(t1, 0.5f);
    P3 smooth = mul(add(t1, t2), 0.25f);
    return creased ? lerp(smooth, sharp, w) : lerp(sharp, smooth, w);
}

// cage edge points: gather :1476-1515 / :1561-1600, one thread per edge;
// scatter :1517-1548 / :1602-1648, one thread per edge adding both halfedge
// contributions (two-term sums are order independent)
__device__ inline void cage_edges_at(DCage c, const P3 *fp, P3 *ep, int gather, int creased, int32_t e) {
    if (e >= c.E) return;
    const int32_t h = c.e2h[e];
    const CageHe x = c.he[h];
    if (gather) {
        const float w = creased ? sat(c.cr[e].sharp) : (x.twin < 0 ? 0.0f : 1.0f);
        ep[e] = edge_gather_value(c.pts[x.vert], c.pts[c.he[x.next].vert], fp[x.face],
                                  fp[c.he[imax(0, x.twin)].face], w, creased);
        return;
    }
    P3 acc = add(zero3(), cage_edge_contrib(c, fp, h, creased));
    if (x.twin >= 0) acc = add(acc, cage_edge_contrib(c, fp, x.twin, creased));
    ep[e] = acc;
}

// cage vertex points, gather (smooth :1661-1711, creased :1769-1869)
__device__ inline void cage_vertices_gather_at(DCage c, const P3 *fp, const P3 *ep, P3 *vp, int creased, int32_t v) {
    if (v >= c.V) return;
    const int32_t h = c.v2h[v];
    const P3 old = c.pts[v];
    if (!creased) {
        P3 acc = add(mul(fp[c.he[h].face], -1.0f), mul(ep[c.he[h].edge], 4.0f));
        float n = 1.0f;
        int32_t it;
        for (it = c.he[c.he[h].prev].twin; it >= 0 && it != h && n < (float)kWalkCap; it = c.he[c.he[it].prev].twin) {
            acc = add(acc, mul(fp[c.he[it].face], -1.0f));
            acc = add(acc, mul(ep[c.he[it].edge], 4.0f));
            n = n + 1.0f;
        }
        P3 sm = add(mul(acc, 1.0f / (n * n)), mul(old, 1.0f - 3.0f / n));
        vp[v] = lerp(old, sm, it != h ? 0.0f : 1.0f);
        return;
    }
    const int32_t p = c.he[h].prev;
    const float ts = c.cr[c.he[h].edge].sharp, ps = c.cr[c.he[p].edge].sharp;
    const float cw = sgn(ts), pcw = sgn(ps);
    const P3 E = ep[c.he[h].edge], PE = ep[c.he[p].edge], PF = fp[c.he[p].face];
    P3 sm = add(mul(PF, -1.0f), mul(PE, 4.0f));
    P3 cr = add(zero3(), mul(PE, pcw));
    float avgS = ps, cc = pcw, n = 1.0f;
    int32_t fw;
    for (fw = c.he[p].twin; fw >= 0 && fw != h && n < (float)kWalkCap; fw = c.he[fw].twin) {
        const int32_t q = c.he[fw].prev;
        const P3 qE = ep[c.he[q].edge], qF = fp[c.he[q].face];
        const float qs = c.cr[c.he[q].edge].sharp, qw = sgn(qs);
        sm = add(sm, mul(qF, -1.0f));
        sm = add(sm, mul(qE, 4.0f));
        n = n + 1.0f;
        cr = add(cr, mul(qE, qw));
        avgS = avgS + qs;
        cc = cc + qw;
        fw = q;
    }
    if (fw < 0) {
        cr = add(cr, mul(E, cw));
        cc = cc + cw;
        n = n + 1.0f;
    }
    sm = add(mul(sm, 1.0f / (n * n)), mul(old, 1.0f - 3.0f / n));
    cr = add(mul(cr, 0.25f), mul(old, 0.5f));
    if (cc <= 1.0f) vp[v] = sm;
    else if (cc >= 3.0f || n == 2.0f) vp[v] = old;
    else vp[v] = lerp(old, cr, sat(avgS * 0.5f));
}

// The scatter-form vertex rule of one cage halfedge from the results of its
// ring walk: smooth (:1742-1748) from the valence n and s = 0 on a boundary;
// creased (:1938-1987) from the walk's valence, crease count and summed
// sharpness, open when the walk met a boundary.
__device__ __forceinline__ P3 cage_smooth_share(P3 old, P3 e, P3 f, int32_t n, float s) {
    const float w = 1.0f / (float)n;
    P3 a;
    a.x = w * (old.x + w * s * (4.0f * e.x - f.x - 3.0f * old.x));
    a.y = w * (old.y + w * s * (4.0f * e.y - f.y - 3.0f * old.y));
    a.z = w * (old.z + w * s * (4.0f * e.z - f.z - 3.0f * old.z));
    return a;
}

__device__ __forceinline__ P3 cage_creased_share(P3 old, P3 E, P3 PE, P3 F, float cw, float pcw, float avgS, float cc,
                                                 float n, bool open) {
    const P3 corner = mul(old, 1.0f / n);
    P3 cr = add(mul(old, 0.25f * cw), mul(E, 0.25f * cw));
    P3 sm = add(mul(F, -1.0f), mul(E, 4.0f));
    sm = add(sm, mul(old, n - 3.0f));
    sm = mul(sm, 1.0f / (n * n));
    if (open) {
        cc = cc + cw;
        n = n + 1.0f;
        cr = add(cr, add(mul(old, 0.25f * pcw), mul(PE, 0.25f * pcw)));
    }
    if (cc <= 1.0f) return sm;
    if (cc >= 3.0f || n == 2.0f) return corner;
    return lerp(corner, cr, sat(avgS * 0.5f));
}

// scatter-form contribution of cage halfedge h to its vertex point (smooth
// :1713-1756, creased :1872-1991), walking h's ring as the reference does
__device__ inline P3 cage_vertex_contrib(const DCage &c, const P3 *fp, const P3 *ep, int creased, int32_t h) {
    const CageHe x = c.he[h];
    const P3 old = c.pts[x.vert];
    if (!creased) {
        int32_t n = 1, fw, bw;
        for (fw = c.he[x.prev].twin; fw >= 0 && fw != h && n < kWalkCap; fw = c.he[c.he[fw].prev].twin) ++n;
        for (bw = x.twin >= 0 ? c.he[x.twin].next : -1; fw < 0 && bw >= 0 && bw != h && n < kWalkCap;
             bw = c.he[bw].twin >= 0 ? c.he[c.he[bw].twin].next : -1)
            ++n;
        return cage_smooth_share(old, ep[x.edge], fp[x.face], n, fw < 0 ? 0.0f : 1.0f);
    }
    const int32_t p = x.prev;
    const float ts = c.cr[x.edge].sharp, ps = c.cr[c.he[p].edge].sharp;
    const P3 PE = ep[c.he[p].edge], E = ep[x.edge], F = fp[x.face];
    float avgS = ps, cc = sgn(ps), n = 1.0f;
    int32_t fw, bw;
    for (fw = c.he[p].twin; fw >= 0 && fw != h && n < (float)kWalkCap; fw = c.he[fw].twin) {
        const int32_t q = c.he[fw].prev;
        const float qs = c.cr[c.he[q].edge].sharp;
        n = n + 1.0f;
        avgS = avgS + qs;
        cc = cc + sgn(qs);
        fw = q;
    }
    for (bw = x.twin; fw < 0 && bw >= 0 && bw != h && n < (float)kWalkCap; bw = c.he[bw].twin) {
        const int32_t q = c.he[bw].next;
        const float qs = c.cr[c.he[q].edge].sharp;
        n = n + 1.0f;
        avgS = avgS + qs;
        cc = cc + sgn(qs);
        bw = q;
    }
    return cage_creased_share(old, E, PE, F, sgn(ts), sgn(ps), avgS, cc, n, fw < 0);
}

// entry j (< kListFan) of v, by selects (no register indexed at run time)
__device__ __forceinline__ float pick_list(const float (&v)[kListFan], int j) {
    float x = v[0];
#pragma unroll
    for (int m = 1; m < kListFan; ++m) x = j == m ? v[m] : x;
    return x;
}

// A closed fan of n <= kListFan halfedges from the cage's ring list R (in
// walk order: R[j+1] = twin(prev(R[j])), indices mod n), when the ring holds
// every halfedge leaving v.  The walk of R[k] visits R[k+1] .. R[k+n-1], and
// the previous edge of R[j] is the edge of R[j+1]: its valence is n, its
// crease count the sum over the ring's edges, and its summed sharpness
// s[k+1] + s[k+2] + ... + s[k+n] in that order (s[j] the sharpness of R[j]'s
// edge), so every contribution follows from the ring's records.
__device__ inline bool cage_vertex_scatter_list(const DCage &c, const P3 *fp, const P3 *ep, P3 *vp, int creased,
                                                int32_t v) {
    const int32_t o = c.vring_off[v], n = c.vring_off[v + 1] - o;
    if (c.vring_bnd[v] != 0 || n < 1 || n > kListFan || c.vout_off[v + 1] - c.vout_off[v] != n) return false;
    int32_t r[kListFan], e[kListFan], f[kListFan];
#pragma unroll
    for (int i = 0; i < kListFan; ++i) r[i] = c.vring[o + (i < n ? i : n - 1)];
#pragma unroll
    for (int i = 0; i < kListFan; ++i) {
        const CageHe x = c.he[r[i]];
        e[i] = x.edge;
        f[i] = x.face;
    }
    float s[kListFan];
    float cc = 0.0f;
#pragma unroll
    for (int i = 0; i < kListFan; ++i) {
        s[i] = creased ? c.cr[e[i]].sharp : 0.0f;
        if (i < n) cc = cc + sgn(s[i]);
    }
    const P3 old = c.pts[v];
    P3 con[kListFan];
#pragma unroll
    for (int k = 0; k < kListFan; ++k) {
        const P3 E = ep[e[k]], F = fp[f[k]];
        if (!creased) {
            con[k] = cage_smooth_share(old, E, F, n, 1.0f);
            continue;
        }
        // k < n, i <= n: k + i < 2n
        const float ps = pick_list(s, k + 1 < n ? k + 1 : k + 1 - n);
        float avgS = ps;
#pragma unroll
        for (int i = 2; i <= kListFan; ++i)
            if (i <= n) avgS = avgS + pick_list(s, k + i < n ? k + i : k + i - n);
        // the previous edge's point is read only on a boundary
        con[k] = cage_creased_share(old, E, E, F, sgn(s[k]), sgn(ps), avgS, cc, (float)n, false);
    }
    P3 acc = zero3();
#pragma unroll
    for (int k = 0; k < kListFan; ++k) {
        if (k >= n) break;
        P3 ck = con[0];
#pragma unroll
        for (int i = 0; i < kListFan; ++i)
            if (i < n && list_rank(r, n, i) == k) ck = con[i];
        acc = add(acc, ck);
    }
    vp[v] = acc;
    return true;
}

// cage vertex points, scatter: one thread per vertex adds the contributions
// of the halfedges leaving it from zero in ascending halfedge order, the
// order of the reference's sequential scatter (:1750, :1986-1987), so the
// result does not depend on scheduling.  With the ring lists: the list form
// above, else every halfedge of the vertex's outgoing list; without them the
// ring walked from v's halfedge.
__device__ inline void cage_vertices_scatter_at(DCage c, const P3 *fp, const P3 *ep, P3 *vp, int creased, int32_t v) {
    if (v >= c.V) return;
    if (c.vring && c.vout) {
        if (cage_vertex_scatter_list(c, fp, ep, vp, creased, v)) return;
        P3 acc = zero3();
        for (int32_t i = c.vout_off[v]; i < c.vout_off[v + 1]; ++i)
            acc = add(acc, cage_vertex_contrib(c, fp, ep, creased, c.vout[i]));
        vp[v] = acc;
        return;
    }
    const int32_t h0 = c.v2h[v];
    P3 acc = zero3();
    int32_t last = -1;
    for (int k = 0; k < kWalkCap; ++k) {
        // the smallest halfedge leaving v above `last`: walk forward
        // (twin(prev(h))) and, past a boundary, backward (next(twin(h)))
        int32_t best = INT32_MAX;
        int32_t h = h0, g = 0;
        for (; h >= 0 && g < kWalkCap; ++g) {
            if (h > last && h < best) best = h;
            h = c.he[c.he[h].prev].twin;
            if (h == h0) break;
        }
        if (h < 0) {
            for (h = c.he[h0].twin >= 0 ? c.he[c.he[h0].twin].next : -1, g = 0; h >= 0 && g < kWalkCap; ++g) {
                if (h > last && h < best) best = h;
                h = c.he[h].twin >= 0 ? c.he[c.he[h].twin].next : -1;
            }
        }
        if (best == INT32_MAX) break;
        acc = add(acc, cage_vertex_contrib(c, fp, ep, creased, best));
        last = best;
    }
    vp[v] = acc;
}

}  // namespace cch
