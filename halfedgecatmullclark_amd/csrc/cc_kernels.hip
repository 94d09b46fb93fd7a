// Per-level refinement kernels for gfx950, one thread per output element
// (halfedge, crease, face, edge or vertex).  Each kernel cites the
// CatmullClark.h rule it implements; the arithmetic keeps the reference's
// operation order (built with -ffp-contract=off).
//
// Level d arrays are the cumulative slices of the device mirror of cc_Subd;
// the cage (level 0) has its own records.  Points of level d+1 are laid out
// [V_d vertex points | F_d face points | E_d edge points].
#include "cc_kernels.h"
#include "cc_level0.h"

namespace cch {

__device__ __forceinline__ uint32_t gid() { return blockIdx.x * blockDim.x + threadIdx.x; }

__global__ void k_cage_halfedges(DCage c, He *out, int topo, int uvs) { cage_halfedges_at(c, out, topo, uvs, (int32_t)gid()); }

// ---- level halfedges + uvs (CatmullClark.h:2749-2796, :2878-2921) ---------
__global__ void k_level_halfedges(const He *__restrict__ in, He *__restrict__ out, int32_t H, int32_t V,
                                  int32_t F, int32_t E, int topo, int uvs) {
    int32_t h = (int32_t)gid();
    if (h >= H) return;
    const HeView q4{in + (h & ~3)};
    const int j = h & 3;
    int4 r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) r[k] = q4[k];
    const int4 x = r[j];
    const int4 p = r[(j + 3) & 3];
    HeFull ch[4];
    if (topo) {
        child_topology(ch, h, x.x, qprev(h), qnext(h), h >> 2, x.y, x.z, p.y, p.x, qnext(x.x), V, F, E);
    }
    if (uvs == 1) {
        UV uv = uv_decode(x.w);
        UV nuv = uv_decode(r[(j + 1) & 3].w);
        UV puv = uv_decode(p.w);
        UV fuv = uv;
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            UV t = uv_decode(r[(j + k) & 3].w);
            fuv.u += t.u;
            fuv.v += t.v;
        }
        fuv.u /= 4.0f;
        fuv.v /= 4.0f;
        ch[0].uv = x.w;
        ch[1].uv = uv_encode_unit(uv_mid(uv, nuv));
        ch[2].uv = uv_encode_unit(fuv);
        ch[3].uv = uv_encode_unit(uv_mid(uv, puv));
    } else if (uvs == 2) {
        ch[0].uv = ch[1].uv = ch[2].uv = ch[3].uv = 0;
    }
    store_children(out + 4 * (int64_t)h, ch, topo, uvs);
}

__global__ void k_creases(const Crease *__restrict__ in, int32_t C, Crease *__restrict__ out) { creases_at(in, C, out, (int32_t)gid()); }

__global__ void k_set_address(const void **slot, const void *p) {
    if (threadIdx.x == 0) *slot = p;
}

__global__ void k_zero(float *p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0f;
}

__global__ void k_cage_faces(DCage c, P3 *fp, int gather) { cage_faces_at(c, fp, gather, (int32_t)gid()); }
__global__ void k_cage_edges(DCage c, const P3 *fp, P3 *ep, int gather, int creased) { cage_edges_at(c, fp, ep, gather, creased, (int32_t)gid()); }
__global__ void k_cage_vertices_gather(DCage c, const P3 *fp, const P3 *ep, P3 *vp, int creased) { cage_vertices_gather_at(c, fp, ep, vp, creased, (int32_t)gid()); }
__global__ void k_cage_vertices_scatter(DCage c, const P3 *fp, const P3 *ep, P3 *vp, int creased) { cage_vertices_scatter_at(c, fp, ep, vp, creased, (int32_t)gid()); }

// ==== level-d points (d >= 1, all faces quads) ===========================
__device__ __forceinline__ float lsharp(const Lvl &L, int32_t e) { return e < L.C ? L.cr[e].sharp : 0.0f; }
__device__ __forceinline__ P3 lpt(const Lvl &L, int32_t h) { return L.old[L.he[h].vert]; }

// face points: gather :2004-2030 ((v0+v1+v2+v3)*0.25), scatter :2032-2052
// (v/4 added in ascending halfedge order)
__global__ void k_faces(Lvl L, int gather) {
    int32_t f = (int32_t)gid();
    if (f >= L.F) return;
    const HeView q4{L.he + 4 * (int64_t)f};
    P3 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = L.old[q4[k].z];
    P3 acc;
    if (gather) {
        acc = add(add(add(v[0], v[1]), v[2]), v[3]);
        acc = mul(acc, 0.25f);
    } else {
        acc = zero3();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            acc.x += v[k].x / 4.0f;
            acc.y += v[k].y / 4.0f;
            acc.z += v[k].z / 4.0f;
        }
    }
    L.fp[f] = acc;
}

// scatter-form edge contribution of halfedge h (:2118-2136 smooth,
// :2205-2238 creased)
__device__ __forceinline__ P3 level_edge_contrib(const Lvl &L, int32_t h, bool creased) {
    const He x = L.he[h];
    const P3 v0 = L.old[x.vert], v1 = lpt(L, qnext(h)), f = L.fp[h >> 2];
    if (creased) {
        const float w = sat(lsharp(L, x.edge));
        P3 sharp = mul(lerp(v0, v1, 0.5f), x.twin < 0 ? 1.0f : 0.5f);
        P3 smooth = mul(lerp(v0, f, 0.5f), 0.5f);
        return lerp(smooth, sharp, w);
    }
    const float w = x.twin >= 0 ? 0.5f : 1.0f;
    return lerp(mul(f, 0.5f), lerp(mul(v0, w), mul(v1, w), 0.5f), w);
}

// edge points, scatter: one thread per halfedge; the halfedge with the larger
// ID of its pair (or a boundary halfedge) writes the sum of both
// contributions, which is what a sequential scatter produces
__global__ void k_edges_scatter(Lvl L, int creased) {
    int32_t h = (int32_t)gid();
    if (h >= L.H) return;
    const He x = L.he[h];
    if (!(h > x.twin)) return;
    P3 acc = zero3();
    if (x.twin >= 0) acc = add(acc, level_edge_contrib(L, x.twin, creased));
    acc = add(acc, level_edge_contrib(L, h, creased));
    L.ep[x.edge] = acc;
}

// edge points, gather :2065-2105 / :2151-2191, one thread per edge through
// the O(depth) edge map
__global__ void k_edges_gather(Lvl L, DCage c, LevelTab t, int d, int creased) {
    int32_t e = (int32_t)gid();
    if (e >= L.E) return;
    const int32_t h = edge_rep(c, t, e, d);
    const int32_t tw = L.he[h].twin;
    const float w = creased ? sat(lsharp(L, e)) : (tw < 0 ? 0.0f : 1.0f);
    L.ep[e] = edge_gather_value(lpt(L, h), lpt(L, qnext(h)), L.fp[h >> 2], L.fp[imax(0, tw) >> 2], w,
                                creased);
}

// vertex points, gather (smooth :2254-2305, creased :2364-2493), one thread
// per vertex through the O(depth) vertex map
__global__ void k_vertices_gather(Lvl L, DCage c, LevelTab t, int d, int creased) {
    int32_t v = (int32_t)gid();
    if (v >= L.V) return;
    const int32_t h = vertex_rep(c, t, v, d);
    const P3 old = L.old[v];
    if (!creased) {
        P3 acc = add(mul(L.fp[h >> 2], -1.0f), mul(L.ep[L.he[h].edge], 4.0f));
        float n = 1.0f;
        int32_t it;
        for (it = L.he[qprev(h)].twin; it >= 0 && it != h && n < (float)kWalkCap; it = L.he[qprev(it)].twin) {
            acc = add(acc, mul(L.fp[it >> 2], -1.0f));
            acc = add(acc, mul(L.ep[L.he[it].edge], 4.0f));
            n = n + 1.0f;
        }
        P3 sm = add(mul(acc, 1.0f / (n * n)), mul(old, 1.0f - 3.0f / n));
        L.vp[v] = lerp(old, sm, it != h ? 0.0f : 1.0f);
        return;
    }
    const int32_t p = qprev(h);
    const He xh = L.he[h], xp = L.he[p];
    const float ts = lsharp(L, xh.edge), ps = lsharp(L, xp.edge);
    const float cw = sgn(ts), pcw = sgn(ps);
    const P3 E = L.ep[xh.edge], PE = L.ep[xp.edge], PF = L.fp[p >> 2];
    P3 sm = add(mul(PF, -1.0f), mul(PE, 4.0f));
    P3 cr = add(zero3(), mul(PE, pcw));
    float avgS = ps, cc = pcw, n = 1.0f;
    int32_t fw, bw;
    for (fw = xp.twin; fw >= 0 && fw != h && n < (float)kWalkCap; fw = L.he[fw].twin) {
        const int32_t q = qprev(fw);
        const int32_t qe = L.he[q].edge;
        const P3 qE = L.ep[qe], qF = L.fp[q >> 2];
        const float qs = lsharp(L, qe), qw = sgn(qs);
        sm = add(sm, mul(qF, -1.0f));
        sm = add(sm, mul(qE, 4.0f));
        n = n + 1.0f;
        cr = add(cr, mul(qE, qw));
        avgS = avgS + qs;
        cc = cc + qw;
        fw = q;
    }
    for (bw = xh.twin; fw < 0 && bw >= 0 && bw != h && n < (float)kWalkCap; bw = L.he[bw].twin) {
        const int32_t q = qnext(bw);
        const int32_t qe = L.he[q].edge;
        const P3 qE = L.ep[qe], qF = L.fp[q >> 2];
        const float qs = lsharp(L, qe), qw = sgn(qs);
        sm = add(sm, mul(qF, -1.0f));
        sm = add(sm, mul(qE, 4.0f));
        n = n + 1.0f;
        cr = add(cr, mul(qE, qw));
        avgS = avgS + qs;
        cc = cc + qw;
        bw = q;
    }
    if (fw < 0) {
        cr = add(cr, mul(E, cw));
        cc = cc + cw;
        n = n + 1.0f;
    }
    sm = add(mul(sm, 1.0f / (n * n)), mul(old, 1.0f - 3.0f / n));
    cr = add(mul(cr, 0.5f / cc), mul(old, 0.5f));
    if (cc <= 1.0f) L.vp[v] = sm;
    else if (cc >= 3.0f || n == 2.0f) L.vp[v] = old;
    else L.vp[v] = lerp(old, cr, sat(avgS * 0.5f));
}

// vertex points, scatter: per-halfedge contributions (smooth :2307-2351,
// creased :2496-2615, whose rule selection tests the corner case first)
// added with float atomics into zeroed vertex points
__global__ void k_vertices_scatter(Lvl L, int creased) {
    int32_t h = (int32_t)gid();
    if (h >= L.H) return;
    const He x = L.he[h];
    const P3 old = L.old[x.vert];
    P3 a;
    if (!creased) {
        int32_t n = 1, fw, bw;
        for (fw = L.he[qprev(h)].twin; fw >= 0 && fw != h && n < kWalkCap; fw = L.he[qprev(fw)].twin) ++n;
        for (bw = qnext(x.twin); fw < 0 && bw >= 0 && bw != h && n < kWalkCap; bw = qnext(L.he[bw].twin)) ++n;
        const float w = 1.0f / (float)n;
        const float s = fw < 0 ? 0.0f : 1.0f;
        const P3 f = L.fp[h >> 2], e = L.ep[x.edge];
        a.x = w * (old.x + w * s * (4.0f * e.x - f.x - 3.0f * old.x));
        a.y = w * (old.y + w * s * (4.0f * e.y - f.y - 3.0f * old.y));
        a.z = w * (old.z + w * s * (4.0f * e.z - f.z - 3.0f * old.z));
    } else {
        const int32_t p = qprev(h);
        const He xp = L.he[p];
        const float ts = lsharp(L, x.edge), ps = lsharp(L, xp.edge);
        const float cw = sgn(ts), pcw = sgn(ps);
        const P3 PE = L.ep[xp.edge], E = L.ep[x.edge], F = L.fp[h >> 2];
        float avgS = ps, cc = pcw, n = 1.0f;
        int32_t fw, bw;
        for (fw = xp.twin; fw >= 0 && fw != h && n < (float)kWalkCap; fw = L.he[fw].twin) {
            const int32_t q = qprev(fw);
            const float qs = lsharp(L, L.he[q].edge);
            n = n + 1.0f;
            avgS = avgS + qs;
            cc = cc + sgn(qs);
            fw = q;
        }
        for (bw = x.twin; fw < 0 && bw >= 0 && bw != h && n < (float)kWalkCap; bw = L.he[bw].twin) {
            const int32_t q = qnext(bw);
            const float qs = lsharp(L, L.he[q].edge);
            n = n + 1.0f;
            avgS = avgS + qs;
            cc = cc + sgn(qs);
            bw = q;
        }
        const P3 corner = mul(old, 1.0f / n);
        P3 cr = add(mul(old, 0.25f * cw), mul(E, 0.25f * cw));
        P3 sm = add(mul(F, -1.0f), mul(E, 4.0f));
        sm = add(sm, mul(old, n - 3.0f));
        sm = mul(sm, 1.0f / (n * n));
        if (fw < 0) {
            cc = cc + cw;
            n = n + 1.0f;
            cr = add(cr, add(mul(old, 0.25f * pcw), mul(PE, 0.25f * pcw)));
        }
        if (cc >= 3.0f || n == 2.0f) a = corner;
        else if (cc <= 1.0f) a = sm;
        else a = lerp(corner, cr, sat(avgS * 0.5f));
    }
    unsafeAtomicAdd(&L.vp[x.vert].x, a.x);
    unsafeAtomicAdd(&L.vp[x.vert].y, a.y);
    unsafeAtomicAdd(&L.vp[x.vert].z, a.z);
}

}  // namespace cch
