// Adjoint of the point refinement.
//
// At fixed topology and sharpness every refined point is a fixed linear
// combination of the points one level up (the weights of the scatter rules,
// CatmullClark.h:1437-1991 and :2032-2615, depend only on valence, crease
// count, boundary and sharpness).  The adjoint applies the transposes level
// by level, finest first.  For a level step d -> d+1 (d >= 1) it runs in
// gather form, without atomics, in four passes:
//   1. weights: the owner of each level-d vertex evaluates the vertex rule's
//      weights (a, b, g, dl) of every halfedge of its ring;
//   2. edges: the owner of each edge adds the vertex rule's shares to the
//      edge's gradient;
//   3. faces: each quad adds the vertex and edge rules' shares to its face's
//      gradient;
//   4. vertices: the owner of each vertex gathers its level-d gradient from
//      the vertex, edge and face rules.
// The cage step (small, polygons of any size) runs as ordered gathers over
// the host-built ring lists (per-halfedge float atomics only for a cage
// whose walks do not close).  Only the stored topology and creases are read; nothing of the
// forward points is kept or recomputed.  Kernels are templated on the scalar
// type so the same code runs in float and, for gradient checks, in double.
#include "cc_adjoint.h"
#include "cc_batch.h"
#include "cc_fan.h"

namespace cch {

namespace {

template <typename T>
struct V3 { T x, y, z; };

template <typename T>
__device__ __forceinline__ void atomic_add3(V3<T> *p, T a, const V3<T> &g) {
    unsafeAtomicAdd(&p->x, a * g.x);
    unsafeAtomicAdd(&p->y, a * g.y);
    unsafeAtomicAdd(&p->z, a * g.z);
}

// weights (alpha, beta, gamma, delta) of one halfedge's vertex contribution
// alpha*V + beta*E(edge h) + gamma*F(face h) + delta*E(edge prev h)
template <typename T>
struct VW { T a, b, g, dl; };

template <typename T>
__device__ __forceinline__ T satT(T x) { return x > (T)1 ? (T)1 : (x < (T)0 ? (T)0 : x); }
template <typename T>
__device__ __forceinline__ T sgnT(T x) { return x < (T)0 ? (T)-1 : (x > (T)0 ? (T)1 : (T)0); }

// sharpness of level-d edge e in the adjoint's precision: the crease records
// in float, a separately refined array in double
template <typename T>
__device__ __forceinline__ T sharp_at(const AdjArgs<T> &a, int32_t e) {
    if (e >= a.C) return (T)0;
    return a.sd ? a.sd[e] : (T)a.cr[e].sharp;
}

template <typename T>
struct RingT {
    T n, cc;
    bool boundary;
    T s1, s2;
    int32_t e1, e2;
    T sall;
};

// level d >= 1 scatter rule (creased :2568-2607, corner test first; smooth
// :2339-2348)
template <typename T>
__device__ __forceinline__ VW<T> level_vertex_weights(const RingT<T> &rp, T ts, T ps, int32_t edge, bool creased) {
    VW<T> w{0, 0, 0, 0};
    const T n = rp.n;
    if (!creased) {
        const T iw = (T)1 / n, s = rp.boundary ? (T)0 : (T)1;
        w.a = iw - (T)3 * iw * iw * s;
        w.b = (T)4 * iw * iw * s;
        w.g = -iw * iw * s;
        return w;
    }
    const T nn = rp.boundary ? n + (T)1 : n;
    if (rp.cc >= (T)3 || nn == (T)2) {
        w.a = (T)1 / n;
        return w;
    }
    if (rp.cc <= (T)1) {
        w.a = (n - (T)3) / (n * n);
        w.b = (T)4 / (n * n);
        w.g = (T)-1 / (n * n);
        return w;
    }
    T avg = rp.sall;
    if (rp.boundary) avg = edge == rp.e1 ? rp.s2 : (edge == rp.e2 ? rp.s1 : rp.s1 + rp.s2);
    const T t = satT(avg * (T)0.5);
    const T cw = sgnT(ts), pcw = rp.boundary ? sgnT(ps) : (T)0;
    w.a = ((T)1 - t) / n + t * (T)0.25 * (cw + pcw);
    w.b = t * (T)0.25 * cw;
    w.dl = t * (T)0.25 * pcw;
    return w;
}

}  // namespace

// ---- level d >= 1, gather form -------------------------------------------------
template <typename T>
struct W4 { T a, b, g, dl; };

// The weights live as four arrays of the level's H scalars, a | b | g | dl,
// so that each pass reads only the one it needs.
template <typename T>
__device__ __forceinline__ void w_store(const AdjArgs<T> &a, int32_t h, T wa, T wb, T wg, T wdl) {
    const int64_t H = a.H;
    a.W[h] = wa;
    a.W[H + h] = wb;
    a.W[2 * H + h] = wg;
    a.W[3 * H + h] = wdl;
}
template <typename T>
__device__ __forceinline__ T w_a(const AdjArgs<T> &a, int32_t h) { return a.W[h]; }
template <typename T>
__device__ __forceinline__ T w_b(const AdjArgs<T> &a, int32_t h) { return a.W[(int64_t)a.H + h]; }
template <typename T>
__device__ __forceinline__ T w_g(const AdjArgs<T> &a, int32_t h) { return a.W[2 * (int64_t)a.H + h]; }
template <typename T>
__device__ __forceinline__ T w_dl(const AdjArgs<T> &a, int32_t h) { return a.W[3 * (int64_t)a.H + h]; }

// weights of one halfedge h in the edge rule (creased :2221-2233, smooth
// :2125-2131): c0 on its vertex, c1 on the next vertex, cf on its face point
template <typename T, int CREASED>
__device__ __forceinline__ void edge_weights(const AdjArgs<T> &a, int32_t edge, bool boundary, T &c0, T &c1, T &cf) {
    if (CREASED) {
        const T w = satT(sharp_at(a, edge));
        const T k = boundary ? (T)1 : (T)0.5;
        c0 = (T)0.25 * ((T)1 - w) + (T)0.5 * k * w;
        c1 = (T)0.5 * k * w;
        cf = (T)0.25 * ((T)1 - w);
    } else {
        const T w = boundary ? (T)1 : (T)0.5;
        c0 = c1 = (T)0.5 * w * w;
        cf = (T)0.5 * ((T)1 - w);
    }
}

// ring parameters of a vertex from its fan (valence, creases, boundary)
template <typename T, int CREASED>
__device__ __forceinline__ RingT<T> ring_params(const AdjArgs<T> &a, const DCage &c, const Fan &f) {
    const He *he = a.he;
    RingT<T> rp;
    rp.n = (T)f.n; rp.cc = 0; rp.boundary = f.boundary; rp.s1 = rp.s2 = 0; rp.e1 = rp.e2 = -1; rp.sall = 0;
    if (CREASED) {
        for (int i = (f.boundary ? -1 : 0); i < f.n; ++i) {
            const int32_t e = i < 0 ? he[fan_at(f, c, 0)].edge : he[qprev(fan_at(f, c, i))].edge;
            const T se = sharp_at(a, e);
            rp.cc += sgnT(se);
            if (se != (T)0) {
                if (rp.e1 < 0) { rp.s1 = se; rp.e1 = e; }
                else if (rp.e2 < 0) { rp.s2 = se; rp.e2 = e; }
            }
            rp.sall += se;
        }
    }
    return rp;
}

// ring crease parameter update for the sharpness se of edge e (the order of
// ring_params)
template <typename T>
__device__ __forceinline__ void ring_add(RingT<T> &rp, T se, int32_t e) {
    rp.cc += sgnT(se);
    if (se != (T)0) {
        if (rp.e1 < 0) { rp.s1 = se; rp.e1 = e; }
        else if (rp.e2 < 0) { rp.s2 = se; rp.e2 = e; }
    }
    rp.sall += se;
}

// pass 1: the vertex rule's weights of every halfedge of a vertex's ring
template <typename T, int CREASED>
__device__ __forceinline__ void adj_ring_weights(const AdjArgs<T> &a, const DCage &c, const Fan &f) {
    const He *he = a.he;
    if (f.n <= 4) {
        // a fan of at most four: every entry's loads issued before any is
        // used, the same sums in the same order as below
        int32_t r[4], re[4], pe[4];
        T ts[4], ps[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = fan_at(f, c, q < f.n ? q : 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            re[q] = he[r[q]].edge;
            pe[q] = he[qprev(r[q])].edge;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            ts[q] = CREASED ? sharp_at(a, re[q]) : (T)0;
            ps[q] = CREASED ? sharp_at(a, pe[q]) : (T)0;
        }
        RingT<T> rp;
        rp.n = (T)f.n; rp.cc = 0; rp.boundary = f.boundary; rp.s1 = rp.s2 = 0; rp.e1 = rp.e2 = -1; rp.sall = 0;
        if (CREASED) {
            if (f.boundary) ring_add(rp, ts[0], re[0]);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (i < f.n) ring_add(rp, ps[i], pe[i]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (i >= f.n) break;
            const VW<T> w = level_vertex_weights<T>(rp, ts[i], ps[i], re[i], CREASED);
            w_store(a, r[i], w.a, w.b, w.g, w.dl);
        }
        return;
    }
    const RingT<T> rp = ring_params<T, CREASED>(a, c, f);
    for (int i = 0; i < f.n; ++i) {
        const int32_t r = fan_at(f, c, i);
        const int32_t re = he[r].edge, pe = he[qprev(r)].edge;
        const T ts = CREASED ? sharp_at(a, re) : (T)0;
        const T ps = CREASED ? sharp_at(a, pe) : (T)0;
        const VW<T> w = level_vertex_weights<T>(rp, ts, ps, re, CREASED);
        w_store(a, r, w.a, w.b, w.g, w.dl);
    }
}

// pass 2: edge(h) for its owner h (the larger halfedge, or a boundary one):
// the b shares of the halfedges on the edge and the dl shares of the
// halfedges whose previous edge it is
template <typename T, int CREASED>
__device__ __forceinline__ void adj_edge_at(const AdjArgs<T> &a, int32_t h) {
    if (h >= a.H) return;
    const He x = a.he[h];
    if (x.twin >= 0 && h < x.twin) return;
    const V3<T> *gV = reinterpret_cast<const V3<T> *>(a.gV);
    V3<T> *gE = reinterpret_cast<V3<T> *>(a.gE);
    V3<T> ge = gE[x.edge];
    const int32_t nh = qnext(h);
    const T bh = w_b(a, h), dn = CREASED ? w_dl(a, nh) : (T)0;   // smooth rules have no dl share
    const V3<T> vh = gV[x.vert], vn = gV[a.he[nh].vert];
    ge.x += bh * vh.x; ge.y += bh * vh.y; ge.z += bh * vh.z;
    ge.x += dn * vn.x; ge.y += dn * vn.y; ge.z += dn * vn.z;
    if (x.twin >= 0) {
        const int32_t t = x.twin, nt = qnext(t);
        const T bt = w_b(a, t), dt = CREASED ? w_dl(a, nt) : (T)0;
        const V3<T> vt = gV[a.he[t].vert], vnt = gV[a.he[nt].vert];
        ge.x += bt * vt.x; ge.y += bt * vt.y; ge.z += bt * vt.z;
        ge.x += dt * vnt.x; ge.y += dt * vnt.y; ge.z += dt * vnt.z;
    }
    gE[x.edge] = ge;
}

// pass 3: face f (halfedges 4f..4f+3): the g shares of the vertex rule and
// the cf shares of the edge rule
template <typename T, int CREASED>
__device__ __forceinline__ void adj_face_at(const AdjArgs<T> &a, int32_t f) {
    if (f >= (a.H >> 2)) return;
    const V3<T> *gV = reinterpret_cast<const V3<T> *>(a.gV);
    const V3<T> *gE = reinterpret_cast<const V3<T> *>(a.gE);
    V3<T> *gF = reinterpret_cast<V3<T> *>(a.gF);
    V3<T> gf = gF[f];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int32_t h = 4 * f + j;
        const He x = a.he[h];
        T c0, c1, cf;
        edge_weights<T, CREASED>(a, x.edge, x.twin < 0, c0, c1, cf);
        const T g = w_g(a, h);
        const V3<T> v = gV[x.vert], e = gE[x.edge];
        gf.x += g * v.x; gf.y += g * v.y; gf.z += g * v.z;
        gf.x += cf * e.x; gf.y += cf * e.y; gf.z += cf * e.z;
    }
    gF[f] = gf;
}

// pass 4: vertex v with fan f gathers a * gV[v] from the vertex rule, c0
// and c1 shares of its incident edges from the edge rule and a quarter of
// each incident face from the face rule
template <typename T, int CREASED>
__device__ __forceinline__ void adj_vertex_gather(const AdjArgs<T> &a, const DCage &c, const Fan &f, int32_t v) {
    const He *he = a.he;
    const V3<T> *gE = reinterpret_cast<const V3<T> *>(a.gE);
    const V3<T> *gF = reinterpret_cast<const V3<T> *>(a.gF);
    const V3<T> gv = reinterpret_cast<const V3<T> *>(a.gV)[v];
    T alpha = 0;
    V3<T> acc{0, 0, 0};
    if (f.n <= 4) {
        // a fan of at most four: every entry's loads issued before any is
        // used, the same sums in the same order as below
        int32_t r[4];
        He x[4], xp[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = fan_at(f, c, q < f.n ? q : 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            x[q] = he[r[q]];
            xp[q] = he[qprev(r[q])];
        }
        T wa[4], c0[4], d1[4];
        V3<T> e[4], ep[4], g[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            T c1, cf, d0, df;
            wa[q] = w_a(a, r[q]);
            edge_weights<T, CREASED>(a, x[q].edge, x[q].twin < 0, c0[q], c1, cf);
            edge_weights<T, CREASED>(a, xp[q].edge, xp[q].twin < 0, d0, d1[q], df);
            e[q] = gE[x[q].edge];
            ep[q] = gE[xp[q].edge];
            g[q] = gF[r[q] >> 2];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (i >= f.n) break;
            alpha += wa[i];
            acc.x += c0[i] * e[i].x; acc.y += c0[i] * e[i].y; acc.z += c0[i] * e[i].z;
            acc.x += d1[i] * ep[i].x; acc.y += d1[i] * ep[i].y; acc.z += d1[i] * ep[i].z;
            acc.x += (T)0.25 * g[i].x; acc.y += (T)0.25 * g[i].y; acc.z += (T)0.25 * g[i].z;
        }
        reinterpret_cast<V3<T> *>(a.gP)[v] = {alpha * gv.x + acc.x, alpha * gv.y + acc.y, alpha * gv.z + acc.z};
        return;
    }
    for (int i = 0; i < f.n; ++i) {
        const int32_t r = fan_at(f, c, i);
        const int32_t p = qprev(r);
        const He x = he[r], xp = he[p];
        alpha += w_a(a, r);
        T c0, c1, cf, d0, d1, df;
        edge_weights<T, CREASED>(a, x.edge, x.twin < 0, c0, c1, cf);     // r leaves v
        edge_weights<T, CREASED>(a, xp.edge, xp.twin < 0, d0, d1, df);  // prev(r) enters v
        const V3<T> e = gE[x.edge], ep = gE[xp.edge], g = gF[r >> 2];
        acc.x += c0 * e.x; acc.y += c0 * e.y; acc.z += c0 * e.z;
        acc.x += d1 * ep.x; acc.y += d1 * ep.y; acc.z += d1 * ep.z;
        acc.x += (T)0.25 * g.x; acc.y += (T)0.25 * g.y; acc.z += (T)0.25 * g.z;
    }
    reinterpret_cast<V3<T> *>(a.gP)[v] = {alpha * gv.x + acc.x, alpha * gv.y + acc.y, alpha * gv.z + acc.z};
}

// Passes 1 and 4 for the vertex v with fan f: for face and interior edge
// vertices born at level >= 2 the ring comes from the child numbering
// (structural_ring), so no ring record is read; the other fans take the
// general walks above.  The sums run over the same ring in the same order as
// the forward rule.
template <typename T, int CREASED>
__device__ __forceinline__ void adj_weights_fan(const AdjArgs<T> &a, const DCage &c, const LevelTab &t, int d,
                                                const Fan &f, int32_t v) {
    int32_t h[4], e[4];
    if (!structural_ring(f, t, d, v, h, e)) {
        adj_ring_weights<T, CREASED>(a, c, f);
        return;
    }
    T s[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) s[q] = CREASED ? sharp_at(a, e[q]) : (T)0;
    RingT<T> rp;
    rp.n = 4; rp.cc = 0; rp.boundary = false; rp.s1 = rp.s2 = 0; rp.e1 = rp.e2 = -1; rp.sall = 0;
    if (CREASED) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {   // the previous edge of h[i] is e[i+1]
            const T se = s[(i + 1) & 3];
            rp.cc += sgnT(se);
            if (se != (T)0) {
                if (rp.e1 < 0) { rp.s1 = se; rp.e1 = e[(i + 1) & 3]; }
                else if (rp.e2 < 0) { rp.s2 = se; rp.e2 = e[(i + 1) & 3]; }
            }
            rp.sall += se;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const VW<T> w = level_vertex_weights<T>(rp, s[i], s[(i + 1) & 3], e[i], CREASED);
        w_store(a, h[i], w.a, w.b, w.g, w.dl);
    }
}

template <typename T, int CREASED>
__device__ __forceinline__ void adj_gather_fan(const AdjArgs<T> &a, const DCage &c, const LevelTab &t, int d,
                                               const Fan &f, int32_t v) {
    int32_t h[4], e[4];
    if (!structural_ring(f, t, d, v, h, e)) {
        adj_vertex_gather<T, CREASED>(a, c, f, v);
        return;
    }
    const V3<T> *gE = reinterpret_cast<const V3<T> *>(a.gE);
    const V3<T> *gF = reinterpret_cast<const V3<T> *>(a.gF);
    const V3<T> gv = reinterpret_cast<const V3<T> *>(a.gV)[v];
    T alpha = 0;
    V3<T> acc{0, 0, 0};
#pragma unroll
    for (int i = 0; i < 4; ++i) {   // interior ring: no boundary edge
        const int32_t r = h[i], re = e[i], pe = e[(i + 1) & 3];
        alpha += w_a(a, r);
        T c0, c1, cf, d0, d1, df;
        edge_weights<T, CREASED>(a, re, false, c0, c1, cf);
        edge_weights<T, CREASED>(a, pe, false, d0, d1, df);
        const V3<T> ev = gE[re], ep = gE[pe], g = gF[r >> 2];
        acc.x += c0 * ev.x; acc.y += c0 * ev.y; acc.z += c0 * ev.z;
        acc.x += d1 * ep.x; acc.y += d1 * ep.y; acc.z += d1 * ep.z;
        acc.x += (T)0.25 * g.x; acc.y += (T)0.25 * g.y; acc.z += (T)0.25 * g.z;
    }
    reinterpret_cast<V3<T> *>(a.gP)[v] = {alpha * gv.x + acc.x, alpha * gv.y + acc.y, alpha * gv.z + acc.z};
}

// The owners of level-d vertices among 256 * SCAN halfedges (structural
// test, no scan of the records), collected in LDS so that every thread then
// works on one, in halfedge order (neighbouring rings, shared cache lines).
// The level tables go to LDS first; the first barrier orders the copy.
template <int SCAN, class Fn, class LP>
__device__ __forceinline__ void owners_block(int32_t H, int d, const DCage &c, const LevelTab &t, const LP &lp,
                                             uint32_t lb, const uint64_t *erep, const Fn &fn) {
    __shared__ int32_t owners[256 * SCAN];
    __shared__ int32_t nown;
    __shared__ LevelTab st;
    __shared__ LevelPtrs slp;
    lds_fill(st, t);
    lds_fill(slp, lp);
    const int32_t base = (int32_t)(lb * 256 * SCAN);
    if (threadIdx.x == 0) nown = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN; ++k) {
        const int32_t h = base + k * 256 + (int32_t)threadIdx.x;
        if (h < H && structural_owner(h, d, c, st, slp.he, erep)) owners[atomicAdd(&nown, 1)] = h;
    }
    __syncthreads();
    const int32_t count = nown;
    for (int32_t i = (int32_t)threadIdx.x; i < count; i += 256) {
        Fan f;
        const int32_t v = marked_fan(owners[i], d, c, st, slp.he, f);
        fn(st, f, v);
    }
}

constexpr int kAdjScan = 4;

// Passes 1 and 4 with one thread per vertex and the fan from the vertex ID
// (vertex_fan): no owner scan and one round of dependent loads fewer, but
// consecutive vertex IDs of split edges and older vertices have rings far
// apart.  Faster while a level has few vertices (latency-bound launches),
// slower on large levels (AdjPlan::per_vertex_max chooses).

template <typename T, int CREASED, int GATHER_PASS>
__device__ __forceinline__ void adj_vertex_pass(const AdjArgs<T> &a, const DCage &c, const LevelTab &t,
                                                const LevelPtrs &lp, int d, int32_t v) {
    __shared__ LevelTab st;
    __shared__ LevelPtrs slp;
    lds_fill(st, t);
    lds_fill(slp, lp);
    __syncthreads();
    if (v >= a.V) return;
    Fan f;
    vertex_fan(v, d, c, st, slp.he, f);
    if (GATHER_PASS) adj_gather_fan<T, CREASED>(a, c, st, d, f, v);
    else adj_weights_fan<T, CREASED>(a, c, st, d, f, v);
}

template <typename T, int CREASED>
struct RingWeightsFn {
    const AdjArgs<T> &a;
    const DCage &c;
    int d;
    __device__ void operator()(const LevelTab &t, const Fan &f, int32_t v) const {
        adj_weights_fan<T, CREASED>(a, c, t, d, f, v);
    }
};

template <typename T, int CREASED>
struct VertexGatherFn {
    const AdjArgs<T> &a;
    const DCage &c;
    int d;
    __device__ void operator()(const LevelTab &t, const Fan &f, int32_t v) const {
        adj_gather_fan<T, CREASED>(a, c, t, d, f, v);
    }
};

// ---- the cage step (level 0 -> 1) -----------------------------------------------
// vertex rule, one thread per cage halfedge with the reference's own walk
// (creased :1884-1988, whose smooth test comes first; smooth :1724-1753)
template <typename T, int CREASED>
__device__ __forceinline__ void adj_cage_vertices_at(const AdjArgs<T> &a, const DCage &c, int32_t h) {
    if (h >= c.H) return;
    const CageHe x = c.he[h];
    const V3<T> gv = reinterpret_cast<const V3<T> *>(a.gV)[x.vert];
    V3<T> *gE = reinterpret_cast<V3<T> *>(a.gE);
    V3<T> *gF = reinterpret_cast<V3<T> *>(a.gF);
    V3<T> *gP = reinterpret_cast<V3<T> *>(a.gP);
    if (!CREASED) {
        int32_t n = 1, fw, bw;
        for (fw = c.he[x.prev].twin; fw >= 0 && fw != h && n < kWalkCap; fw = c.he[c.he[fw].prev].twin) ++n;
        for (bw = x.twin >= 0 ? c.he[x.twin].next : -1; fw < 0 && bw >= 0 && bw != h && n < kWalkCap;
             bw = c.he[bw].twin >= 0 ? c.he[c.he[bw].twin].next : -1)
            ++n;
        const T iw = (T)1 / (T)n, s = fw < 0 ? (T)0 : (T)1;
        atomic_add3(&gP[x.vert], iw - (T)3 * iw * iw * s, gv);
        if (s != (T)0) {
            atomic_add3(&gE[x.edge], (T)4 * iw * iw * s, gv);
            atomic_add3(&gF[x.face], -iw * iw * s, gv);
        }
        return;
    }
    const int32_t p = x.prev;
    const T ts = (T)c.cr[x.edge].sharp, ps = (T)c.cr[c.he[p].edge].sharp;
    T avgS = ps, cc = sgnT(ps), n = 1;
    int32_t fw, bw;
    for (fw = c.he[p].twin; fw >= 0 && fw != h && n < (T)kWalkCap; fw = c.he[fw].twin) {
        const int32_t q = c.he[fw].prev;
        const T qs = (T)c.cr[c.he[q].edge].sharp;
        n = n + (T)1;
        avgS = avgS + qs;
        cc = cc + sgnT(qs);
        fw = q;
    }
    for (bw = x.twin; fw < 0 && bw >= 0 && bw != h && n < (T)kWalkCap; bw = c.he[bw].twin) {
        const int32_t q = c.he[bw].next;
        const T qs = (T)c.cr[c.he[q].edge].sharp;
        n = n + (T)1;
        avgS = avgS + qs;
        cc = cc + sgnT(qs);
        bw = q;
    }
    const bool boundary = fw < 0;
    const T nn = n;
    const T ccAll = boundary ? cc + sgnT(ts) : cc;
    const T nPost = boundary ? nn + (T)1 : nn;
    if (ccAll <= (T)1) {
        atomic_add3(&gP[x.vert], (nn - (T)3) / (nn * nn), gv);
        atomic_add3(&gE[x.edge], (T)4 / (nn * nn), gv);
        atomic_add3(&gF[x.face], (T)-1 / (nn * nn), gv);
    } else if (ccAll >= (T)3 || nPost == (T)2) {
        atomic_add3(&gP[x.vert], (T)1 / nn, gv);
    } else {
        const T tt = satT(avgS * (T)0.5);
        const T cw = sgnT(ts), pcw = boundary ? sgnT(ps) : (T)0;
        atomic_add3(&gP[x.vert], ((T)1 - tt) / nn + tt * (T)0.25 * (cw + pcw), gv);
        if (cw != (T)0) atomic_add3(&gE[x.edge], tt * (T)0.25 * cw, gv);
        if (pcw != (T)0) atomic_add3(&gE[c.he[p].edge], tt * (T)0.25 * pcw, gv);
    }
}

// edge rule (creased :1612-1645, smooth :1527-1545)
template <typename T, int CREASED>
__device__ __forceinline__ void adj_cage_edges_at(const AdjArgs<T> &a, const DCage &c, int32_t h) {
    if (h >= c.H) return;
    const CageHe x = c.he[h];
    const V3<T> ge = reinterpret_cast<const V3<T> *>(a.gE)[x.edge];
    const bool boundary = x.twin < 0;
    T c0, c1, cf;
    if (CREASED) {
        const T w = satT((T)c.cr[x.edge].sharp);
        const T k = boundary ? (T)1 : (T)0.5;
        c0 = (T)0.25 * ((T)1 - w) + (T)0.5 * k * w;
        c1 = (T)0.5 * k * w;
        cf = (T)0.25 * ((T)1 - w);
    } else {
        const T w = boundary ? (T)1 : (T)0.5;
        c0 = c1 = (T)0.5 * w * w;
        cf = (T)0.5 * ((T)1 - w);
    }
    V3<T> *gP = reinterpret_cast<V3<T> *>(a.gP);
    atomic_add3(&gP[x.vert], c0, ge);
    atomic_add3(&gP[c.he[x.next].vert], c1, ge);
    if (cf != (T)0) atomic_add3(reinterpret_cast<V3<T> *>(a.gF) + x.face, cf, ge);
}

// face rule (:1444-1460): each corner receives 1/m of its face's gradient
template <typename T>
__device__ __forceinline__ void adj_cage_faces_at(const AdjArgs<T> &a, const DCage &c, int32_t h) {
    if (h >= c.H) return;
    const CageHe x = c.he[h];
    int32_t m = 1;
    for (int32_t it = x.next; it != h && m < kWalkCap; it = c.he[it].next) ++m;
    const V3<T> gf = reinterpret_cast<const V3<T> *>(a.gF)[x.face];
    atomic_add3(reinterpret_cast<V3<T> *>(a.gP) + x.vert, (T)1 / (T)m, gf);
}

// ---- the cage step without atomics -------------------------------------------------
// The same transposed rules as above, in four gathers with a fixed order of
// additions (so the adjoint is reproducible bit for bit): (1) per halfedge
// the vertex rule's weights (to its vertex, edge, face and previous edge);
// (2) per edge, its two halfedges' shares; (3) per face, the vertex and edge
// rules' shares; (4) per vertex, over its outgoing halfedges in ascending
// order, the vertex, edge and face rules' shares.  Needs the cage's lists.

// the edge rule's weights of cage halfedge x (c0: its vertex, c1: the next
// vertex, cf: its face)
template <typename T, int CREASED>
__device__ __forceinline__ void cage_edge_weights(const DCage &c, const CageHe &x, T &c0, T &c1, T &cf) {
    const bool boundary = x.twin < 0;
    if (CREASED) {
        const T w = satT((T)c.cr[x.edge].sharp);
        const T k = boundary ? (T)1 : (T)0.5;
        c0 = (T)0.25 * ((T)1 - w) + (T)0.5 * k * w;
        c1 = (T)0.5 * k * w;
        cf = (T)0.25 * ((T)1 - w);
    } else {
        const T w = boundary ? (T)1 : (T)0.5;
        c0 = c1 = (T)0.5 * w * w;
        cf = (T)0.5 * ((T)1 - w);
    }
}

// (1) W[h] = (to its vertex, to its edge, to its face, to its previous edge);
// 0 marks a share the rule does not add
template <typename T, int CREASED>
__device__ __forceinline__ void adj_cage_weights_at(const AdjArgs<T> &a, const DCage &c, int32_t h) {
    if (h >= c.H) return;
    const CageHe x = c.he[h];
    W4<T> w{0, 0, 0, 0};
    if (!CREASED) {
        int32_t n = 1, fw, bw;
        for (fw = c.he[x.prev].twin; fw >= 0 && fw != h && n < kWalkCap; fw = c.he[c.he[fw].prev].twin) ++n;
        for (bw = x.twin >= 0 ? c.he[x.twin].next : -1; fw < 0 && bw >= 0 && bw != h && n < kWalkCap;
             bw = c.he[bw].twin >= 0 ? c.he[c.he[bw].twin].next : -1)
            ++n;
        const T iw = (T)1 / (T)n, s = fw < 0 ? (T)0 : (T)1;
        w.a = iw - (T)3 * iw * iw * s;
        if (s != (T)0) {
            w.b = (T)4 * iw * iw * s;
            w.g = -iw * iw * s;
        }
    } else {
        const int32_t p = x.prev;
        const T ts = (T)c.cr[x.edge].sharp, ps = (T)c.cr[c.he[p].edge].sharp;
        T avgS = ps, cc = sgnT(ps), n = 1;
        int32_t fw, bw;
        for (fw = c.he[p].twin; fw >= 0 && fw != h && n < (T)kWalkCap; fw = c.he[fw].twin) {
            const int32_t q = c.he[fw].prev;
            const T qs = (T)c.cr[c.he[q].edge].sharp;
            n = n + (T)1;
            avgS = avgS + qs;
            cc = cc + sgnT(qs);
            fw = q;
        }
        for (bw = x.twin; fw < 0 && bw >= 0 && bw != h && n < (T)kWalkCap; bw = c.he[bw].twin) {
            const int32_t q = c.he[bw].next;
            const T qs = (T)c.cr[c.he[q].edge].sharp;
            n = n + (T)1;
            avgS = avgS + qs;
            cc = cc + sgnT(qs);
            bw = q;
        }
        const bool boundary = fw < 0;
        const T nn = n;
        const T ccAll = boundary ? cc + sgnT(ts) : cc;
        const T nPost = boundary ? nn + (T)1 : nn;
        if (ccAll <= (T)1) {
            w.a = (nn - (T)3) / (nn * nn);
            w.b = (T)4 / (nn * nn);
            w.g = (T)-1 / (nn * nn);
        } else if (ccAll >= (T)3 || nPost == (T)2) {
            w.a = (T)1 / nn;
        } else {
            const T tt = satT(avgS * (T)0.5);
            const T cw = sgnT(ts), pcw = boundary ? sgnT(ps) : (T)0;
            w.a = ((T)1 - tt) / nn + tt * (T)0.25 * (cw + pcw);
            if (cw != (T)0) w.b = tt * (T)0.25 * cw;
            if (pcw != (T)0) w.dl = tt * (T)0.25 * pcw;
        }
    }
    w_store(a, h, w.a, w.b, w.g, w.dl);
}

template <typename T>
__device__ __forceinline__ void add_scaled(V3<T> &acc, T w, const V3<T> &g) {
    if (w == (T)0) return;
    acc.x += w * g.x; acc.y += w * g.y; acc.z += w * g.z;
}

// (2) edge e: the vertex rule's shares of its halfedges (ascending) and of
// the halfedges whose previous edge it is
template <typename T>
__device__ __forceinline__ void adj_cage_edge_gather(const AdjArgs<T> &a, const DCage &c, int32_t e) {
    if (e >= c.E) return;
    const V3<T> *gV = reinterpret_cast<const V3<T> *>(a.gV);
    V3<T> *gE = reinterpret_cast<V3<T> *>(a.gE);
    const int32_t h0 = c.e2h[e], t0 = c.he[h0].twin;
    int32_t hs[2] = {h0, t0};
    int k = 1;
    if (t0 >= 0) {
        k = 2;
        if (t0 < h0) { hs[0] = t0; hs[1] = h0; }
    }
    V3<T> acc = gE[e];
    for (int i = 0; i < k; ++i) add_scaled(acc, w_b(a, hs[i]), gV[c.he[hs[i]].vert]);
    for (int i = 0; i < k; ++i) {
        const int32_t nx = c.he[hs[i]].next;
        add_scaled(acc, w_dl(a, nx), gV[c.he[nx].vert]);
    }
    gE[e] = acc;
}

// (3) face f: the vertex rule's shares of its halfedges, then the edge
// rule's face shares (in its halfedge order from f2h)
template <typename T, int CREASED>
__device__ __forceinline__ void adj_cage_face_gather(const AdjArgs<T> &a, const DCage &c, int32_t f) {
    if (f >= c.F) return;
    const V3<T> *gV = reinterpret_cast<const V3<T> *>(a.gV);
    const V3<T> *gE = reinterpret_cast<const V3<T> *>(a.gE);
    V3<T> *gF = reinterpret_cast<V3<T> *>(a.gF);
    const int32_t o = c.fring_off[f], m = c.fring_off[f + 1] - o;
    V3<T> acc = gF[f];
    for (int i = 0; i < m; ++i) {
        const int32_t h = c.fring[o + i];
        add_scaled(acc, w_g(a, h), gV[c.he[h].vert]);
    }
    for (int i = 0; i < m; ++i) {
        const CageHe x = c.he[c.fring[o + i]];
        T c0, c1, cf;
        cage_edge_weights<T, CREASED>(c, x, c0, c1, cf);
        add_scaled(acc, cf, gE[x.edge]);
    }
    gF[f] = acc;
}

// (4) vertex v: over its outgoing halfedges h, ascending: the vertex rule's
// own share, the edge rule's c0 of h and c1 of prev(h), and 1/m of h's face
template <typename T, int CREASED>
__device__ __forceinline__ void adj_cage_vertex_gather(const AdjArgs<T> &a, const DCage &c, int32_t v) {
    if (v >= c.V) return;
    const V3<T> *gE = reinterpret_cast<const V3<T> *>(a.gE);
    const V3<T> *gF = reinterpret_cast<const V3<T> *>(a.gF);
    const V3<T> gv = reinterpret_cast<const V3<T> *>(a.gV)[v];
    V3<T> acc{0, 0, 0};
    for (int32_t i = c.vout_off[v]; i < c.vout_off[v + 1]; ++i) {
        const int32_t h = c.vout[i];
        const CageHe x = c.he[h], xp = c.he[x.prev];
        T c0, c1, cf, d0, d1, df;
        cage_edge_weights<T, CREASED>(c, x, c0, c1, cf);
        cage_edge_weights<T, CREASED>(c, xp, d0, d1, df);
        add_scaled(acc, w_a(a, h), gv);
        add_scaled(acc, c0, gE[x.edge]);
        add_scaled(acc, d1, gE[xp.edge]);
        const int32_t m = c.fring_off[x.face + 1] - c.fring_off[x.face];
        add_scaled(acc, (T)1 / (T)m, gF[x.face]);
    }
    reinterpret_cast<V3<T> *>(a.gP)[v] = acc;
}

// ---- kernels: one thread per halfedge ------------------------------------------
template <typename T, int CREASED>
__global__ __launch_bounds__(256) void k_adj_weights(AdjArgs<T> a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    owners_block<kAdjScan>(a.H, d, c, t, lp, blockIdx.x, a.erep, RingWeightsFn<T, CREASED>{a, c, d});
}

template <typename T, int CREASED>
__global__ __launch_bounds__(256) void k_adj_edges(AdjArgs<T> a) {
    adj_edge_at<T, CREASED>(a, (int32_t)(blockIdx.x * blockDim.x + threadIdx.x));
}

template <typename T, int CREASED>
__global__ __launch_bounds__(256) void k_adj_faces(AdjArgs<T> a) {
    adj_face_at<T, CREASED>(a, (int32_t)(blockIdx.x * blockDim.x + threadIdx.x));
}

template <typename T, int CREASED>
__global__ __launch_bounds__(256) void k_adj_vertices(AdjArgs<T> a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    owners_block<kAdjScan>(a.H, d, c, t, lp, blockIdx.x, a.erep, VertexGatherFn<T, CREASED>{a, c, d});
}

template <typename T, int CREASED, int GATHER_PASS>
__global__ __launch_bounds__(256) void k_adj_per_vertex(AdjArgs<T> a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    adj_vertex_pass<T, CREASED, GATHER_PASS>(a, c, t, lp, d, (int32_t)(blockIdx.x * 256 + threadIdx.x));
}

template <typename T, int CREASED>
__global__ __launch_bounds__(256) void k_adj_cage_vertices(AdjArgs<T> a, DCage c) {
    adj_cage_vertices_at<T, CREASED>(a, c, (int32_t)(blockIdx.x * blockDim.x + threadIdx.x));
}

template <typename T, int CREASED>
__global__ __launch_bounds__(256) void k_adj_cage_edges(AdjArgs<T> a, DCage c) {
    adj_cage_edges_at<T, CREASED>(a, c, (int32_t)(blockIdx.x * blockDim.x + threadIdx.x));
}

template <typename T>
__global__ __launch_bounds__(256) void k_adj_cage_faces(AdjArgs<T> a, DCage c) {
    adj_cage_faces_at<T>(a, c, (int32_t)(blockIdx.x * blockDim.x + threadIdx.x));
}

template <typename T, int CREASED>
__global__ __launch_bounds__(256) void k_adj_cage_weights(AdjArgs<T> a, DCage c) {
    adj_cage_weights_at<T, CREASED>(a, c, (int32_t)(blockIdx.x * blockDim.x + threadIdx.x));
}

template <typename T>
__global__ __launch_bounds__(256) void k_adj_cage_edge_gather(AdjArgs<T> a, DCage c) {
    adj_cage_edge_gather<T>(a, c, (int32_t)(blockIdx.x * blockDim.x + threadIdx.x));
}

template <typename T, int CREASED>
__global__ __launch_bounds__(256) void k_adj_cage_face_gather(AdjArgs<T> a, DCage c) {
    adj_cage_face_gather<T, CREASED>(a, c, (int32_t)(blockIdx.x * blockDim.x + threadIdx.x));
}

template <typename T, int CREASED>
__global__ __launch_bounds__(256) void k_adj_cage_vertex_gather(AdjArgs<T> a, DCage c) {
    adj_cage_vertex_gather<T, CREASED>(a, c, (int32_t)(blockIdx.x * blockDim.x + threadIdx.x));
}

// crease sharpness of the next level in T (CatmullClark.h:2962-2980), for
// the double-precision adjoint; links come from the float crease records
template <typename T>
__global__ void k_sharp_next(const Crease *__restrict__ links, const T *__restrict__ sin, int32_t C,
                             T *__restrict__ sout) {
    const int32_t e = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= C) return;
    const Crease x = links[e];
    const T s3 = (T)3 * sin[e];
    const T ns = sin[x.next], ps = sin[x.prev];
    const T c0 = (ps + s3) / (T)4 - (T)1, c1 = (s3 + ns) / (T)4 - (T)1;
    sout[2 * (int64_t)e + 0] = c0 > (T)0 ? c0 : (T)0;
    sout[2 * (int64_t)e + 1] = c1 > (T)0 ? c1 : (T)0;
}

template <typename T>
__global__ void k_sharp_cage(const Crease *__restrict__ cr, int32_t E, T *__restrict__ s) {
    const int32_t e = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e < E) s[e] = (T)cr[e].sharp;
}

template <typename T>
__global__ void k_copy(const T *__restrict__ src, T *__restrict__ dst, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

template <typename T>
__global__ void k_fill0(T *dst, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = (T)0;
}

// ---- driver -----------------------------------------------------------------------
namespace {
unsigned grid_for(int64_t n) { return (unsigned)((n + 255) / 256); }
unsigned grid_stride(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 65536 ? b : 65536);
}
}  // namespace

template <typename T>
void adjoint_levels(const AdjPlan &pl, const T *gradFinest, T *gradCage, T *bufA, T *bufB, T *W, T *sharp,
                    int creased, hipStream_t st) {
    const Counts &n = *pl.n;
    const int D = pl.D;
    // sharpness of levels 0..D-1 in T when a buffer is given (cumulative,
    // level d at offset offC[d] + E0 with level 0 first)
    if (sharp && creased) {
        hipLaunchKernelGGL(k_sharp_cage<T>, dim3(grid_for(pl.cage.E)), dim3(256), 0, st, pl.cage.cr, pl.cage.E, sharp);
        for (int d = 0; d + 1 < D; ++d) {
            const Crease *links = d == 0 ? pl.cage.cr : pl.cr + n.offC[d];
            const T *sin = sharp + (d == 0 ? 0 : pl.cage.E + n.offC[d]);
            hipLaunchKernelGGL(k_sharp_next<T>, dim3(grid_for(n.C[d])), dim3(256), 0, st, links, sin, (int32_t)n.C[d],
                               sharp + pl.cage.E + n.offC[d + 1]);
        }
    }
    // G holds the gradient of level d+1 (modified in place), P receives level d
    T *G = bufA, *P = bufB;
    hipLaunchKernelGGL(k_copy<T>, dim3(grid_stride(3 * n.V[D])), dim3(256), 0, st, gradFinest, G, 3 * n.V[D]);
    for (int d = D - 1; d >= 1; --d) {
        AdjArgs<T> a;
        a.he = pl.he + n.offH[d];
        a.cr = pl.cr + n.offC[d];
        a.H = (int32_t)n.H[d];
        a.C = (int32_t)n.C[d];
        a.V = (int32_t)n.V[d];
        a.sd = (sharp && creased) ? sharp + pl.cage.E + n.offC[d] : nullptr;
        a.gV = G;
        a.gF = G + 3 * n.V[d];
        a.gE = G + 3 * (n.V[d] + n.F[d]);
        a.gP = P;
        a.W = W;
        a.erep = pl.erep;
        const dim3 g(grid_for(n.H[d])), gq(grid_for(n.H[d] / 4)), gv(grid_for(n.V[d])), b(256);
        const bool pv = n.V[d] <= pl.per_vertex_max;
        if (creased) {
            if (pv) hipLaunchKernelGGL((k_adj_per_vertex<T, 1, 0>), gv, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            else hipLaunchKernelGGL((k_adj_weights<T, 1>), gq, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            hipLaunchKernelGGL((k_adj_edges<T, 1>), g, b, 0, st, a);
            hipLaunchKernelGGL((k_adj_faces<T, 1>), gq, b, 0, st, a);
            if (pv) hipLaunchKernelGGL((k_adj_per_vertex<T, 1, 1>), gv, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            else hipLaunchKernelGGL((k_adj_vertices<T, 1>), gq, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
        } else {
            if (pv) hipLaunchKernelGGL((k_adj_per_vertex<T, 0, 0>), gv, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            else hipLaunchKernelGGL((k_adj_weights<T, 0>), gq, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            hipLaunchKernelGGL((k_adj_edges<T, 0>), g, b, 0, st, a);
            hipLaunchKernelGGL((k_adj_faces<T, 0>), gq, b, 0, st, a);
            if (pv) hipLaunchKernelGGL((k_adj_per_vertex<T, 0, 1>), gv, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            else hipLaunchKernelGGL((k_adj_vertices<T, 0>), gq, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
        }
        T *tmp = G; G = P; P = tmp;
    }
    AdjArgs<T> a;
    a.he = nullptr;
    a.cr = nullptr;
    a.H = pl.cage.H;
    a.C = pl.cage.E;
    a.V = pl.cage.V;
    a.sd = nullptr;
    a.gV = G;
    a.gF = G + 3 * (int64_t)pl.cage.V;
    a.gE = G + 3 * ((int64_t)pl.cage.V + pl.cage.F);
    a.gP = gradCage;
    a.W = W;
    const dim3 g(grid_for(pl.cage.H)), b(256);
    if (pl.cage.vout) {   // gathers in a fixed order
        const dim3 ge(grid_for(pl.cage.E)), gf(grid_for(pl.cage.F)), gv(grid_for(pl.cage.V));
        if (creased) {
            hipLaunchKernelGGL((k_adj_cage_weights<T, 1>), g, b, 0, st, a, pl.cage);
            hipLaunchKernelGGL(k_adj_cage_edge_gather<T>, ge, b, 0, st, a, pl.cage);
            hipLaunchKernelGGL((k_adj_cage_face_gather<T, 1>), gf, b, 0, st, a, pl.cage);
            hipLaunchKernelGGL((k_adj_cage_vertex_gather<T, 1>), gv, b, 0, st, a, pl.cage);
        } else {
            hipLaunchKernelGGL((k_adj_cage_weights<T, 0>), g, b, 0, st, a, pl.cage);
            hipLaunchKernelGGL(k_adj_cage_edge_gather<T>, ge, b, 0, st, a, pl.cage);
            hipLaunchKernelGGL((k_adj_cage_face_gather<T, 0>), gf, b, 0, st, a, pl.cage);
            hipLaunchKernelGGL((k_adj_cage_vertex_gather<T, 0>), gv, b, 0, st, a, pl.cage);
        }
        return;
    }
    a.W = nullptr;
    hipLaunchKernelGGL(k_fill0<T>, dim3(grid_stride(3 * (int64_t)pl.cage.V)), dim3(256), 0, st, gradCage,
                       3 * (int64_t)pl.cage.V);
    if (creased) {
        hipLaunchKernelGGL((k_adj_cage_vertices<T, 1>), g, b, 0, st, a, pl.cage);
        hipLaunchKernelGGL((k_adj_cage_edges<T, 1>), g, b, 0, st, a, pl.cage);
    } else {
        hipLaunchKernelGGL((k_adj_cage_vertices<T, 0>), g, b, 0, st, a, pl.cage);
        hipLaunchKernelGGL((k_adj_cage_edges<T, 0>), g, b, 0, st, a, pl.cage);
    }
    hipLaunchKernelGGL(k_adj_cage_faces<T>, g, b, 0, st, a, pl.cage);
}

// ---- batched (float): one launch per rule and level over every cage ------------
namespace {

template <int CREASED>
struct OpAdjWeights {
    int d;
    char *arena, *adj;
    __device__ void operator()(const AdjJob &J, int64_t i) const {
        AdjArgs<float> a;
        DCage c;
        unpack_adj(J, arena, adj, a, c);
        const JobLevelPtrs lp{J.lp, arena};
        owners_block<kAdjScan>(a.H, d, c, J.t, lp, (uint32_t)(i >> 8), a.erep, RingWeightsFn<float, CREASED>{a, c, d});
    }
};
template <int CREASED>
struct OpAdjEdges {
    char *arena, *adj;
    __device__ void operator()(const AdjJob &J, int64_t i) const {
        AdjArgs<float> a;
        DCage c;
        unpack_adj(J, arena, adj, a, c);
        adj_edge_at<float, CREASED>(a, (int32_t)i);
    }
};
template <int CREASED>
struct OpAdjFaces {
    char *arena, *adj;
    __device__ void operator()(const AdjJob &J, int64_t i) const {
        AdjArgs<float> a;
        DCage c;
        unpack_adj(J, arena, adj, a, c);
        adj_face_at<float, CREASED>(a, (int32_t)i);
    }
};
template <int CREASED>
struct OpAdjVertices {
    int d;
    char *arena, *adj;
    __device__ void operator()(const AdjJob &J, int64_t i) const {
        AdjArgs<float> a;
        DCage c;
        unpack_adj(J, arena, adj, a, c);
        const JobLevelPtrs lp{J.lp, arena};
        owners_block<kAdjScan>(a.H, d, c, J.t, lp, (uint32_t)(i >> 8), a.erep, VertexGatherFn<float, CREASED>{a, c, d});
    }
};
template <int CREASED>
struct OpAdjCageVertices {
    char *arena, *adj;
    __device__ void operator()(const AdjJob &J, int64_t i) const {
        AdjArgs<float> a;
        DCage c;
        unpack_adj(J, arena, adj, a, c);
        adj_cage_vertices_at<float, CREASED>(a, c, (int32_t)i);
    }
};
template <int CREASED>
struct OpAdjCageEdges {
    char *arena, *adj;
    __device__ void operator()(const AdjJob &J, int64_t i) const {
        AdjArgs<float> a;
        DCage c;
        unpack_adj(J, arena, adj, a, c);
        adj_cage_edges_at<float, CREASED>(a, c, (int32_t)i);
    }
};
struct OpAdjCageFaces {
    char *arena, *adj;
    __device__ void operator()(const AdjJob &J, int64_t i) const {
        AdjArgs<float> a;
        DCage c;
        unpack_adj(J, arena, adj, a, c);
        adj_cage_faces_at<float>(a, c, (int32_t)i);
    }
};
// the cage step as gathers; `pass` 0 weights (per halfedge), 1 edges,
// 2 faces, 3 vertices (each thread index below the element count)
template <int CREASED, int PASS>
struct OpAdjCageGather {
    char *arena, *adj;
    __device__ void operator()(const AdjJob &J, int64_t i) const {
        AdjArgs<float> a;
        DCage c;
        unpack_adj(J, arena, adj, a, c);
        if (PASS == 0) adj_cage_weights_at<float, CREASED>(a, c, (int32_t)i);
        else if (PASS == 1) adj_cage_edge_gather<float>(a, c, (int32_t)i);
        else if (PASS == 2) adj_cage_face_gather<float, CREASED>(a, c, (int32_t)i);
        else adj_cage_vertex_gather<float, CREASED>(a, c, (int32_t)i);
    }
};

}  // namespace

// level d of every job: weights, edges, faces, vertices in turn; `quads`
// has one workgroup per 1024 halfedges of each job, `halfedges` one per 256
void batch_adj_level(const BatchTable &quads, const BatchTable &halfedges, char *arena, char *adj, int d, int creased,
                     hipStream_t st) {
    if (creased) launch_batch<AdjJob>(quads, OpAdjWeights<1>{d, arena, adj}, st);
    else launch_batch<AdjJob>(quads, OpAdjWeights<0>{d, arena, adj}, st);
    if (creased) launch_batch<AdjJob>(halfedges, OpAdjEdges<1>{arena, adj}, st);
    else launch_batch<AdjJob>(halfedges, OpAdjEdges<0>{arena, adj}, st);
    if (creased) {
        launch_batch<AdjJob>(quads, OpAdjFaces<1>{arena, adj}, st);
        launch_batch<AdjJob>(quads, OpAdjVertices<1>{d, arena, adj}, st);
    } else {
        launch_batch<AdjJob>(quads, OpAdjFaces<0>{arena, adj}, st);
        launch_batch<AdjJob>(quads, OpAdjVertices<0>{d, arena, adj}, st);
    }
}

void batch_adj_cage_gather(const BatchTable &halfedges, const BatchTable &edges, const BatchTable &faces,
                           const BatchTable &vertices, char *arena, char *adj, int creased, hipStream_t st) {
    if (creased) {
        launch_batch<AdjJob>(halfedges, OpAdjCageGather<1, 0>{arena, adj}, st);
        launch_batch<AdjJob>(edges, OpAdjCageGather<1, 1>{arena, adj}, st);
        launch_batch<AdjJob>(faces, OpAdjCageGather<1, 2>{arena, adj}, st);
        launch_batch<AdjJob>(vertices, OpAdjCageGather<1, 3>{arena, adj}, st);
    } else {
        launch_batch<AdjJob>(halfedges, OpAdjCageGather<0, 0>{arena, adj}, st);
        launch_batch<AdjJob>(edges, OpAdjCageGather<0, 1>{arena, adj}, st);
        launch_batch<AdjJob>(faces, OpAdjCageGather<0, 2>{arena, adj}, st);
        launch_batch<AdjJob>(vertices, OpAdjCageGather<0, 3>{arena, adj}, st);
    }
}

void batch_adj_cage(const BatchTable &t, char *arena, char *adj, int creased, hipStream_t st) {
    if (creased) {
        launch_batch<AdjJob>(t, OpAdjCageVertices<1>{arena, adj}, st);
        launch_batch<AdjJob>(t, OpAdjCageEdges<1>{arena, adj}, st);
    } else {
        launch_batch<AdjJob>(t, OpAdjCageVertices<0>{arena, adj}, st);
        launch_batch<AdjJob>(t, OpAdjCageEdges<0>{arena, adj}, st);
    }
    launch_batch<AdjJob>(t, OpAdjCageFaces{arena, adj}, st);
}

template void adjoint_levels<float>(const AdjPlan &, const float *, float *, float *, float *, float *, float *,
                                    int, hipStream_t);
template void adjoint_levels<double>(const AdjPlan &, const double *, double *, double *, double *, double *,
                                     double *, int, hipStream_t);

}  // namespace cch
