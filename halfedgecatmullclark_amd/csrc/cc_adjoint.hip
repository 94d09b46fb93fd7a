// Adjoint of the point refinement.
//
// At fixed topology and sharpness every refined point is a fixed linear
// combination of the points one level up (the weights of the scatter rules,
// CatmullClark.h:1437-1991 and :2032-2615, depend only on valence, crease
// count, boundary and sharpness).  The adjoint applies the transposes level
// by level, finest first.  For a level step d -> d+1 (d >= 1) it runs in
// gather form, without atomics, in three passes:
//   P. parameters: the owner of each level-d vertex packs its vertex rule's
//      parameters (valence, boundary, rule, the first two sharp ring edges:
//      16 bytes, VParam) from its ring;
//   Q. quads: the four lanes of each quad form, from their corners'
//      parameters, the vertex rule's shares of each of the quad's edges and
//      the updated edge gradients (the larger halfedge of a pair stores its
//      edge's), then the updated face gradient;
//   R. vertices: the owner of each vertex gathers its level-d gradient from
//      the vertex rule (its own weight), its edges' and its faces' updated
//      gradients.
// The cage step (small, polygons of any size) runs as four ordered gathers
// over the host-built ring lists with per-halfedge weights (per-halfedge
// float atomics only for a cage whose walks do not close).  Only the stored
// topology and creases are read; nothing of the forward points is kept or
// recomputed.  Kernels are templated on the scalar type so the same code
// runs in float and, for gradient checks, in double.
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

}  // namespace

// ---- level d >= 1, gather form -------------------------------------------------
template <typename T>
struct W4 { T a, b, g, dl; };

// The cage step's weights live as four arrays of the cage's H scalars,
// a | b | g | dl, so that each pass reads only the one it needs.
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

// ring crease parameter update for the sharpness se of edge e, in ring
// order (the edge of the first halfedge on a boundary, then the previous
// edge of every fan halfedge)
template <typename T>
__device__ __forceinline__ void ring_add(RingT<T> &rp, T se, int32_t e) {
    rp.cc += sgnT(se);
    if (se != (T)0) {
        if (rp.e1 < 0) { rp.s1 = se; rp.e1 = e; }
        else if (rp.e2 < 0) { rp.s2 = se; rp.e2 = e; }
    }
    rp.sall += se;
}

template <typename T>
__device__ __forceinline__ RingT<T> ring_init(int n, bool boundary) {
    RingT<T> rp;
    rp.n = (T)n; rp.cc = 0; rp.boundary = boundary; rp.s1 = rp.s2 = 0; rp.e1 = rp.e2 = -1; rp.sall = 0;
    return rp;
}

// ring parameters of any fan from its records (valence, creases, boundary)
template <typename T, int CREASED>
__device__ __forceinline__ RingT<T> ring_params(const AdjArgs<T> &a, const DCage &c, const Fan &f) {
    const He *he = a.he;
    RingT<T> rp = ring_init<T>(f.n, f.boundary);
    if (!CREASED) return rp;
    if (f.n <= 4) {
        // a fan of at most four: every entry's loads issued before any is used
        int32_t r[4], re0 = 0, pe[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = fan_at(f, c, q < f.n ? q : 0);
        re0 = he[r[0]].edge;
#pragma unroll
        for (int q = 0; q < 4; ++q) pe[q] = he[qprev(r[q])].edge;
        T s0 = sharp_at(a, re0), ps[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) ps[q] = sharp_at(a, pe[q]);
        if (f.boundary) ring_add(rp, s0, re0);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < f.n) ring_add(rp, ps[q], pe[q]);
        return rp;
    }
    for (int i = (f.boundary ? -1 : 0); i < f.n; ++i) {
        const int32_t e = i < 0 ? he[fan_at(f, c, 0)].edge : he[qprev(fan_at(f, c, i))].edge;
        ring_add(rp, sharp_at(a, e), e);
    }
    return rp;
}

// ring parameters of a structural ring (four interior halfedges h[q] with
// edges e[q] and sharpness s[q]; the previous edge of h[q] is e[q+1])
template <typename T, int CREASED>
__device__ __forceinline__ RingT<T> ring_params4(const int32_t (&e)[4], const T (&s)[4]) {
    RingT<T> rp = ring_init<T>(4, false);
    if (CREASED) {
#pragma unroll
        for (int i = 0; i < 4; ++i) ring_add(rp, s[(i + 1) & 3], e[(i + 1) & 3]);
    }
    return rp;
}

// The packed parameters of a ring and the rule they select (level d >= 1
// scatter rule, CatmullClark.h:2568-2607, corner test first).
template <typename T>
__device__ __forceinline__ VParam<T> vparam_of(const RingT<T> &rp, bool creased) {
    int rule = 0;
    if (creased) {
        const T nn = rp.boundary ? rp.n + (T)1 : rp.n;
        rule = (rp.cc >= (T)3 || nn == (T)2) ? 1 : (rp.cc <= (T)1 ? 0 : 2);
    }
    VParam<T> p;
    p.nb = (int32_t)rp.n | (rule << 28) | (rp.boundary ? (1 << 30) : 0);
    p.e1 = rp.e1;
    const T n = rp.n;
    if (!creased) {   // the same operations as the rule's weights in vparam_weights
        const T iw = (T)1 / n, s = rp.boundary ? (T)0 : (T)1;
        p.s1 = iw - (T)3 * iw * iw * s;
        p.s2 = -iw * iw * s;
    } else if (rule == 0) {
        p.s1 = (n - (T)3) / (n * n);
        p.s2 = (T)-1 / (n * n);
    } else if (rule == 1) {
        p.s1 = (T)1 / n;
        p.s2 = (T)0;
    } else {
        p.s1 = rp.s1;
        p.s2 = rp.s2;
    }
    return p;
}

// weights (alpha, beta, gamma, delta) of one ring halfedge's contribution
// alpha*V + beta*E(its edge) + gamma*F(its face) + delta*E(its previous
// edge) to the vertex rule (creased :2568-2607; smooth :2339-2348), from
// the vertex's parameters and the halfedge's own (ts, edge) and previous
// (ps) sharpness
//
// The smooth and corner weights come stored in p (vparam_of): the smooth
// rule's are 1/n - 3/n^2, 4/n^2, -1/n^2 uncreased (times 0 but the first on
// a boundary) and (n - 3)/n^2, 4/n^2, -1/n^2 creased, the corner rule's 1/n.
// b = -4 g holds bit for bit: both are correctly rounded quotients (or
// products) and the factor 4 is exact, signed zeros included (a boundary's
// g is -0, its b +0).
template <typename T>
__device__ __forceinline__ VW<T> vparam_weights(const VParam<T> &p, T ts, T ps, int32_t edge, bool creased) {
    VW<T> w{0, 0, 0, 0};
    const int rule = creased ? (p.nb >> 28) & 3 : 0;
    if (rule == 0) {
        w.a = p.s1;
        w.b = (T)-4 * p.s2;
        w.g = p.s2;
        return w;
    }
    if (rule == 1) {
        w.a = p.s1;
        return w;
    }
    const T n = (T)(p.nb & 0x0FFFFFFF);
    const bool boundary = ((p.nb >> 30) & 1) != 0;
    // two sharp ring edges: the sum of the sharpness this halfedge's walk
    // visits, every ring edge but its own on a boundary
    const T avg = boundary ? (edge == p.e1 ? p.s2 : (ts != (T)0 ? p.s1 : p.s1 + p.s2)) : p.s1 + p.s2;
    const T t = satT(avg * (T)0.5);
    const T cw = sgnT(ts), pcw = boundary ? sgnT(ps) : (T)0;
    w.a = ((T)1 - t) / n + t * (T)0.25 * (cw + pcw);
    w.b = t * (T)0.25 * cw;
    w.dl = t * (T)0.25 * pcw;
    return w;
}

template <typename T>
__device__ __forceinline__ bool vparam_boundary_crease(const VParam<T> &p) {
    return ((p.nb >> 30) & 1) != 0 && ((p.nb >> 28) & 3) == 2;
}

// pass 1: the parameters of the vertex rule of vertex v with fan f.  Face
// and interior edge vertices born at level >= 2 take their ring from the
// child numbering (structural_ring), so no ring record is read.
template <typename T, int CREASED>
__device__ __forceinline__ void adj_params_fan(const AdjArgs<T> &a, const DCage &c, const LevelTab &t, int d,
                                               const Fan &f, int32_t v) {
    int32_t h[4], e[4];
    RingT<T> rp;
    if (structural_ring(f, t, d, v, h, e, c)) {
        T s[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] = CREASED ? sharp_at(a, e[q]) : (T)0;
        rp = ring_params4<T, CREASED>(e, s);
    } else {
        rp = ring_params<T, CREASED>(a, c, f);
    }
    a.VP[v] = vparam_of<T>(rp, CREASED);
}

// value held by the lane of the next halfedge of the same quad, for any
// dword-sized record (DPP quad_perm)
template <class S>
__device__ __forceinline__ S quad_next_rec(const S &v) {
    static_assert(sizeof(S) % 4 == 0, "whole dwords");
    S r;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(S) / 4); ++i)
        reinterpret_cast<int *>(&r)[i] = quad_rel<kQuadNext>(reinterpret_cast<const int *>(&v)[i]);
    return r;
}
template <int K, class S>
__device__ __forceinline__ S quad_bcast_rec(const S &v) {
    static_assert(sizeof(S) % 4 == 0, "whole dwords");
    S r;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(S) / 4); ++i)
        reinterpret_cast<int *>(&r)[i] = quad_bcast<K>(reinterpret_cast<const int *>(&v)[i]);
    return r;
}
template <class S>
__device__ __forceinline__ S quad_prev_rec(const S &v) {
    static_assert(sizeof(S) % 4 == 0, "whole dwords");
    S r;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(S) / 4); ++i)
        reinterpret_cast<int *>(&r)[i] = quad_rel<kQuadPrev>(reinterpret_cast<const int *>(&v)[i]);
    return r;
}

template <typename T>
__device__ __forceinline__ void add_w(V3<T> &acc, T w, const V3<T> &g) {
    acc.x += w * g.x; acc.y += w * g.y; acc.z += w * g.z;
}

// pass 2: one lane per halfedge h of a quad (the four lanes of a quad run
// together).  Edge e = edge(h) joins A = vertex(h) and B = vertex(next h),
// both corners of this quad.  The vertex rule adds E(e) to V(A) through h
// (beta) and through next(twin h) (delta), and to V(B) through next(h)
// (delta) and twin(h) (beta); all four weights follow from the parameters
// of A and B and the sharpness of e (the edge of next(twin h) is read only
// for a boundary crease vertex).  So every quad forms the updated gradient
// of each of its edges itself, in the order of the edge's larger halfedge
// (both quads of an edge get the same bits), and from them and its corners
// the updated gradient of its face; the larger halfedge stores its edge's.
template <typename T, int CREASED>
__device__ __forceinline__ void adj_quad_at(const AdjArgs<T> &a, int32_t h) {
    if (a.rare && h == 0) a.rare[0] = 0;   // the vertex pass that follows appends
    const bool live = h < a.H;   // H is a multiple of 4: quads are whole
    const int4 x = live ? he_gld(a.he, h) : make_int4(0, 0, 0, 0);
    const V3<T> *gV = reinterpret_cast<const V3<T> *>(a.gV);
    const VParam<T> P = a.VP[x.z];   // a dead lane's record is all zeros: vertex 0
    const V3<T> gv = gV[x.z];
    const V3<T> ge0 = live ? reinterpret_cast<const V3<T> *>(a.gE)[x.y] : V3<T>{0, 0, 0};
    const T ts = (CREASED && live) ? sharp_at(a, x.y) : (T)0;
    const T ps = CREASED ? quad_prev_rec(ts) : (T)0;
    const VW<T> w = vparam_weights<T>(P, ts, ps, x.y, CREASED);
    T c0, c1, cf;
    edge_weights<T, CREASED>(a, x.y, x.x < 0, c0, c1, cf);
    const VParam<T> PB = quad_next_rec(P);
    const V3<T> gvB = quad_next_rec(gv);
    const T dlN = CREASED ? quad_next_rec(w.dl) : (T)0;   // next(h) leaves B, its previous edge is e
    T bt = 0, dlt = 0;
    if (x.x >= 0) {
        bt = vparam_weights<T>(PB, ts, (T)0, x.y, CREASED).b;   // twin(h) leaves B along e
        if (CREASED && ts != (T)0 && vparam_boundary_crease(P)) {
            // next(twin h) leaves A; its previous edge is e, its own edge
            // decides which sharpness its walk sums
            const int32_t e2 = a.he[qnext(x.x)].edge;
            dlt = vparam_weights<T>(P, sharp_at(a, e2), ts, e2, CREASED).dl;
        }
    }
    const bool owner = x.x < 0 || h > x.x;
    V3<T> ge = ge0;
    if (owner) {
        add_w(ge, w.b, gv);
        add_w(ge, dlN, gvB);
        if (x.x >= 0) {
            add_w(ge, bt, gvB);
            add_w(ge, dlt, gv);
        }
    } else {   // the larger halfedge is twin(h), leaving B
        add_w(ge, bt, gvB);
        add_w(ge, dlt, gv);
        add_w(ge, w.b, gv);
        add_w(ge, dlN, gvB);
    }
    if (live && owner) reinterpret_cast<V3<T> *>(a.gEu)[x.y] = ge;
    // the face: its corners' gamma shares and its edges' face shares, in
    // halfedge order
    V3<T> pg{w.g * gv.x, w.g * gv.y, w.g * gv.z}, pe{cf * ge.x, cf * ge.y, cf * ge.z};
    const V3<T> g0 = quad_bcast_rec<0>(pg), e0 = quad_bcast_rec<0>(pe);
    const V3<T> g1 = quad_bcast_rec<1>(pg), e1 = quad_bcast_rec<1>(pe);
    const V3<T> g2 = quad_bcast_rec<2>(pg), e2 = quad_bcast_rec<2>(pe);
    const V3<T> g3 = quad_bcast_rec<3>(pg), e3 = quad_bcast_rec<3>(pe);
    if (!live || (h & 3) != 0) return;
    const V3<T> *gF = reinterpret_cast<const V3<T> *>(a.gF);
    V3<T> *gFo = reinterpret_cast<V3<T> *>(a.gFu ? a.gFu : a.gF);
    V3<T> gf = gF[h >> 2];
    gf.x += g0.x; gf.y += g0.y; gf.z += g0.z;
    gf.x += e0.x; gf.y += e0.y; gf.z += e0.z;
    gf.x += g1.x; gf.y += g1.y; gf.z += g1.z;
    gf.x += e1.x; gf.y += e1.y; gf.z += e1.z;
    gf.x += g2.x; gf.y += g2.y; gf.z += g2.z;
    gf.x += e2.x; gf.y += e2.y; gf.z += e2.z;
    gf.x += g3.x; gf.y += g3.y; gf.z += g3.z;
    gf.x += e3.x; gf.y += e3.y; gf.z += e3.z;
    gFo[h >> 2] = gf;
}

// pass 3: vertex v with fan f gathers alpha * gV[v] from the vertex rule
// (its own parameters, from the ring), the c0 and c1 shares of its incident
// edges from the edge rule (edge gradients after pass 2) and a quarter of
// each incident face (face gradients after pass 2)
// the gather of vertex v whose interior ring of four, h[q] and their edges
// e[q], follows from the child numbering (structural_ring)
template <typename T, int CREASED>
__device__ __forceinline__ void adj_gather_ring(const AdjArgs<T> &a, int32_t v, const int32_t (&h)[4],
                                                const int32_t (&e)[4]) {
    const V3<T> *gE = reinterpret_cast<const V3<T> *>(a.gEu);
    const V3<T> *gF = reinterpret_cast<const V3<T> *>(a.gFu ? a.gFu : a.gF);
    const V3<T> gv = reinterpret_cast<const V3<T> *>(a.gV)[v];
    T alpha = 0;
    V3<T> acc{0, 0, 0};
    T s[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) s[q] = CREASED ? sharp_at(a, e[q]) : (T)0;
    V3<T> ge[4], gf[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        ge[q] = gE[e[q]];
        gf[q] = gF[h[q] >> 2];
    }
    const VParam<T> vp = vparam_of<T>(ring_params4<T, CREASED>(e, s), CREASED);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int n1 = (i + 1) & 3;
        alpha += vparam_weights<T>(vp, s[i], s[n1], e[i], CREASED).a;
        T c0, c1, cf, d0, d1, df;
        edge_weights<T, CREASED>(a, e[i], false, c0, c1, cf);
        edge_weights<T, CREASED>(a, e[n1], false, d0, d1, df);
        add_w(acc, c0, ge[i]);
        add_w(acc, d1, ge[n1]);
        add_w(acc, (T)0.25, gf[i]);
    }
    reinterpret_cast<V3<T> *>(a.gP)[v] = {alpha * gv.x + acc.x, alpha * gv.y + acc.y, alpha * gv.z + acc.z};
}

template <typename T, int CREASED>
__device__ __forceinline__ void adj_gather_fan(const AdjArgs<T> &a, const DCage &c, const LevelTab &t, int d,
                                               const Fan &f, int32_t v) {
    const He *he = a.he;
    const V3<T> *gE = reinterpret_cast<const V3<T> *>(a.gEu);
    const V3<T> *gF = reinterpret_cast<const V3<T> *>(a.gFu ? a.gFu : a.gF);
    const V3<T> gv = reinterpret_cast<const V3<T> *>(a.gV)[v];
    T alpha = 0;
    V3<T> acc{0, 0, 0};
    {
        int32_t h[4], e[4];
        if (structural_ring(f, t, d, v, h, e, c)) {   // interior ring: no boundary edge
            adj_gather_ring<T, CREASED>(a, v, h, e);
            return;
        }
    }
    // the parameters the P pass stored for v (every vertex of the level has
    // them by now; the quad pass read them too)
    const VParam<T> vp = a.VP[v];
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
        V3<T> ev[4], ep[4], g[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            T c1, cf, d0, df;
            const T ts = CREASED ? sharp_at(a, x[q].edge) : (T)0, ps = CREASED ? sharp_at(a, xp[q].edge) : (T)0;
            wa[q] = vparam_weights<T>(vp, ts, ps, x[q].edge, CREASED).a;
            edge_weights<T, CREASED>(a, x[q].edge, x[q].twin < 0, c0[q], c1, cf);
            edge_weights<T, CREASED>(a, xp[q].edge, xp[q].twin < 0, d0, d1[q], df);
            ev[q] = gE[x[q].edge];
            ep[q] = gE[xp[q].edge];
            g[q] = gF[r[q] >> 2];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (i >= f.n) break;
            alpha += wa[i];
            add_w(acc, c0[i], ev[i]);
            add_w(acc, d1[i], ep[i]);
            add_w(acc, (T)0.25, g[i]);
        }
        reinterpret_cast<V3<T> *>(a.gP)[v] = {alpha * gv.x + acc.x, alpha * gv.y + acc.y, alpha * gv.z + acc.z};
        return;
    }
    for (int i = 0; i < f.n; ++i) {
        const int32_t r = fan_at(f, c, i);
        const int32_t p = qprev(r);
        const He x = he[r], xp = he[p];
        const T ts = CREASED ? sharp_at(a, x.edge) : (T)0, ps = CREASED ? sharp_at(a, xp.edge) : (T)0;
        alpha += vparam_weights<T>(vp, ts, ps, x.edge, CREASED).a;
        T c0, c1, cf, d0, d1, df;
        edge_weights<T, CREASED>(a, x.edge, x.twin < 0, c0, c1, cf);     // r leaves v
        edge_weights<T, CREASED>(a, xp.edge, xp.twin < 0, d0, d1, df);  // prev(r) enters v
        add_w(acc, c0, gE[x.edge]);
        add_w(acc, d1, gE[xp.edge]);
        add_w(acc, (T)0.25, gF[r >> 2]);
    }
    reinterpret_cast<V3<T> *>(a.gP)[v] = {alpha * gv.x + acc.x, alpha * gv.y + acc.y, alpha * gv.z + acc.z};
}

constexpr int kAdjScan = 4;

// Passes 1 and 3 with one thread per vertex and the fan from the vertex ID
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
    else adj_params_fan<T, CREASED>(a, c, st, d, f, v);
}

template <typename T, int CREASED>
struct RingParamsFn {
    static constexpr bool kOwners = false;
    const AdjArgs<T> &a;
    const DCage &c;
    int d;
    __device__ void operator()(const LevelTab &t, const Fan &f, int32_t v) const {
        adj_params_fan<T, CREASED>(a, c, t, d, f, v);
    }
};

// Pass 3 for the owners of a block.  RING false: every owner's vertex, any
// fan.  RING true: only vertices whose interior ring of four follows from the
// child numbering (no general fan code in the kernel, so it keeps fewer
// registers and runs more waves); the other owners go to the rare list
// (a.rare, which must be set) for k_adj_vertices_rare.
template <typename T, int CREASED, bool RING = false>
struct VertexGatherFn {
    static constexpr bool kOwners = RING;
    const AdjArgs<T> &a;
    const DCage &c;
    int d;
    __device__ void operator()(const LevelTab &t, const Fan &f, int32_t v) const {
        adj_gather_fan<T, CREASED>(a, c, t, d, f, v);
    }
    __device__ void owner(const LevelTab &t, const He *const *levels, int32_t h) const {
        Fan f;
        int32_t v, r[4], e[4];
        if (!marked_ring_fan(h, d, c, t, levels, f, v) || !structural_ring(f, t, d, v, r, e, c)) {
            const int32_t slot = atomicAdd(a.rare, 1);
            if (slot < a.rare_cap) a.rare[1 + slot] = h;
            return;
        }
        adj_gather_ring<T, CREASED>(a, v, r, e);
    }
};

// the vertices the vertex pass listed (a.rare: their owner halfedges), one
// thread each, with the fan from the owner; the caller passes this thread's
// list index
template <typename T, int CREASED, class LP>
__device__ __forceinline__ void adj_vertices_rare(const AdjArgs<T> &a, const DCage &c, const LevelTab &t, const LP &lp,
                                                  int d, int32_t i) {
    __shared__ LevelTab st;
    __shared__ LevelPtrs slp;
    lds_fill(st, t);
    lds_fill(slp, lp);
    __syncthreads();
    const int32_t n = a.rare[0] < a.rare_cap ? a.rare[0] : a.rare_cap;
    if (i >= n) return;
    Fan f;
    const int32_t v = marked_fan(a.rare[1 + i], d, c, st, slp.he, f);
    adj_gather_fan<T, CREASED>(a, c, st, d, f, v);
}

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
__global__ __launch_bounds__(256) void k_adj_params(AdjArgs<T> a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    owners_block<kAdjScan>(a.H, d, c, t, lp, xcd_block(blockIdx.x, gridDim.x, 1), a.erep,
                           RingParamsFn<T, CREASED>{a, c, d});
}

template <typename T, int CREASED>
__global__ __launch_bounds__(256) void k_adj_quads(AdjArgs<T> a) {
    adj_quad_at<T, CREASED>(a, (int32_t)(xcd_block(blockIdx.x, gridDim.x, 1) * blockDim.x + threadIdx.x));
}

template <typename T, int CREASED, bool RING>
__global__ __launch_bounds__(256) void k_adj_vertices(AdjArgs<T> a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    owners_block<kAdjScan>(a.H, d, c, t, lp, xcd_block(blockIdx.x, gridDim.x, 1), a.erep,
                           VertexGatherFn<T, CREASED, RING>{a, c, d});
}

template <typename T, int CREASED>
__global__ __launch_bounds__(256) void k_adj_vertices_rare(AdjArgs<T> a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    adj_vertices_rare<T, CREASED>(a, c, t, lp, d, (int32_t)(blockIdx.x * 256 + threadIdx.x));
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
    // G holds the gradient of level d+1, P receives level d.  The level
    // passes only read G (the quad pass writes the updated face gradients to
    // W), so the finest level reads the caller's gradient itself; the cage
    // step updates G in place, so a cage refined once works on a copy.
    T *G = const_cast<T *>(gradFinest), *P = bufB;
    VParam<T> *cache = sizeof(T) == sizeof(float) ? static_cast<VParam<T> *>(pl.vp_cache) : nullptr;
    const bool params = !(cache && pl.vp_valid);   // run the parameter pass
    if (D == 1) {
        G = bufA;
        hipLaunchKernelGGL(k_copy<T>, dim3(grid_stride(3 * n.V[D])), dim3(256), 0, st, gradFinest, G, 3 * n.V[D]);
    }
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
        a.VP = cache ? cache + n.offV[d] : reinterpret_cast<VParam<T> *>(W);
        a.gEu = W + 4 * n.V[d];
        a.gFu = W + 4 * n.V[d] + 3 * n.E[d];
        a.erep = pl.erep;
        const dim3 g(grid_for(n.H[d])), gq(grid_for(n.H[d] / 4)), gv(grid_for(n.V[d])), b(256);
        const bool pv = n.V[d] <= pl.per_vertex_max;
        if (!pv && pl.rare) {
            a.rare = pl.rare;
            a.rare_cap = pl.rare_cap;
        }
        const dim3 gr(grid_for(pl.rare_cap > 0 ? pl.rare_cap : 1));
        if (creased) {
            if (!params) {
            } else if (pv) {
                hipLaunchKernelGGL((k_adj_per_vertex<T, 1, 0>), gv, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            } else {
                hipLaunchKernelGGL((k_adj_params<T, 1>), gq, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            }
            hipLaunchKernelGGL((k_adj_quads<T, 1>), g, b, 0, st, a);
            if (pv) hipLaunchKernelGGL((k_adj_per_vertex<T, 1, 1>), gv, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            else if (a.rare) hipLaunchKernelGGL((k_adj_vertices<T, 1, true>), gq, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            else hipLaunchKernelGGL((k_adj_vertices<T, 1, false>), gq, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            if (a.rare && a.rare_cap > 0) hipLaunchKernelGGL((k_adj_vertices_rare<T, 1>), gr, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
        } else {
            if (!params) {
            } else if (pv) {
                hipLaunchKernelGGL((k_adj_per_vertex<T, 0, 0>), gv, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            } else {
                hipLaunchKernelGGL((k_adj_params<T, 0>), gq, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            }
            hipLaunchKernelGGL((k_adj_quads<T, 0>), g, b, 0, st, a);
            if (pv) hipLaunchKernelGGL((k_adj_per_vertex<T, 0, 1>), gv, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            else if (a.rare) hipLaunchKernelGGL((k_adj_vertices<T, 0, true>), gq, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            else hipLaunchKernelGGL((k_adj_vertices<T, 0, false>), gq, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
            if (a.rare && a.rare_cap > 0) hipLaunchKernelGGL((k_adj_vertices_rare<T, 0>), gr, b, 0, st, a, pl.cage, pl.tab, pl.lp, d);
        }
        // the next level reads P; the one after writes bufA (never the
        // caller's gradient), then the two alternate
        T *const next = d == D - 1 ? bufA : G;
        G = P;
        P = next;
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
    a.VP = nullptr;
    a.gEu = nullptr;
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
struct OpAdjParams {
    int d;
    char *arena, *adj;
    __device__ void operator()(const AdjJob &J, int64_t i) const {
        AdjArgs<float> a;
        DCage c;
        unpack_adj(J, arena, adj, a, c);
        const JobLevelPtrs lp{J.lp, arena};
        owners_block<kAdjScan>(a.H, d, c, J.t, lp, (uint32_t)(i >> 8), a.erep, RingParamsFn<float, CREASED>{a, c, d});
    }
};
template <int CREASED>
struct OpAdjQuads {
    char *arena, *adj;
    __device__ void operator()(const AdjJob &J, int64_t i) const {
        AdjArgs<float> a;
        DCage c;
        unpack_adj(J, arena, adj, a, c);
        adj_quad_at<float, CREASED>(a, (int32_t)i);
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
        // every job's level d >= 1 has a rare list
        owners_block<kAdjScan>(a.H, d, c, J.t, lp, (uint32_t)(i >> 8), a.erep,
                               VertexGatherFn<float, CREASED, true>{a, c, d});
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
// the vertices job k's vertex pass listed, one thread each
template <int CREASED>
__global__ __launch_bounds__(256) void k_batch_adj_vertices_rare(const AdjJob *__restrict__ jobs,
                                                                 const int32_t *__restrict__ start, int n, int d,
                                                                 char *arena, char *adj) {
    const int k = find_job(start, n, blockIdx.x);
    const AdjJob &J = jobs[k];
    AdjArgs<float> a;
    DCage c;
    unpack_adj(J, arena, adj, a, c);
    const JobLevelPtrs lp{J.lp, arena};
    adj_vertices_rare<float, CREASED>(a, c, J.t, lp, d, (int32_t)((blockIdx.x - (uint32_t)start[k]) * 256 + threadIdx.x));
}

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

// level d of every job: vertex parameters, quads (edges and faces) and
// vertex gathers in turn; `quads` has one workgroup per 1024 halfedges of
// each job (the owner passes), `halfedges` one per 256
void batch_adj_level(const BatchTable &quads, const BatchTable &halfedges, const BatchTable &rare, char *arena, char *adj,
                     int d, int creased, int params, hipStream_t st) {
    const auto rare_pass = [&](auto kernel) {
        if (rare.blocks == 0 || rare.n == 0) return;
        hipLaunchKernelGGL(kernel, dim3(rare.blocks), dim3(256), 0, st, static_cast<const AdjJob *>(rare.jobs),
                           rare.start, (int)rare.n, d, arena, adj);
    };
    if (creased) {
        if (params) launch_batch<AdjJob>(quads, OpAdjParams<1>{d, arena, adj}, st);
        launch_batch<AdjJob>(halfedges, OpAdjQuads<1>{arena, adj}, st);
        launch_batch<AdjJob>(quads, OpAdjVertices<1>{d, arena, adj}, st);
        rare_pass(k_batch_adj_vertices_rare<1>);
    } else {
        if (params) launch_batch<AdjJob>(quads, OpAdjParams<0>{d, arena, adj}, st);
        launch_batch<AdjJob>(halfedges, OpAdjQuads<0>{arena, adj}, st);
        launch_batch<AdjJob>(quads, OpAdjVertices<0>{d, arena, adj}, st);
        rare_pass(k_batch_adj_vertices_rare<0>);
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
