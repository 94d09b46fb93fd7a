// Device helpers shared by the fused forward kernels and the adjoint:
// launch remapping, quad-lane exchanges, the point rules and the structural
// enumeration of vertex rings.
#pragma once

#include "cc_common.h"
#include "cc_fused.h"

namespace cch {

// Copies a small table (kernel argument or job field) into LDS, one dword
// per thread.  Reads of LevelTab / LevelPtrs at a per-lane index become
// vector memory loads when the table is a kernel argument, with one memory
// round trip per step of the level loops; from LDS they are cheap.  The
// caller synchronizes the block before the first read.
template <class T>
__device__ __forceinline__ void lds_fill(T &dst, const T &src) {
    static_assert(sizeof(T) % 4 == 0, "tables are whole dwords");
    for (int i = (int)threadIdx.x; i < (int)(sizeof(T) / 4); i += (int)blockDim.x)
        reinterpret_cast<int32_t *>(&dst)[i] = reinterpret_cast<const int32_t *>(&src)[i];
}

// value of `v` held by lane (quad base + k) for k = 0..3 (DPP quad_perm)
template <int K>
__device__ __forceinline__ int quad_bcast(int v) {
    return __builtin_amdgcn_update_dpp(0, v, K * 0x55, 0xF, 0xF, false);
}
template <int K>
__device__ __forceinline__ float quad_bcastf(float v) {
    return __int_as_float(quad_bcast<K>(__float_as_int(v)));
}
template <int K>
__device__ __forceinline__ P3 quad_bcast3(P3 p) {
    return {quad_bcastf<K>(p.x), quad_bcastf<K>(p.y), quad_bcastf<K>(p.z)};
}

// A row of 16 lanes rotated right by N: lane i receives lane (i - N) mod 16
// of its row (DPP row_ror).
template <int N>
__device__ __forceinline__ int row_rori(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x120 + N, 0xF, 0xF, false);
}
template <int N>
__device__ __forceinline__ float row_ror(float v) {
    return __int_as_float(row_rori<N>(__float_as_int(v)));
}
template <int N>
__device__ __forceinline__ P3 row_ror3(P3 v) {
    return P3{row_ror<N>(v.x), row_ror<N>(v.y), row_ror<N>(v.z)};
}

// value held by the lane of the next (+1), opposite (+2) or previous (+3)
// halfedge of the same quad (DPP quad_perm [1,2,3,0], [2,3,0,1], [3,0,1,2])
constexpr int kQuadNext = 0x39, kQuadOpp = 0x4E, kQuadPrev = 0x93;
template <int CTRL>
__device__ __forceinline__ int quad_rel(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float quad_relf(float v) {
    return __int_as_float(quad_rel<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ int4 quad_rel4(int4 v) {
    return make_int4(quad_rel<CTRL>(v.x), quad_rel<CTRL>(v.y), quad_rel<CTRL>(v.z), quad_rel<CTRL>(v.w));
}
template <int CTRL>
__device__ __forceinline__ UV quad_reluv(UV t) {
    return {quad_relf<CTRL>(t.u), quad_relf<CTRL>(t.v)};
}

template <typename T>
__device__ __forceinline__ T pick4(T a0, T a1, T a2, T a3, int k) {
    return k == 0 ? a0 : (k == 1 ? a1 : (k == 2 ? a2 : a3));
}

// Trailing zero base-4 digits of h >= 0 (children 0 of children 0 ...), at
// most d - 1: h = q << 2k with the vertex of h born at level d - k.
__device__ __forceinline__ int zero_digits(int32_t h, int d) {
    const int cap = d > 1 ? d - 1 : 0;
    const int z = h ? (__builtin_ctz((uint32_t)h) >> 1) : cap;
    return z < cap ? z : cap;
}

__device__ __forceinline__ float lsharp(const Crease *cr, int32_t C, int32_t e) {
    return e < C ? cr[e].sharp : 0.0f;
}

// face point of a level-d quad from its four corner points in halfedge
// order: gather form (:2015-2025) or scatter form (v/4 accumulated, :2048)
__device__ __forceinline__ P3 face_rule(P3 p0, P3 p1, P3 p2, P3 p3, bool gather) {
    if (gather) return mul(add(add(add(p0, p1), p2), p3), 0.25f);
    P3 a = zero3();
    a.x += p0.x / 4.0f; a.y += p0.y / 4.0f; a.z += p0.z / 4.0f;
    a.x += p1.x / 4.0f; a.y += p1.y / 4.0f; a.z += p1.z / 4.0f;
    a.x += p2.x / 4.0f; a.y += p2.y / 4.0f; a.z += p2.z / 4.0f;
    a.x += p3.x / 4.0f; a.y += p3.y / 4.0f; a.z += p3.z / 4.0f;
    return a;
}

__device__ __forceinline__ P3 quad_face(const He *he, const P3 *old, int32_t quad_base, bool gather) {
    const HeView q4{he + quad_base};
    return face_rule(old[q4[0].z], old[q4[1].z], old[q4[2].z], old[q4[3].z], gather);
}

// scatter-form edge contribution (:2118-2136 smooth, :2205-2238 creased)
__device__ __forceinline__ P3 edge_contrib(P3 v0, P3 v1, P3 f, bool boundary, float sharp, bool creased) {
    if (creased) {
        const float w = sat(sharp);
        P3 sp = mul(lerp(v0, v1, 0.5f), boundary ? 1.0f : 0.5f);
        P3 sm = mul(lerp(v0, f, 0.5f), 0.5f);
        return lerp(sm, sp, w);
    }
    const float w = boundary ? 1.0f : 0.5f;
    return lerp(mul(f, 0.5f), lerp(mul(v0, w), mul(v1, w), 0.5f), w);
}

// gather-form edge point (:2094-2102 smooth, :2180-2188 creased)
__device__ __forceinline__ P3 edge_gather(P3 v0, P3 v1, P3 f0, P3 f1, float w, bool creased) {
    P3 t1 = add(v0, v1), t2 = add(f0, f1);
    P3 sp = mul(t1, 0.5f);
    P3 sm = mul(add(t1, t2), 0.25f);
    return creased ? lerp(sm, sp, w) : lerp(sp, sm, w);
}

__device__ __forceinline__ int4 child_record(int k, int32_t h, int4 x, int4 p, int32_t V, int32_t F, int32_t E) {
    const int32_t prev = qprev(h);
    switch (k) {
    case 0: return make_int4(i4(qnext(x.x)) + 3, 2 * x.y + (h > x.x ? 0 : 1), x.z, 0);
    case 1: return make_int4(i4(qnext(h)) + 2, 2 * E + h, V + F + x.y, 0);
    case 2: return make_int4(i4(prev) + 1, 2 * E + prev, V + (h >> 2), 0);
    default: return make_int4(i4(p.x) + 0, 2 * p.y + (prev > p.x ? 1 : 0), V + F + p.y, 0);
    }
}


// int4s of LDS per parent halfedge that stage its four child records
constexpr int kStageInt4 = kHeBytes / 4;

// Slot of child k of thread t among the thread's four 16-byte LDS slots.
// Written in place (slot k), the four store passes put lanes t, t + 4, ...
// on the same banks (a 64-byte stride); rotating by s(t) = t/2 + t/8 makes
// every group of 8 lanes (128 bytes a cycle) and of 16 lanes (256 bytes)
// cover distinct banks in each pass.
__device__ __forceinline__ int stage_slot(uint32_t t, int k) {
    return (int)(4 * t) + ((k + (int)(t >> 1) + (int)(t >> 3)) & 3);
}

// Child records staged through LDS: this thread's four children go to its
// slots of `stage` (kStageInt4 int4s per thread), then, after the block's
// barrier, the block's contiguous run of 4 * NT records (NT threads) is
// written as whole int4s, each store instruction covering 1 KiB of
// consecutive memory.  The run starts at record 4 * base, a multiple of 16
// records, so it is 16-byte aligned for either record size.
template <int NT>
__device__ __forceinline__ void staged_children_store(int4 *stage, const int4 (&c)[4], He *out, int32_t base,
                                                      int32_t H) {
    if (kHasUv) {
#pragma unroll
        for (int k = 0; k < 4; ++k) stage[stage_slot(threadIdx.x, k)] = c[k];
    } else {
        int32_t *w = reinterpret_cast<int32_t *>(stage) + 12 * threadIdx.x;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            w[3 * k] = c[k].x;
            w[3 * k + 1] = c[k].y;
            w[3 * k + 2] = c[k].z;
        }
    }
    __syncthreads();
    // the block's run from a block-uniform base (scalar registers), lanes at
    // 32-bit offsets within it
    int4 *const run = reinterpret_cast<int4 *>(out) + kStageInt4 * (int64_t)base;
    const int64_t left = kStageInt4 * ((int64_t)H - base);   // int4s of the run inside the array
#pragma unroll
    for (int k = 0; k < kStageInt4; ++k) {
        const uint32_t q = threadIdx.x + (uint32_t)k * NT;   // record q & 3 of thread q / 4
        const int4 v = kHasUv ? stage[stage_slot(q >> 2, (int)(q & 3))] : stage[q];
        if ((int64_t)q < left) store_stream(at32(run, q), v);
    }
}

// ---- vertex points ----------------------------------------------------------
// Per-halfedge scatter contribution (creased :2509-2611 with the corner test
// first; smooth :2320-2348) from the quantities the ring walk gathered.
struct RingParams {
    float n;        // halfedges around the vertex (valence before the boundary correction)
    float cc;       // crease count including the boundary correction
    bool boundary;
    float s1, s2;   // the first two nonzero sharpness values seen ...
    int32_t e1, e2; // ... and their edges
    float sall;     // sum of every incident sharpness
};

__device__ __forceinline__ P3 vertex_scatter_contrib(const RingParams &rp, P3 old, P3 E, P3 PE, P3 Fc, float ts,
                                                     float ps, int32_t edge, bool creased) {
    if (!creased) {
        const float w = 1.0f / rp.n;
        const float s = rp.boundary ? 0.0f : 1.0f;
        P3 a;
        a.x = w * (old.x + w * s * (4.0f * E.x - Fc.x - 3.0f * old.x));
        a.y = w * (old.y + w * s * (4.0f * E.y - Fc.y - 3.0f * old.y));
        a.z = w * (old.z + w * s * (4.0f * E.z - Fc.z - 3.0f * old.z));
        return a;
    }
    const float cw = sgn(ts), pcw = sgn(ps);
    const float n = rp.n;
    const P3 corner = mul(old, 1.0f / n);
    P3 cr = add(mul(old, 0.25f * cw), mul(E, 0.25f * cw));
    P3 sm = add(mul(Fc, -1.0f), mul(E, 4.0f));
    sm = add(sm, mul(old, n - 3.0f));
    sm = mul(sm, 1.0f / (n * n));
    const float nn = rp.boundary ? n + 1.0f : n;
    if (rp.boundary) cr = add(cr, add(mul(old, 0.25f * pcw), mul(PE, 0.25f * pcw)));
    if (rp.cc >= 3.0f || nn == 2.0f) return corner;
    if (rp.cc <= 1.0f) return sm;
    // two creased edges: the sum of the incident sharpness this halfedge's
    // walk visits (every edge but its own on a boundary)
    float avg = rp.sall;
    if (rp.boundary) avg = edge == rp.e1 ? rp.s2 : (edge == rp.e2 ? rp.s1 : rp.s1 + rp.s2);
    return lerp(corner, cr, sat(avg * 0.5f));
}

// ---- vertex points with structurally enumerated rings -------------------------
// The halfedges around a vertex of level d follow from the level c at which
// the vertex appeared (CatmullClark.h:2770-2793): children 0 of a parent keep
// the parent's vertex, so a ring at level d is the ring at level c shifted
// by 2(d - c) bits; a face vertex of a quad face f' (level c-1 >= 1) has the
// ring 16f'+2, +6, +10, +14; an edge vertex of edge(p) with twin t has
// 4p+1, 4next(t)+3, 4t+1, 4next(p)+3 (or 4next(p)+3, 4p+1 on a boundary);
// cage vertices and faces of the cage take their ring from the cage.  The
// ring is listed in the reference's forward walk order (twin(prev(x))),
// starting at the vertex's first boundary halfedge or at its representative
// (CatmullClark.h:1336), so no step waits on the previous one.
struct Fan {
    int kind;       // 0 cage vertex, 1 face vertex of a quad, 2 face vertex of a cage face, 3 edge vertex
    int shift;      // level-d halfedge = level-c halfedge << shift
    int n;          // halfedges around the vertex
    int ri;         // index of the representative in the fan
    bool boundary;
    int32_t a0, a1, a2, a3;   // kinds 0 and 2 with cage rings: a1 = offset of the ring's list
};

__device__ __forceinline__ int32_t fan_at(const Fan &f, const DCage &c, int i) {
    const int j = i < f.n ? (i > 0 ? i : 0) : (f.n > 1 ? f.n - 1 : 0);   // ring lists: stay in the list
    switch (f.kind) {
    case 1: return (f.a0 + 4 * i) << f.shift;
    case 3: return pick4(f.a0, f.a1, f.a2, f.a3, i) << f.shift;
    case 2: {  // the face's cage halfedges from f2h, children 2
        if (c.fring) return (4 * c.fring[f.a1 + j] + 2) << f.shift;
        int32_t p = f.a0;
        for (int k = 0; k < i && k < kWalkCap; ++k) p = c.he[p].next;
        return (4 * p + 2) << f.shift;
    }
    default: {  // the vertex's cage halfedges in forward order, children 0
        if (c.vring) return (4 * c.vring[f.a1 + j]) << f.shift;
        int32_t g = f.a0;
        for (int k = 0; k < i && k < kWalkCap; ++k) g = c.he[c.he[g].prev].twin;
        return (4 * g) << f.shift;
    }
    }
}

// the fan of cage vertex v with first halfedge p (kind 0): from its ring list
// or by walking twin(prev(g)) (CatmullClark.h:1336-1345)
__device__ __forceinline__ void cage_vertex_fan(const DCage &c, int32_t v, int32_t p, Fan &f) {
    f.kind = 0;
    f.a0 = p;
    if (p < 0) {   // a vertex without halfedges: an empty fan
        f.n = 0;
        f.a1 = 0;
        return;
    }
    if (c.vring) {
        f.a1 = c.vring_off[v];
        f.n = c.vring_off[v + 1] - f.a1;
        f.boundary = c.vring_bnd[v] != 0;
        return;
    }
    int n = 1;
    int32_t g = c.he[c.he[p].prev].twin;
    for (; g >= 0 && g != p && n < kWalkCap; g = c.he[c.he[g].prev].twin) ++n;
    f.n = n;
    f.boundary = g < 0;
}

// the fan of the face point of cage face `face` with first halfedge p (kind 2)
__device__ __forceinline__ void cage_face_fan(const DCage &c, int32_t face, int32_t p, Fan &f) {
    f.kind = 2;
    f.a0 = p;
    if (c.fring) {
        f.a1 = c.fring_off[face];
        f.n = c.fring_off[face + 1] - f.a1;
        return;
    }
    int n = 1;
    for (int32_t it = c.he[p].next; it != p && n < kWalkCap; it = c.he[it].next) ++n;
    f.n = n;
}

// Is h the representative of its vertex (CatmullClark.h:1336)?  The test of
// owner_fan without building the fan (no twin record is read).
__device__ __forceinline__ bool owner_test(int32_t h, int32_t vert, int d, const DCage &c, const LevelTab &t,
                                           const uint64_t *erep = nullptr) {
    const int k = zero_digits(h, d);
    const int32_t q = h >> (2 * k);
    const int lvl = d - k;
    const int digit = q & 3;
    const int32_t p = q >> 2;
    if (digit == 0) return c.v2h[vert] == p;
    if (digit == 3) return false;
    if (digit == 2) return lvl - 1 >= 1 ? (p & 3) == 0 : c.f2h[c.he[p].face] == p;
    const int32_t e = vert - t.V[lvl - 1] - t.F[lvl - 1];
    if (lvl - 1 >= 1 && erep) return ((erep[t.R[lvl - 1] + (p >> 6)] >> (p & 63)) & 1ull) != 0ull;
    return lvl - 1 >= 1 ? is_edge_rep(c, t, e, lvl - 1, p) : c.e2h[e] == p;
}

// OR of v over the 16 lanes of this lane's DPP row, in the row's last lane
// (an inclusive OR scan by row_shr 1, 2, 4, 8; lanes before the row start
// read 0)
__device__ __forceinline__ uint32_t row_or(uint32_t v) {
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);
    v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);
    return v;
}

// Is h the representative of its vertex?  If so, describe the fan.
__device__ __forceinline__ bool owner_fan(int32_t h, int32_t vert, int d, const DCage &c, const LevelTab &t,
                                          const He *const *levels, Fan &f) {
    const int k = zero_digits(h, d);
    const int32_t q = h >> (2 * k);
    const int lvl = d - k;            // level at which the vertex appeared
    const int digit = q & 3;
    const int32_t p = q >> 2;         // parent halfedge, level lvl-1
    f.shift = 2 * k;
    f.boundary = false;
    f.ri = 0;
    if (digit == 0) {                 // lvl == 1: a cage vertex
        if (c.v2h[vert] != p) return false;
        cage_vertex_fan(c, vert, p, f);
        return true;
    }
    if (digit == 3) return false;
    if (digit == 2) {
        if (lvl - 1 >= 1) {
            if ((p & 3) != 0) return false;
            f.kind = 1;
            f.a0 = 4 * p + 2;
            f.n = 4;
            return true;
        }
        const int32_t face = c.he[p].face;
        if (c.f2h[face] != p) return false;
        cage_face_fan(c, face, p, f);
        return true;
    }
    // digit 1: an edge vertex of edge(p)
    const int32_t e = vert - t.V[lvl - 1] - t.F[lvl - 1];
    int32_t tw, np, nt;
    if (lvl - 1 >= 1) {
        if (!is_edge_rep(c, t, e, lvl - 1, p)) return false;
        np = qnext(p);
        if (k == 0) {
            // born at this level: child 0 of p (halfedge 4p of level d, next
            // to h = 4p + 1) has the twin 4 next(twin(p)) + 3, negative on a
            // boundary (CatmullClark.h:2777), so the parent level is not read
            const int32_t tc = levels[d][4 * p].twin;
            nt = tc >> 2;
            tw = tc < 0 ? -1 : qprev(nt);
        } else {
            tw = levels[lvl - 1][p].twin;
            nt = qnext(tw);
        }
    } else {
        if (c.e2h[e] != p) return false;
        tw = c.he[p].twin;
        np = c.he[p].next;
        nt = tw >= 0 ? c.he[tw].next : -1;
    }
    f.kind = 3;
    if (tw >= 0) {
        f.n = 4;
        f.a0 = 4 * p + 1; f.a1 = 4 * nt + 3; f.a2 = 4 * tw + 1; f.a3 = 4 * np + 3;
    } else {
        f.n = 2;
        f.boundary = true;
        f.ri = 1;
        f.a0 = 4 * np + 3; f.a1 = 4 * p + 1; f.a2 = f.a3 = 0;
    }
    return true;
}

// Is level-d halfedge h the representative of its vertex?  owner_test
// without h's record: a face vertex born at level >= 2 is owned by child 2
// of its quad's first halfedge, an edge vertex by the child 1 whose parent
// represents the edge (one parent record read), a cage-level vertex as the
// cage maps say.
__device__ __forceinline__ bool structural_owner(int32_t h, int d, const DCage &c, const LevelTab &t,
                                                 const He *const *levels, const uint64_t *erep = nullptr) {
    const int k = zero_digits(h, d);
    const int32_t q = h >> (2 * k);
    const int lvl = d - k;
    const int digit = q & 3;
    const int32_t p = q >> 2;
    if (digit == 3) return false;
    if (lvl >= 2) {
        if (digit == 2) return (p & 3) == 0;
        // the parent level's edge-representative bit when it is kept
        if (erep) return ((erep[t.R[lvl - 1] + (p >> 6)] >> (p & 63)) & 1ull) != 0ull;
        return is_edge_rep(c, t, levels[lvl - 1][p].edge, lvl - 1, p);
    }
    const CageHe x = c.he[p];
    if (digit == 0) return c.v2h[x.vert] == p;
    if (digit == 2) return c.f2h[x.face] == p;
    return c.e2h[x.edge] == p;
}

// The fan and vertex ID of a marked owner h when its vertex is a face
// vertex born at level >= 2, the face vertex of a cage quad with halfedges
// p .. p+3 in next order, a cage vertex with a closed ring of four, or an
// interior edge vertex (the fans structural_ring accepts), reading at most
// the parent's record (cage data for cage vertices and faces); false for any
// other vertex (other cage vertices and faces, boundary edges).
__device__ __forceinline__ bool marked_ring_fan(int32_t h, int d, const DCage &c, const LevelTab &t,
                                                const He *const *levels, Fan &f, int32_t &v) {
    const int k = zero_digits(h, d);
    const int32_t q = h >> (2 * k);
    const int lvl = d - k;
    const int digit = q & 3;
    const int32_t p = q >> 2;
    f.shift = 2 * k;
    f.boundary = false;
    f.ri = 0;
    f.n = 4;
    if (digit == 2) {
        if (lvl >= 2) {
            f.kind = 1;
            f.a0 = 4 * p + 2;
            v = (int32_t)t.V[lvl - 1] + (p >> 2);
            return true;
        }
        // the face vertex of cage face face(p): the same ring when the face
        // is a quad whose halfedges are p .. p+3 in next order (p, its
        // representative, first)
        if (p + 3 >= c.H) return false;
        const CageHe x = c.he[p];
        const int32_t n1 = c.he[p + 1].next, n2 = c.he[p + 2].next, n3 = c.he[p + 3].next;
        if (x.next != p + 1 || n1 != p + 2 || n2 != p + 3 || n3 != p) return false;
        f.kind = 1;
        f.a0 = 4 * p + 2;
        v = c.V + x.face;
        return true;
    }
    if (digit == 0) {
        // a cage vertex (p, its representative, leaves it): the ring of four
        // from the cage's ring list when that ring is closed
        if (!c.vring) return false;
        const int32_t vv = c.he[p].vert;
        const int32_t o = c.vring_off[vv];
        if (c.vring_off[vv + 1] - o != 4 || c.vring_bnd[vv] != 0) return false;
        f.kind = 0;
        f.a0 = p;
        f.a1 = o;
        v = vv;
        return true;
    }
    if (digit != 1) return false;
    int32_t tw, np, nw, e;
    if (lvl >= 2) {
        const int4 x = he_load(levels[lvl - 1], p);
        tw = x.x;
        e = x.y;
        np = qnext(p);
        nw = qnext(tw);
    } else {
        const CageHe x = c.he[p];
        tw = x.twin;
        e = x.edge;
        np = x.next;
        nw = tw >= 0 ? c.he[tw].next : -1;
    }
    if (tw < 0) return false;
    f.kind = 3;
    f.a0 = 4 * p + 1; f.a1 = 4 * nw + 3; f.a2 = 4 * tw + 1; f.a3 = 4 * np + 3;
    v = (int32_t)(t.V[lvl - 1] + t.F[lvl - 1]) + e;
    return true;
}

// The fan and vertex ID of a level-d halfedge h known to be the
// representative of its vertex (marked by the fused kernel).  A face or edge
// vertex born at level lvl >= 2 takes its ID from the child numbering
// (child 2 of p: V + face(p), child 1: V + F + edge(p), parent-level counts),
// so only an edge vertex reads one record, its parent's (twin and edge);
// cage-level vertices read h's record and go through owner_fan.
__device__ __forceinline__ int32_t marked_fan(int32_t h, int d, const DCage &c, const LevelTab &t,
                                              const He *const *levels, Fan &f) {
    const int k = zero_digits(h, d);
    const int32_t q = h >> (2 * k);
    const int lvl = d - k;
    const int digit = q & 3;
    const int32_t p = q >> 2;
    if (lvl >= 2 && digit == 2) {
        f.kind = 1;
        f.shift = 2 * k;
        f.boundary = false;
        f.ri = 0;
        f.n = 4;
        f.a0 = 4 * p + 2;
        return (int32_t)t.V[lvl - 1] + (p >> 2);
    }
    if (lvl >= 2 && digit == 1) {
        const int4 x = he_load(levels[lvl - 1], p);
        const int32_t tw = x.x, np = qnext(p);
        f.kind = 3;
        f.shift = 2 * k;
        if (tw >= 0) {
            f.boundary = false;
            f.ri = 0;
            f.n = 4;
            f.a0 = 4 * p + 1; f.a1 = 4 * qnext(tw) + 3; f.a2 = 4 * tw + 1; f.a3 = 4 * np + 3;
        } else {
            f.boundary = true;
            f.ri = 1;
            f.n = 2;
            f.a0 = 4 * np + 3; f.a1 = 4 * p + 1; f.a2 = f.a3 = 0;
        }
        return (int32_t)(t.V[lvl - 1] + t.F[lvl - 1]) + x.y;
    }
    const int32_t v = levels[d][h].vert;
    owner_fan(h, v, d, c, t, levels, f);
    return v;
}

// The owners of level-d vertices among 256 * SCAN halfedges (structural
// test, no scan of the records), collected in LDS so that every thread then
// works on one, in halfedge order (neighbouring rings, shared cache lines):
// fn(tables, fan, vertex), or with Fn::kOwners fn.owner(tables, level
// records, owner halfedge), which finds the fan itself.  The level tables go
// to LDS first; the first barrier orders the copy.
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
        if constexpr (Fn::kOwners) {
            fn.owner(st, slp.he, owners[i]);
        } else {
            Fan f;
            const int32_t v = marked_fan(owners[i], d, c, st, slp.he, f);
            fn(st, f, v);
        }
    }
}


// The fan of vertex v of level d >= 1, equal to what owner_fan gives its
// representative (CatmullClark.h:1336): v was born at the level b with
// V[b-1] <= v < V[b] as the face vertex of face v - V[b-1] or the edge
// vertex of edge v - V[b-1] - F[b-1] of level b-1 (a cage vertex when
// v < V[0]); its ring at level d is the ring at level b shifted by 2(d-b).
__device__ __forceinline__ void vertex_fan(int32_t v, int d, const DCage &c, const LevelTab &t,
                                           const He *const *levels, Fan &f) {
    int b = d;
    while (b > 0 && v < t.V[b - 1]) --b;
    f.boundary = false;
    f.ri = 0;
    if (b == 0) {   // a cage vertex
        f.shift = 2 * (d - 1);
        cage_vertex_fan(c, v, c.v2h[v], f);
        return;
    }
    f.shift = 2 * (d - b);
    const int32_t Vp = b - 1 >= 1 ? t.V[b - 1] : c.V;
    const int32_t Fp = b - 1 >= 1 ? t.F[b - 1] : c.F;
    if (v < Vp + Fp) {   // a face vertex
        const int32_t face = v - Vp;
        if (b - 1 >= 1) {
            f.kind = 1;
            f.a0 = 16 * face + 2;
            f.n = 4;
            return;
        }
        cage_face_fan(c, face, c.f2h[face], f);
        return;
    }
    const int32_t e = v - Vp - Fp;   // an edge vertex
    int32_t p, tw, np, nt;
    if (b - 1 >= 1) {
        p = edge_rep(c, t, e, b - 1);
        tw = levels[b - 1][p].twin;
        np = qnext(p);
        nt = qnext(tw);
    } else {
        p = c.e2h[e];
        tw = c.he[p].twin;
        np = c.he[p].next;
        nt = tw >= 0 ? c.he[tw].next : -1;
    }
    f.kind = 3;
    if (tw >= 0) {
        f.n = 4;
        f.a0 = 4 * p + 1; f.a1 = 4 * nt + 3; f.a2 = 4 * tw + 1; f.a3 = 4 * np + 3;
    } else {
        f.n = 2;
        f.boundary = true;
        f.ri = 1;
        f.a0 = 4 * np + 3; f.a1 = 4 * p + 1; f.a2 = f.a3 = 0;
    }
}

// Everything the vertex rules need about one fan halfedge r: the points of
// its own edge and of its previous edge, its face point, both sharpnesses.
struct FanEntry {
    P3 E, PE, F;
    float ts, ps;
    int32_t re, pe;
    int32_t r;   // the fan halfedge itself (orders the scatter sum)
};

__device__ __forceinline__ FanEntry fan_entry(const FusedArgs &a, int32_t r, int4 rec, int4 prv) {
    FanEntry e;
    e.r = r;
    e.re = rec.y;
    e.pe = prv.y;
    e.E = a.ep[e.re];
    e.PE = a.ep[e.pe];
    e.F = a.fp[r >> 2];
    e.ts = lsharp(a.cr, a.C, e.re);
    e.ps = lsharp(a.cr, a.C, e.pe);
    return e;
}

// fans of up to four halfedges (face and edge vertices) with every load
// issued before any is used
struct FanPreloaded {
    FanEntry e0, e1, e2, e3;
    __device__ __forceinline__ FanEntry get(int i) const { return pick4(e0, e1, e2, e3, i); }
    __device__ __forceinline__ int32_t key(int i) const { return pick4(e0.r, e1.r, e2.r, e3.r, i); }
};

// any fan, entries loaded when asked for
struct FanOnDemand {
    const FusedArgs *a;
    const Fan *f;
    const DCage *c;
    __device__ __forceinline__ FanEntry get(int i) const {
        const int32_t r = fan_at(*f, *c, i);
        const HeView he4{a->he};
        return fan_entry(*a, r, he4[r], he4[qprev(r)]);
    }
    __device__ __forceinline__ int32_t key(int i) const { return fan_at(*f, *c, i); }
};

// The vertex rules over a fan (see k_level_vertices for the reference walk
// they reproduce).  MAXN > 0 bounds the fan size at compile time so the loops
// unroll.
template <int GATHER, int CREASED, int MAXN, class Acc>
__device__ __forceinline__ P3 fan_vertex_point(const Fan &f, const Acc &acc, P3 old) {
    const int n = f.n;
    const int ri = f.ri;
    const int lim = MAXN > 0 ? MAXN : n;
    if (GATHER) {
        if (!CREASED) {
            P3 sum = zero3();
            float nv = 1.0f;
            for (int i = 0; i < lim; ++i) {
                if (i < ri || i >= n) continue;
                const FanEntry e = acc.get(i);
                if (i == ri) {
                    sum = add(mul(e.F, -1.0f), mul(e.E, 4.0f));
                } else {
                    sum = add(sum, mul(e.F, -1.0f));
                    sum = add(sum, mul(e.E, 4.0f));
                    nv = nv + 1.0f;
                }
            }
            P3 sm = add(mul(sum, 1.0f / (nv * nv)), mul(old, 1.0f - 3.0f / nv));
            return lerp(old, sm, f.boundary ? 0.0f : 1.0f);
        }
        P3 sm = zero3(), cr = zero3(), Erep = zero3();
        float avgS = 0.0f, cc = 0.0f, nv = 0.0f, cwRep = 0.0f;
        for (int i = 0; i < lim; ++i) {
            if (i < ri || i >= n) continue;
            const FanEntry e = acc.get(i);
            const float pw = sgn(e.ps);
            if (i == ri) {
                cwRep = sgn(e.ts);
                Erep = e.E;
                sm = add(mul(e.F, -1.0f), mul(e.PE, 4.0f));
                cr = add(zero3(), mul(e.PE, pw));
                avgS = e.ps;
                cc = pw;
                nv = 1.0f;
            } else {
                sm = add(sm, mul(e.F, -1.0f));
                sm = add(sm, mul(e.PE, 4.0f));
                nv = nv + 1.0f;
                cr = add(cr, mul(e.PE, pw));
                avgS = avgS + e.ps;
                cc = cc + pw;
            }
        }
        if (f.boundary) {
            for (int k = 0; k < lim; ++k) {
                const int i = ri - 1 - k;
                if (i < 0) break;
                const FanEntry e = acc.get(i);
                const float qw = sgn(e.ts);
                sm = add(sm, mul(e.F, -1.0f));
                sm = add(sm, mul(e.E, 4.0f));
                nv = nv + 1.0f;
                cr = add(cr, mul(e.E, qw));
                avgS = avgS + e.ts;
                cc = cc + qw;
            }
            cr = add(cr, mul(Erep, cwRep));
            cc = cc + cwRep;
            nv = nv + 1.0f;
        }
        sm = add(mul(sm, 1.0f / (nv * nv)), mul(old, 1.0f - 3.0f / nv));
        cr = add(mul(cr, 0.5f / cc), mul(old, 0.5f));
        if (cc <= 1.0f) return sm;
        if (cc >= 3.0f || nv == 2.0f) return old;
        return lerp(old, cr, sat(avgS * 0.5f));
    }
    // scatter rules: fan parameters, then per-halfedge contributions in fan
    // order (incident edges: the own edge of the first fan halfedge on a
    // boundary, and the previous edge of every fan halfedge)
    RingParams rp;
    rp.n = (float)n; rp.cc = 0.0f; rp.boundary = f.boundary; rp.s1 = rp.s2 = 0.0f; rp.e1 = rp.e2 = -1; rp.sall = 0.0f;
    if (CREASED) {
        for (int i = -1; i < lim; ++i) {
            if (i >= n || (i < 0 && !f.boundary)) continue;
            const FanEntry e = acc.get(i < 0 ? 0 : i);
            const float se = i < 0 ? e.ts : e.ps;
            const int32_t ed = i < 0 ? e.re : e.pe;
            rp.cc += sgn(se);
            if (se != 0.0f) {
                if (rp.e1 < 0) { rp.s1 = se; rp.e1 = ed; }
                else if (rp.e2 < 0) { rp.s2 = se; rp.e2 = ed; }
            }
            rp.sall += se;
        }
    }
    // contributions added from zero in ascending halfedge order, the order
    // of the reference's sequential scatter (CatmullClark.h:2345-2347,
    // :2610-2611)
    P3 sum = zero3();
    if constexpr (MAXN == 4) {
        // fans of at most four: every contribution in fan order, the rank of
        // each fan halfedge among the fan, then the sum by rank
        P3 con[4];
        int32_t key[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const FanEntry e = acc.get(i);
            con[i] = vertex_scatter_contrib(rp, old, e.E, e.PE, e.F, CREASED ? e.ts : 0.0f, CREASED ? e.ps : 0.0f,
                                            e.re, CREASED);
            key[i] = i < n ? e.r : INT32_MAX;
        }
        int rank[4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            rank[i] = (key[0] < key[i]) + (key[1] < key[i]) + (key[2] < key[i]) + (key[3] < key[i]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k >= n) break;
            const P3 ck = rank[0] == k ? con[0] : (rank[1] == k ? con[1] : (rank[2] == k ? con[2] : con[3]));
            sum = add(sum, ck);
        }
        return sum;
    }
    if constexpr (MAXN == 8) {
        // the same for fans of at most eight: the entries' points are dead
        // once the contributions are formed
        P3 con[8];
        int32_t key[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const FanEntry e = acc.get(i);
            con[i] = vertex_scatter_contrib(rp, old, e.E, e.PE, e.F, CREASED ? e.ts : 0.0f, CREASED ? e.ps : 0.0f,
                                            e.re, CREASED);
            key[i] = i < n ? e.r : INT32_MAX;
        }
        int rank[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            int r = 0;
#pragma unroll
            for (int j = 0; j < 8; ++j) r += key[j] < key[i] ? 1 : 0;
            rank[i] = r;
        }
        // rank k belongs to exactly one entry: add it there (no select
        // between entries, which would index con at run time)
#pragma unroll
        for (int k = 0; k < 8; ++k) {
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (k < n && rank[i] == k) sum = add(sum, con[i]);
        }
        return sum;
    }
    int32_t last = INT32_MIN;
    for (int k = 0; k < lim; ++k) {
        if (k >= n) break;
        int32_t bk = INT32_MAX;
        int bi = 0;
        for (int i = 0; i < lim; ++i) {
            if (i >= n) break;
            const int32_t r = acc.key(i);
            if (r > last && r < bk) { bk = r; bi = i; }
        }
        const FanEntry e = acc.get(bi);
        sum = add(sum, vertex_scatter_contrib(rp, old, e.E, e.PE, e.F, CREASED ? e.ts : 0.0f, CREASED ? e.ps : 0.0f,
                                              e.re, CREASED));
        last = bk;
    }
    return sum;
}

// The entries of a closed fan of n <= 4 halfedges r[q] and their edges
// re[q] (indices mod n): the previous edge of r[q] is the edge of r[q+1].
// Entries q >= n repeat entry 0 and are never summed.
__device__ __forceinline__ FanPreloaded closed_fan(const FusedArgs &a, const int32_t (&r)[4], const int32_t (&re)[4],
                                                   int n = 4) {
    P3 E[4], F[4];
    float s[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        E[q] = a.ep[re[q]];
        F[q] = a.fp[r[q] >> 2];
        s[q] = lsharp(a.cr, a.C, re[q]);
    }
    FanEntry e[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const bool wrap = q + 1 >= n;   // q + 1 < 4 whenever it is false
        const int n1 = (q + 1) & 3;
        e[q].r = r[q];
        e[q].re = re[q];
        e[q].pe = wrap ? re[0] : re[n1];
        e[q].E = E[q];
        e[q].PE = wrap ? E[0] : E[n1];
        e[q].F = F[q];
        e[q].ts = s[q];
        e[q].ps = wrap ? s[0] : s[n1];
    }
    return FanPreloaded{e[0], e[1], e[2], e[3]};
}

// The ring of level-d vertex v with fan f when it is a face vertex born at
// level lvl >= 2 or an interior edge vertex (born at any level):
// halfedges h[q] in fan order and their edges e[q] (the previous edge of
// h[q] is e[q+1]); false for any other fan.
__device__ __forceinline__ bool structural_ring(const Fan &f, const LevelTab &t, int d, int32_t v, int32_t (&h)[4],
                                                int32_t (&e)[4], const DCage &c) {
    const int lvl = d - (f.shift >> 1);   // the level the vertex was born at
    // (kind 1 at level 1: a cage quad with halfedges p .. p+3 in next order,
    // as marked_ring_fan checked)
    if (!(f.kind == 1 || (f.kind == 3 && !f.boundary) || (f.kind == 0 && f.n == 4 && !f.boundary && c.vring)))
        return false;
    // The ring at level lvl are children of level-(lvl-1) halfedges (the
    // cage's for lvl = 1: the cage step numbers children the same way,
    // CatmullClark.h:2693-2740), and the ring at
    // level d their children 0, 00, ...  Edges and twins follow from the
    // child numbering (CatmullClark.h:2770-2793, E the parent level's
    // edge count): child 0 of h: edge 2 edge(h) + (h > twin(h) ? 0 : 1),
    // twin 4 next(twin(h)) + 3; child 1: edge 2E + h, twin 4 next(h) + 2;
    // child 2: edge 2E + prev(h), twin 4 prev(h) + 1; child 3: edge
    // 2 edge(prev h) + (prev h > twin(prev h)), twin 4 twin(prev h).  So
    // no ring record is read.  Face ring of quad 4f': children 2 of
    // 4f'+q.  Edge ring of e = edge(p) = edge(w), w = twin(p): 4p+1,
    // 4 next(w)+3, 4w+1, 4 next(p)+3, with next from the fan (the cage's
    // next for lvl = 1).
    const int32_t Ep2 = 2 * (int32_t)t.E[lvl - 1];
    int32_t tw[4];
    if (f.kind == 0) {
        // A cage vertex with the closed ring R[q] (the cage's ring list):
        // children 0 at level 1, 4 R[q], with edge 2 edge(R) + (R > twin(R)
        // ? 0 : 1) and twin 4 next(twin(R)) + 3; the descent below needs
        // that twin outside R's quad (next(twin(R)) != R).
        int32_t R[4], w[4], en[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) R[q] = c.vring[f.a1 + q];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const CageHe x = c.he[R[q]];
            w[q] = x.twin;
            en[q] = x.edge;
        }
        bool ok = true;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t nt = c.he[w[q] >= 0 ? w[q] : R[q]].next;
            ok = ok && w[q] >= 0 && nt != R[q];
            h[q] = 4 * R[q];
            e[q] = 2 * en[q] + (R[q] > w[q] ? 0 : 1);
            tw[q] = 4 * nt + 3;
        }
        if (!ok) return false;
    } else if (f.kind == 1) {
        const int32_t p = (f.a0 - 2) >> 2;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t pp = p + ((q + 3) & 3);
            h[q] = 4 * (p + q) + 2;
            e[q] = Ep2 + pp;
            tw[q] = 4 * pp + 1;
        }
    } else {
        const int32_t p = (f.a0 - 1) >> 2, w = (f.a2 - 1) >> 2;
        const int32_t nw = (f.a1 - 3) >> 2, np = (f.a3 - 3) >> 2;
        const int32_t e2 = 2 * (v - (int32_t)t.V[lvl - 1] - (int32_t)t.F[lvl - 1]);
        h[0] = f.a0; h[1] = f.a1; h[2] = f.a2; h[3] = f.a3;
        e[0] = Ep2 + p;
        e[1] = e2 + (w > p ? 1 : 0);
        e[2] = Ep2 + w;
        e[3] = e2 + (p > w ? 1 : 0);
        tw[0] = 4 * np + 2;
        tw[1] = 4 * p;
        tw[2] = 4 * nw + 2;
        tw[3] = 4 * w;
    }
    // Down to level d through children 0: child 0 of x has edge
    // 2 edge(x) + (x > twin(x) ? 0 : 1) and twin 4 next(twin(x)) + 3.  x and
    // its twin lie in different quads, so 4x and 4 next(twin(x)) + 3 keep
    // their order (and again lie in different quads): the same edge bit
    // comes n times.
    const int n = d - lvl;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        e[q] = (e[q] << n) + (h[q] > tw[q] ? 0 : (1 << n) - 1);
        h[q] <<= 2 * n;
    }
    return true;
}

// fans of up to eight halfedges (cage vertices and faces), every load issued
// before any is used
struct FanPreloaded8 {
    FanEntry e0, e1, e2, e3, e4, e5, e6, e7;
    __device__ __forceinline__ FanEntry get(int i) const {
        return i < 4 ? pick4(e0, e1, e2, e3, i) : pick4(e4, e5, e6, e7, i - 4);
    }
    __device__ __forceinline__ int32_t key(int i) const {
        return i < 4 ? pick4(e0.r, e1.r, e2.r, e3.r, i) : pick4(e4.r, e5.r, e6.r, e7.r, i - 4);
    }
};

// The entries of a fan of n <= 8 halfedges r[q] in forward order
// (r[q+1] = twin(prev(r[q]))) and their edges re[q]: the previous edge of
// r[q] is the edge of r[q+1], so one record per entry is read and its edge
// point and sharpness serve two entries.  The last entry's previous edge is
// the first entry's edge (closed fan, indices mod n) or, on a boundary,
// plast, the edge of prev(r[n-1]).  Entries q >= n repeat entry 0 and are
// never summed.
__device__ __forceinline__ FanPreloaded8 fan8(const FusedArgs &a, const int32_t (&r)[8], const int32_t (&re)[8], int n,
                                              bool boundary, int32_t plast) {
    P3 E[8], F[8];
    float s[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        E[q] = a.ep[re[q]];
        F[q] = a.fp[r[q] >> 2];
        s[q] = lsharp(a.cr, a.C, re[q]);
    }
    const int32_t pl = boundary ? plast : re[0];
    const P3 El = a.ep[pl];
    const float sl = lsharp(a.cr, a.C, pl);
    FanEntry e[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const bool last = q + 1 >= n;   // q + 1 < 8 whenever it is false
        const int n1 = q + 1 < 8 ? q + 1 : 0;
        e[q].r = r[q];
        e[q].re = re[q];
        e[q].pe = last ? pl : re[n1];
        e[q].E = E[q];
        e[q].PE = last ? El : E[n1];
        e[q].F = F[q];
        e[q].ts = s[q];
        e[q].ps = last ? sl : s[n1];
    }
    return FanPreloaded8{e[0], e[1], e[2], e[3], e[4], e[5], e[6], e[7]};
}

// vertex point of vertex v of level d with fan f: fans of at most four (and
// with PRE8 of at most eight) load every entry up front, longer fans entry
// by entry; with PRE8 closed fans and fans of five to eight read one record
// per entry
template <int GATHER, int CREASED, int PRE8 = 0>
__device__ __forceinline__ void fan_vertex_store(const FusedArgs &a, const DCage &c, const LevelTab &t, int d,
                                                 const Fan &f, int32_t v) {
    const HeView he4{a.he};
    const P3 old = a.old[v];
    int32_t h[4], e[4];
    if (structural_ring(f, t, d, v, h, e, c)) {
        const FanPreloaded pre = closed_fan(a, h, e);
        a.vp[v] = fan_vertex_point<GATHER, CREASED, 4>(f, pre, old);
    } else if (f.kind == 1 || (f.kind == 3 && !f.boundary)) {
        // a closed fan of four in forward order: r[q+1] = twin(prev(r[q])),
        // so the previous edge of r[q] is the edge of r[q+1] and only the
        // four fan records are read
        int32_t r[4], re[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = fan_at(f, c, q);
#pragma unroll
        for (int q = 0; q < 4; ++q) re[q] = he4[r[q]].y;
        const FanPreloaded pre = closed_fan(a, r, re);
        a.vp[v] = fan_vertex_point<GATHER, CREASED, 4>(f, pre, old);
    } else if (PRE8 && !f.boundary && f.n <= 4) {
        int32_t r[4], re[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = fan_at(f, c, q < f.n ? q : 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) re[q] = he4[r[q]].y;
        a.vp[v] = fan_vertex_point<GATHER, CREASED, 4>(f, closed_fan(a, r, re, f.n), old);
    } else if (PRE8 && f.n > 4 && f.n <= 8) {
        int32_t r[8], re[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) r[q] = fan_at(f, c, q < f.n ? q : 0);
#pragma unroll
        for (int q = 0; q < 8; ++q) re[q] = he4[r[q]].y;
        const int32_t plast = f.boundary ? he4[qprev(r[f.n - 1])].y : 0;
        a.vp[v] = fan_vertex_point<GATHER, CREASED, 8>(f, fan8(a, r, re, f.n, f.boundary, plast), old);
    } else if (f.n <= 4) {
        int32_t r[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = fan_at(f, c, q < f.n ? q : 0);
        int4 rec[4], prv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            rec[q] = he4[r[q]];
            prv[q] = he4[qprev(r[q])];
        }
        FanPreloaded pre;
        pre.e0 = fan_entry(a, r[0], rec[0], prv[0]);
        pre.e1 = fan_entry(a, r[1], rec[1], prv[1]);
        pre.e2 = fan_entry(a, r[2], rec[2], prv[2]);
        pre.e3 = fan_entry(a, r[3], rec[3], prv[3]);
        a.vp[v] = fan_vertex_point<GATHER, CREASED, 4>(f, pre, old);
    } else {
        FanOnDemand od{&a, &f, &c};
        a.vp[v] = fan_vertex_point<GATHER, CREASED, 0>(f, od, old);
    }
}

}  // namespace cch
