// Shared host/device definitions of the refinement path: record layouts,
// exact per-level counts, and the scalar rules every kernel uses.
//
// Every floating-point helper reproduces the operation order of
// CatmullClark.h so that, compiled with -ffp-contract=off, results are
// bitwise equal to the reference where the reference itself is
// deterministic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cch {

// Deepest level whose IDs fit int32 for the smallest cage (3 halfedges).
constexpr int kMaxDepth = 15;

// Record layouts, byte-identical to CatmullClark.h:30-55.
struct CageHe { int32_t twin, next, prev, face, edge, vert, uv; };  // 28 B
struct Crease { int32_t next, prev; float sharp; };                 // 12 B
struct P3 { float x, y, z; };                                       // 12 B
struct UV { float u, v; };                                          //  8 B

// The level record cc_Halfedge_SemiRegular (CatmullClark.h:47-55): 16 bytes
// with the uv code, or 12 without it in a library built with CC_DISABLE_UV
// (the reference's switch, which its GL demo uses).  Kernels go through
// he_load / he_store / he_uv, which give and take the record as an int4
// whose w is the uv code (0 without uvs).
template <int BYTES>
struct HeRec;
template <>
struct alignas(16) HeRec<16> { int32_t twin, edge, vert, uv; };
template <>
struct HeRec<12> { int32_t twin, edge, vert; };
#ifdef CC_DISABLE_UV
constexpr int kHeBytes = 12;
#else
constexpr int kHeBytes = 16;
#endif
using He = HeRec<kHeBytes>;
constexpr bool kHasUv = kHeBytes == 16;

static_assert(sizeof(CageHe) == 28, "cc_Halfedge layout");
static_assert(sizeof(He) == kHeBytes, "cc_Halfedge_SemiRegular layout");
static_assert(sizeof(Crease) == 12, "cc_Crease layout");
static_assert(sizeof(P3) == 12, "cc_VertexPoint layout");

// record i of a level array as (twin, edge, vertex, uv code or 0)
__host__ __device__ __forceinline__ int4 he_load(const HeRec<16> *p, int64_t i) {
    return reinterpret_cast<const int4 *>(p)[i];
}
// element i < 2^28 of p: a 32-bit unsigned byte offset, so that a uniform p
// stays in scalar registers (global loads with a scalar base)
template <class T>
__host__ __device__ __forceinline__ const T *at32(const T *p, uint32_t i) {
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + i * (uint32_t)sizeof(T));
}
template <class T>
__host__ __device__ __forceinline__ T *at32(T *p, uint32_t i) {
    return reinterpret_cast<T *>(reinterpret_cast<char *>(p) + i * (uint32_t)sizeof(T));
}
__host__ __device__ __forceinline__ int4 he_load(const HeRec<12> *p, int64_t i) {
    const int32_t *q = reinterpret_cast<const int32_t *>(p) + 3 * i;
    return make_int4(q[0], q[1], q[2], 0);
}
// the whole record i (the uv code is dropped without uvs)
__host__ __device__ __forceinline__ void he_store(HeRec<16> *p, int64_t i, int4 r) {
    reinterpret_cast<int4 *>(p)[i] = r;
}
__host__ __device__ __forceinline__ void he_store(HeRec<12> *p, int64_t i, int4 r) {
    int32_t *q = reinterpret_cast<int32_t *>(p) + 3 * i;
    q[0] = r.x; q[1] = r.y; q[2] = r.z;
}
// twin, edge and vertex of record i, leaving its uv code
template <int B>
__host__ __device__ __forceinline__ void he_store_ids(HeRec<B> *p, int64_t i, int4 r) {
    p[i].twin = r.x; p[i].edge = r.y; p[i].vert = r.z;
}
// the uv code of record i alone (nothing without uvs)
__host__ __device__ __forceinline__ void he_store_uv(HeRec<16> *p, int64_t i, int32_t uv) { p[i].uv = uv; }
__host__ __device__ __forceinline__ void he_store_uv(HeRec<12> *, int64_t, int32_t) {}
__host__ __device__ __forceinline__ int32_t he_uv(const HeRec<16> &x) { return x.uv; }
__host__ __device__ __forceinline__ int32_t he_uv(const HeRec<12> &) { return 0; }

// Exact per-level counts (CatmullClark.h:455-585, :811-936) from the
// recurrences F' = H, E' = 2E + H, H' = 4H, C' = 2C, V' = V + E + F.
// off*[d] is the offset of level d (d >= 1) in the cumulative arrays.
struct Counts {
    int64_t H[kMaxDepth + 2], C[kMaxDepth + 2], F[kMaxDepth + 2], E[kMaxDepth + 2], V[kMaxDepth + 2];
    int64_t offH[kMaxDepth + 2], offC[kMaxDepth + 2], offV[kMaxDepth + 2];

    void init(int64_t V0, int64_t H0, int64_t E0, int64_t F0) {
        H[0] = H0; C[0] = E0; F[0] = F0; E[0] = E0; V[0] = V0;
        offH[0] = offC[0] = offV[0] = 0;
        offH[1] = offC[1] = offV[1] = 0;
        for (int d = 1; d <= kMaxDepth + 1; ++d) {
            H[d] = 4 * H[d - 1];
            C[d] = 2 * C[d - 1];
            F[d] = H[d - 1];
            E[d] = 2 * E[d - 1] + H[d - 1];
            V[d] = V[d - 1] + E[d - 1] + F[d - 1];
            if (d + 1 <= kMaxDepth + 1) {
                offH[d + 1] = offH[d] + H[d];
                offC[d + 1] = offC[d] + C[d];
                offV[d + 1] = offV[d] + V[d];
            }
        }
    }
    // cumulative counts over levels 1..d
    int64_t cumH(int d) const { return offH[d + 1]; }
    int64_t cumC(int d) const { return offC[d + 1]; }
    int64_t cumV(int d) const { return offV[d + 1]; }
};

// int32 per-level table passed by value to kernels that walk the O(depth)
// edge/vertex maps.
struct LevelTab {
    int32_t E[kMaxDepth + 1], V[kMaxDepth + 1], F[kMaxDepth + 1];
    // word offset of level L's edge-representative bits (one bit per level-L
    // halfedge, levels 1 ..) in a mesh's bit array: the words of the levels
    // before it, ceil(H / 64) each
    int32_t R[kMaxDepth + 1];
};

// Upper bound on the steps of any face or vertex walk.  Valid meshes never
// reach it; it only guarantees that corrupt topology cannot keep a wave
// running forever.
constexpr int kWalkCap = 1 << 16;

// Device view of a cage (kernel argument).
struct DCage {
    const CageHe *he;
    const int32_t *v2h, *e2h, *f2h;
    const P3 *pts;
    const UV *uvs;
    const Crease *cr;
    // Cage rings, listed once on the host (null: walk the cage instead).
    // Vertex v: the halfedges of the forward walk twin(prev(g)) from v2h[v]
    // are vring[vring_off[v] .. vring_off[v + 1]), vring_bnd[v] != 0 when the
    // walk ends on the boundary; face f: its halfedges from f2h[f] by next
    // are fring[fring_off[f] .. fring_off[f + 1]).
    const int32_t *vring_off, *vring, *vring_bnd, *fring_off, *fring;
    // every halfedge leaving vertex v, ascending: vout[vout_off[v] ..
    // vout_off[v + 1]) (null when the rings are)
    const int32_t *vout_off, *vout;
    int32_t V, U, H, E, F;
};

// Loads through explicit address spaces.  A pointer that may point to LDS
// or global memory compiles to flat instructions, which also count against
// the LDS wait counter; these keep global loads global and LDS loads LDS.
template <class T>
__device__ __forceinline__ T gld(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "whole dwords");
    T r;
    const __attribute__((address_space(1))) uint32_t *q = (const __attribute__((address_space(1))) uint32_t *)(p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) reinterpret_cast<uint32_t *>(&r)[i] = q[i];
    return r;
}
template <class T>
__device__ __forceinline__ T sld(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "whole dwords");
    T r;
    const __attribute__((address_space(3))) uint32_t *q = (const __attribute__((address_space(3))) uint32_t *)(p);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) reinterpret_cast<uint32_t *>(&r)[i] = q[i];
    return r;
}

// record i through an explicitly global load (see gld)
__device__ __forceinline__ int4 he_gld(const HeRec<16> *p, int64_t i) {
    return gld(reinterpret_cast<const int4 *>(p) + i);
}
__device__ __forceinline__ int4 he_gld(const HeRec<12> *p, int64_t i) {
    const int32_t *q = reinterpret_cast<const int32_t *>(p) + 3 * i;
    return make_int4(gld(q), gld(q + 1), gld(q + 2), 0);
}

// a level array read record by record as int4s (he_load)
struct HeView {
    const He *p;
    __host__ __device__ __forceinline__ int4 operator[](int64_t i) const { return he_load(p, i); }
};

// Contiguous block ranges per XCD: hardware hands consecutive workgroups to
// the 8 XCDs in turn; this returns the logical block so that XCD x sees
// blocks [x*q + min(x, r) ...) of a launch with nb = 8q + r blocks.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb, int remap) {
    if (!remap) return b;
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, i = b >> 3;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// ---- scalar rules (CatmullClark.h:296-392, :667-689) ----------------------
__host__ __device__ __forceinline__ int32_t qnext(int32_t h) {
    return (int32_t)(((uint32_t)h & ~3u) | (((uint32_t)h + 1u) & 3u));
}
__host__ __device__ __forceinline__ int32_t qprev(int32_t h) {
    return (int32_t)(((uint32_t)h & ~3u) | (((uint32_t)h - 1u) & 3u));
}
__host__ __device__ __forceinline__ int32_t i4(int32_t x) { return (int32_t)((uint32_t)x * 4u); }
__host__ __device__ __forceinline__ int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
__host__ __device__ __forceinline__ float fminr(float a, float b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ float fmaxr(float a, float b) { return a > b ? a : b; }
__host__ __device__ __forceinline__ float sat(float x) { return fmaxr(0.0f, fminr(x, 1.0f)); }
__host__ __device__ __forceinline__ float sgn(float x) { return x < 0.0f ? -1.0f : (x > 0.0f ? 1.0f : 0.0f); }
__host__ __device__ __forceinline__ float lerp1(float x, float y, float u) { return x + u * (y - x); }

__host__ __device__ __forceinline__ P3 add(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ P3 mul(P3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ P3 lerp(P3 a, P3 b, float u) {
    return {lerp1(a.x, b.x, u), lerp1(a.y, b.y, u), lerp1(a.z, b.z, u)};
}
__host__ __device__ __forceinline__ P3 zero3() { return {0.0f, 0.0f, 0.0f}; }

// A 16-byte record store marked non-temporal: the child records of a level
// are read again only by the next level's kernel, after gigabytes of other
// traffic, so they need not displace the points that the vertex kernel is
// about to gather.
__device__ __forceinline__ void store_stream(int4 *dst, int4 v) {
    typedef int v4i __attribute__((ext_vector_type(4)));
    const v4i w = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(w, reinterpret_cast<v4i *>(dst));
}

// UV codec (CatmullClark.h:374-392).  The encode keeps the low 32 bits of a
// 64-bit truncating conversion, which is what x86-64 compilers emit for the
// reference's float -> uint32_t conversion, so out-of-range inputs agree too.
//
// The decode's single-precision k / 65535 (k < 2^16) is evaluated as the
// product with the rounded reciprocal followed by one Newton correction with
// fused multiply-adds (the residual k - q * 65535 is exact): three
// instructions instead of a full division, and equal to the correctly
// rounded quotient for every one of the 65536 codes.
__host__ __device__ __forceinline__ float u16_unit(uint32_t k) {
    const float x = (float)k, r = 1.0f / 65535.0f;
    const float q = x * r;
    return fmaf(fmaf(-q, 65535.0f, x), r, q);
}
__host__ __device__ __forceinline__ UV uv_decode(int32_t code) {
    uint32_t t = (uint32_t)code;
    return {u16_unit(t & 0xFFFFu), u16_unit((t >> 16) & 0xFFFFu)};
}
__host__ __device__ __forceinline__ uint32_t f2u_trunc(float x) { return (uint32_t)(int64_t)x; }
__host__ __device__ __forceinline__ int32_t uv_encode(UV uv) {
    uint32_t u = f2u_trunc(uv.u * 65535.0f);
    uint32_t v = f2u_trunc(uv.v * 65535.0f);
    return (int32_t)((u & 0xFFFFu) | ((v & 0xFFFFu) << 16));
}
// uv_encode of coordinates in [0, 1] (decoded codes, their midpoints and
// face averages: every uv of a level >= 1): the products lie in [0, 65535],
// where a 32-bit truncating conversion equals the 64-bit one.
__host__ __device__ __forceinline__ int32_t uv_encode_unit(UV uv) {
    uint32_t u = (uint32_t)(int32_t)(uv.u * 65535.0f);
    uint32_t v = (uint32_t)(int32_t)(uv.v * 65535.0f);
    return (int32_t)((u & 0xFFFFu) | ((v & 0xFFFFu) << 16));
}
__host__ __device__ __forceinline__ UV uv_mid(UV a, UV b) {
    return {lerp1(a.u, b.u, 0.5f), lerp1(a.v, b.v, 0.5f)};
}

// O(depth) maps (CatmullClark.h:1206-1395): a representative halfedge of an
// edge / vertex at level d >= 1.
__host__ __device__ inline int32_t edge_rep_first(const DCage &c, int32_t e) {
    if (e >= 2 * c.E) {
        int32_t h = e - 2 * c.E;
        return imax(4 * h + 1, 4 * c.he[h].next + 2);
    } else if (e & 1) {
        return 4 * c.he[c.e2h[e >> 1]].next + 3;
    }
    return 4 * c.e2h[e >> 1] + 0;
}

__host__ __device__ inline int32_t edge_rep(const DCage &c, const LevelTab &t, int32_t e, int d) {
    uint32_t path = 1u;
    int32_t rep = 0;
    int k = d;
    for (; k > 1; --k) {
        int32_t Ep = t.E[k - 1];
        if (e >= 2 * Ep) {
            int32_t h = e - 2 * Ep;
            rep = imax(4 * h + 1, 4 * qnext(h) + 2);
            break;
        }
        path = (path << 1) | (uint32_t)(e & 1);
        e >>= 1;
    }
    if (k == 1) rep = edge_rep_first(c, e);
    for (; path > 1u; path >>= 1) rep = (path & 1u) ? 4 * qnext(rep) + 3 : 4 * rep + 0;
    return rep;
}

// Is level-d halfedge p (d >= 1) the representative edge_rep(e, d) of its
// edge e?  Compares level by level from fine to coarse and stops at the
// first mismatch, so most calls end after a few integer steps: child 0 of
// rep(e') represents the even child edge 2e', child 3 of next(rep(e')) the
// odd one 2e' + 1.
__host__ __device__ inline bool is_edge_rep(const DCage &c, const LevelTab &t, int32_t e, int d, int32_t p) {
    int k = d;
    for (; k > 1; --k) {
        const int32_t Ep = t.E[k - 1];
        if (e >= 2 * Ep) {
            const int32_t h = e - 2 * Ep;
            return p == imax(4 * h + 1, 4 * qnext(h) + 2);
        }
        if (e & 1) {
            if ((p & 3) != 3) return false;
            p = qprev(p >> 2);
        } else {
            if ((p & 3) != 0) return false;
            p >>= 2;
        }
        e >>= 1;
    }
    return edge_rep_first(c, e) == p;
}

__host__ __device__ inline int32_t vertex_rep(const DCage &c, const LevelTab &t, int32_t v, int d) {
    int shift = 0;
    int32_t rep = 0;
    int k = d;
    for (; k > 1; --k) {
        int32_t Vp = t.V[k - 1], Fp = t.F[k - 1];
        if (v >= Vp + Fp) { rep = 4 * edge_rep(c, t, v - Vp - Fp, k - 1) + 1; break; }
        if (v >= Vp) { rep = 4 * (4 * (v - Vp)) + 2; break; }
        shift += 2;
    }
    if (k == 1) {
        if (v >= c.V + c.F) rep = 4 * c.e2h[v - c.V - c.F] + 1;
        else if (v >= c.V) rep = 4 * c.f2h[v - c.V] + 2;
        else rep = 4 * c.v2h[v] + 0;
    }
    return rep << shift;
}

}  // namespace cch
