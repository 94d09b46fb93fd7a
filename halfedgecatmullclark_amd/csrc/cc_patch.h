// Patch-local level step (levels d >= 1) as device functions shared by the
// per-cage kernels and the batched kernels.
//
// A tile is NT consecutive halfedges of level d.  For d >= 4 and NT = 256 it
// is the level-d descendants of one halfedge four levels up, an 8x8 patch of
// quads; NT = 128 is half such a patch, NT = 512 two neighbouring ones.  One
// thread per halfedge:
//   1. loads its record and vertex point and stages both in LDS;
//   2. writes its four children (topology and uvIDs, CatmullClark.h
//      :2749-2796, :2878-2921) through an LDS transpose;
//   3. computes its quad's face point by DPP (:2004-2052) and, as the larger
//      halfedge of its pair, its edge's point (:2065-2241) with the twin's
//      face point from LDS when the twin is in the tile;
//   4. as the representative of its vertex (:1336), evaluates the vertex
//      rule (:2254-2615) from the LDS face and edge points when the whole
//      fan and every fan edge's owner lie in the tile; otherwise it marks
//      the vertex in a per-wave bit mask; deferred_group finishes those
//      vertices from the face and edge points in memory.
// Results equal those of the per-level kernels bit for bit: every point is
// computed by the same expressions from the same inputs.
//
// UVMODE: 0 leave uvIDs, 1 refine them, 2 write 0, 3 refine or write 0 as
// a.uv_refine says (batches mixing cages with and without uvs).
#pragma once

#include "cc_fan.h"
#include "cc_fused.h"

namespace cch {

template <int NT>
__device__ __forceinline__ bool in_tile(int32_t x, int32_t base) {
    return x >= base && x < base + NT;
}

// the halfedge that writes edge(x): the larger of the pair, or x on a boundary
__device__ __forceinline__ int32_t edge_owner(int32_t x, int32_t twin) {
    return (twin < 0 || x > twin) ? x : twin;
}

// Marks the wave's deferred vertices: one 64-bit word per wave, written by
// its first lane (a plain store; no counter is shared between waves).
__device__ __forceinline__ void defer_mark(const FusedArgs &a, bool want, int32_t h) {
    const uint64_t m = __ballot(want);
    if ((threadIdx.x & 63u) == 0 && h < a.H) a.defer_mask[h >> 6] = m;
}

// LDS of one tile: staged children, records, points, face and edge points
template <int NT, bool STAGED>
struct PatchLds {
    int4 stage[STAGED ? kStageInt4 * NT : 1];
    int4 srec[NT];
    P3 spts[NT];
    P3 sF[NT / 4];
    P3 sE[NT];
};

template <int GATHER, int CREASED, int TOPO, int UVMODE, int NT, class LP>
__device__ __forceinline__ void patch_tile(const FusedArgs &a, const DCage &c, const LevelTab &t, const LP &lp, int d,
                                           uint32_t tile) {
    constexpr bool kStaged = TOPO && (UVMODE != 0 || !kHasUv);   // whole child records
    __shared__ PatchLds<NT, kStaged> L;
    __shared__ LevelTab st;     // read after the tile's first barrier
    __shared__ LevelPtrs slp;
    lds_fill(st, t);
    lds_fill(slp, lp);
    const int32_t base = (int32_t)(tile * NT);
    const int tid = (int)threadIdx.x;
    const int32_t h = base + tid;
    const bool live = h < a.H;   // H is a multiple of 4: quads are whole
    const int j = h & 3;
    const HeView he4{a.he};
    const int4 x = live ? he4[h] : make_int4(0, 0, 0, 0);
    const int32_t twin = x.x;
    const bool eowner = live && h > twin;
    const bool twin_here = eowner && in_tile<NT>(twin, base);
    const bool twin_far = eowner && twin >= 0 && !twin_here;

    // loads of the point work first: this vertex point and a far twin quad
    const P3 P = live ? a.old[x.z] : zero3();
    const int32_t tq = twin_far ? (twin & ~3) : 0;
    const int4 tr0 = he4[tq], tr1 = he4[tq + 1], tr2 = he4[tq + 2], tr3 = he4[tq + 3];
    L.srec[tid] = x;
    L.spts[tid] = P;

    const int4 p = quad_rel4<kQuadPrev>(x);
    if (TOPO || UVMODE) {
        int4 ch[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) ch[k] = child_record(k, h, x, p, a.V, a.F, a.E);
        if (UVMODE == 1 || (UVMODE == 3 && a.uv_refine)) {
            const UV uv = uv_decode(x.w);
            const UV nuv = quad_reluv<kQuadNext>(uv), ouv = quad_reluv<kQuadOpp>(uv), puv = quad_reluv<kQuadPrev>(uv);
            UV fuv = uv;
            fuv.u += nuv.u; fuv.v += nuv.v;
            fuv.u += ouv.u; fuv.v += ouv.v;
            fuv.u += puv.u; fuv.v += puv.v;
            fuv.u /= 4.0f;
            fuv.v /= 4.0f;
            ch[0].w = x.w;
            ch[1].w = uv_encode_unit(uv_mid(uv, nuv));
            ch[2].w = uv_encode_unit(fuv);
            ch[3].w = uv_encode_unit(uv_mid(uv, puv));
        }
        if (kStaged) {
            staged_children_store<NT>(L.stage, ch, a.out_he, base, a.H);
        } else {
            __syncthreads();
            if (live) {
                if (TOPO) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) he_store_ids(a.out_he, 4 * (int64_t)h + k, ch[k]);
                } else {
#pragma unroll
                    for (int k = 0; k < 4; ++k) he_store_uv(a.out_he, 4 * (int64_t)h + k, ch[k].w);
                }
            }
        }
    } else {
        __syncthreads();
    }

    // face point of the quad, by DPP
    const P3 q0 = quad_bcast3<0>(P), q1 = quad_bcast3<1>(P), q2 = quad_bcast3<2>(P), q3 = quad_bcast3<3>(P);
    const P3 F = face_rule(q0, q1, q2, q3, GATHER);
    if (live && j == 0) {
        a.fp[h >> 2] = F;
        L.sF[tid >> 2] = F;
    }
    // edge point of the pair's larger halfedge
    if (eowner) {
        const P3 v1 = pick4(q0, q1, q2, q3, (j + 1) & 3);
        const float s = CREASED ? lsharp(a.cr, a.C, x.y) : 0.0f;
        P3 Ev;
        if (twin >= 0) {
            P3 t0, t1, t2, t3;
            if (twin_here) {
                const int tb = (twin & ~3) - base;
                t0 = L.spts[tb]; t1 = L.spts[tb + 1]; t2 = L.spts[tb + 2]; t3 = L.spts[tb + 3];
            } else {
                t0 = a.old[tr0.z]; t1 = a.old[tr1.z]; t2 = a.old[tr2.z]; t3 = a.old[tr3.z];
            }
            const P3 Ft = face_rule(t0, t1, t2, t3, GATHER);
            if (GATHER) {
                Ev = edge_gather(P, v1, F, Ft, CREASED ? sat(s) : 1.0f, CREASED);
            } else {
                const int tj = twin & 3;
                const P3 ct = edge_contrib(pick4(t0, t1, t2, t3, tj), pick4(t0, t1, t2, t3, (tj + 1) & 3), Ft, false,
                                           s, CREASED);
                Ev = add(add(zero3(), ct), edge_contrib(P, v1, F, false, s, CREASED));
            }
        } else if (GATHER) {
            Ev = edge_gather(P, v1, F, quad_face(a.he, a.old, 0, true), CREASED ? sat(s) : 0.0f, CREASED);
        } else {
            Ev = add(zero3(), edge_contrib(P, v1, F, true, s, CREASED));
        }
        a.ep[x.y] = Ev;
        L.sE[tid] = Ev;
    }
    __syncthreads();

    // vertex point of the vertex this halfedge represents
    Fan f;
    const bool owner = live && owner_fan(h, x.z, d, c, st, slp.he, f);
    bool local = owner && (f.kind == 1 || f.kind == 3);
    int32_t r[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        r[q] = local ? fan_at(f, c, q < f.n ? q : 0) : base;
        const int32_t rp = qprev(r[q]);
        local = local && in_tile<NT>(r[q], base);
        if (local) {
            local = in_tile<NT>(edge_owner(r[q], L.srec[r[q] - base].x), base) &&
                    in_tile<NT>(edge_owner(rp, L.srec[rp - base].x), base);
        }
    }
    defer_mark(a, owner && !local, h);
    if (!local) return;
    FanEntry e[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int32_t rb = r[q] - base, pb = qprev(r[q]) - base;
        const int4 rr = L.srec[rb], pr = L.srec[pb];
        e[q].r = r[q];
        e[q].re = rr.y;
        e[q].pe = pr.y;
        e[q].E = L.sE[edge_owner(r[q], rr.x) - base];
        e[q].PE = L.sE[edge_owner(qprev(r[q]), pr.x) - base];
        e[q].F = L.sF[rb >> 2];
        e[q].ts = CREASED ? lsharp(a.cr, a.C, rr.y) : 0.0f;
        e[q].ps = CREASED ? lsharp(a.cr, a.C, pr.y) : 0.0f;
    }
    FanPreloaded pre{e[0], e[1], e[2], e[3]};
    a.vp[x.z] = fan_vertex_point<GATHER, CREASED, 4>(f, pre, P);
}

// a vertex left by patch_tile, from the face and edge points in memory
template <int GATHER, int CREASED>
__device__ __forceinline__ void deferred_vertex(const FusedArgs &a, const DCage &c, const LevelTab &t,
                                                const LevelPtrs &lp, int d, int32_t h) {
    const HeView he4{a.he};
    const int32_t v = he4[h].z;
    Fan f;
    owner_fan(h, v, d, c, t, lp.he, f);
    fan_vertex_store<GATHER, CREASED>(a, c, t, d, f, v);
}

// Source mask words per wave of the deferred kernel: about a quarter of the
// halfedges represent a vertex and about a quarter of those are deferred,
// so 16 words give a wave about 64 vertices.
constexpr int kDeferGroup = 16;

// The vertices marked in mask words [w0, w0 + kDeferGroup) of `words`, one
// per lane: the masks' popcounts are prefix-summed across lanes 0..15 and
// lane l takes the l-th marked halfedge, in rounds of 64.
template <int GATHER, int CREASED, class LP>
__device__ __forceinline__ void deferred_group(const FusedArgs &a, const DCage &c, const LevelTab &tg, const LP &lpg,
                                               int d, int32_t w0, int32_t words) {
    __shared__ LevelTab t;
    __shared__ LevelPtrs lp;
    lds_fill(t, tg);
    lds_fill(lp, lpg);
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63u);
    const int32_t w = w0 + lane;
    const uint64_t m = (lane < kDeferGroup && w < words) ? a.defer_mask[w] : 0ull;
    int32_t pre = __popcll(m);
#pragma unroll
    for (int off = 1; off < kDeferGroup; off <<= 1) {
        const int32_t y = __shfl_up(pre, off);
        if (lane >= off) pre += y;
    }
    const int32_t total = __shfl(pre, kDeferGroup - 1);
    for (int32_t base = 0; base < total; base += 64) {
        const int32_t e = base + lane;
        int32_t s = 0;
#pragma unroll
        for (int j = 0; j < kDeferGroup; ++j) s += __shfl(pre, j) <= e ? 1 : 0;
        const int src = s < kDeferGroup ? s : kDeferGroup - 1;
        const int32_t before = __shfl(pre, s > 0 ? s - 1 : 0);
        uint64_t ms = __shfl(m, src);
        if (e < total) {
            const int32_t r = e - (s > 0 ? before : 0);
            for (int32_t k = 0; k < r; ++k) ms &= ms - 1ull;
            const int32_t h = (w0 + s) * 64 + (__ffsll((unsigned long long)ms) - 1);
            deferred_vertex<GATHER, CREASED>(a, c, t, lp, d, h);
        }
    }
}

}  // namespace cch
