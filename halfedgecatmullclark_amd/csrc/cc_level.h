// Bodies of the fused level kernel and of the compacted vertex-point kernel
// (levels d >= 1) as device functions of the logical block index, shared by
// the per-cage kernels (cc_fused.hip) and the batched kernels (cc_batch.hip).
//
// Fused: one thread per parent halfedge writes its four children
// (CatmullClark.h:2749-2796, uvs :2878-2921) and the face point of its quad
// (:2004-2052, by DPP) and, as the larger halfedge of its pair, its edge's
// point (:2065-2241).  UVMODE: 0 leave uvIDs, 1 refine, 2 write 0, 3 refine
// or write 0 as a.uv_refine says.
#pragma once

#include "cc_fan.h"
#include "cc_fused.h"

namespace cch {

// This halfedge's side of its edge, which the lanes of the twin's quad read
// from LDS when the twin owns the edge: the face point of h's quad (gather
// rules) or h's own scatter contribution to the edge point, with v1 the
// point of the next halfedge's vertex and s the edge's sharpness.
template <int GATHER, int CREASED>
__device__ __forceinline__ P3 edge_side(P3 P, P3 v1, P3 F, float s) {
    return GATHER ? F : edge_contrib(P, v1, F, false, s, CREASED);
}

// face point F of h's quad (stored by lane 0 of the quad) and, for the
// larger halfedge of each pair, the edge point Ev; mine: this lane's
// edge_side.  The twin's side: SIDES, spts holds every lane's edge_side and
// a twin inside the block gives its own; else spts holds every lane's vertex
// point and a twin quad inside the block gives its corners, from which its
// face point follows as for a twin quad outside the block (g2, g3: the
// points of its two corners away from the edge, loaded by the caller).
template <int GATHER, int CREASED, bool SIDES>
__device__ __forceinline__ void level_fused_points(const FusedArgs &a, int32_t h, int4 x, P3 P, P3 v1, P3 F, P3 mine,
                                                   bool owner, bool twin_here, int32_t base, const P3 *spts, P3 g2,
                                                   P3 g3, float s, P3 &Ev) {
    const int32_t twin = x.x;
    if ((h & 3) == 0) *at32(a.fp + (base >> 2), threadIdx.x >> 2) = F;   // quad h / 4 (base: a multiple of 256)
    if (!owner) return;
    // this halfedge owns its edge: both sides
    if (twin >= 0) {
        P3 Ft = zero3(), ct = zero3();   // the twin quad's face point (gather), the twin's contribution (scatter)
        if (SIDES && twin_here) {
            const P3 o = sld(spts + (twin - base));
            if (GATHER) Ft = o;
            else ct = o;
        } else {
            // The twin quad's corners, from the twin on: the twin's vertex
            // is the head of h (v1), the twin's next vertex is h's own (P),
            // the other two come from LDS when the quad is in this block
            // (the reads are valid for every lane; selecting afterwards
            // keeps LDS and global reads in their own instructions), else
            // from the loads above.  Corner k of the twin quad is corner
            // (k - tj) & 3 counted from the twin: rotate by tj in two steps
            // (by 1, then by 2).
            const int tj = twin & 3;
            P3 o2 = g2, o3 = g3;
            if (!SIDES) {
                const int tb = twin_here ? (twin & ~3) - base : 0;
                const P3 l2 = sld(spts + tb + ((tj + 2) & 3)), l3 = sld(spts + tb + ((tj + 3) & 3));
                o2 = twin_here ? l2 : g2;
                o3 = twin_here ? l3 : g3;
            }
            const bool r1 = (tj & 1) != 0, r2 = (tj & 2) != 0;
            const P3 y0 = r1 ? o3 : v1, y1 = r1 ? v1 : P, y2 = r1 ? P : o2, y3 = r1 ? o2 : o3;
            Ft = face_rule(r2 ? y2 : y0, r2 ? y3 : y1, r2 ? y0 : y2, r2 ? y1 : y3, GATHER);
            if (!GATHER) ct = edge_contrib(v1, P, Ft, false, s, CREASED);
        }
        if (GATHER) {
            const float w = CREASED ? sat(s) : 1.0f;
            Ev = edge_gather(P, v1, F, Ft, w, CREASED);
        } else {
            Ev = add(add(zero3(), ct), mine);
        }
    } else {
        if (GATHER) {
            const float w = CREASED ? sat(s) : 0.0f;
            Ev = edge_gather(P, v1, F, quad_face(a.he, a.old, 0, true), w, CREASED);
        } else {
            Ev = add(zero3(), edge_contrib(P, v1, F, true, s, CREASED));
        }
    }
    a.ep[x.y] = Ev;
}

// Vertex points of the face vertices born at level d >= 2, inside the fused
// kernel.  The face vertex of level-(d-1) quad Q (halfedges p = 4Q .. 4Q+3)
// has the ring of children 2 of p + q, level-d halfedges 16Q + 4q + 2: one
// row of 16 lanes.  Their quads' face points are in the ring lanes; the edge
// of ring halfedge q, 2E' + prev(p + q) (E' the level-(d-1) edge count),
// is owned by the ring lane itself for q = 1, 2, 3 and by lane 16Q + 13
// (child 1 of p + 3, its twin) for q = 0.  Lane 16Q + 2, the vertex's
// representative, gathers the four face points and four edge points by row
// rotations and evaluates the masked vertex kernel's own function on the
// entries that kernel would load (sharpness 0: edges from 2E' on have no
// crease record, 2E' >= C), so the point is the same bit for bit.  Every
// lane of the wave runs the rotations; a row is live or dead as a whole
// (H is a multiple of 16 from level 2 on).  The scatter rules add one
// contribution per ring halfedge from zero in ascending order: each ring
// lane computes its own (the ring's parameters are those of a smooth
// valence-4 vertex), and lane 16Q + 2 adds the four.
template <int GATHER, int CREASED>
__device__ __forceinline__ void face_vertex_in_row(const FusedArgs &a, const LevelTab &t, int d, int32_t h,
                                                   int32_t base, bool live, P3 old, P3 F, P3 E) {
    // the vertex point of parent-level quad h / 16, from a block-uniform base
    P3 *const vps = a.vp + ((int32_t)t.V[d - 1] + (base >> 4));
    if (!GATHER) {
        // ring lane q (offset 4q + 2): its edge point (offset 2: from 13)
        // and the previous edge's, the next ring lane's (offset 14: from 13)
        const int off = (int)(h & 15);
        const P3 E5 = row_ror3<5>(E), E12 = row_ror3<12>(E), E1 = row_ror3<1>(E);
        const P3 Eown = off == 2 ? E5 : E, PE = off == 14 ? E1 : E12;
        RingParams rp;
        rp.n = 4.0f; rp.cc = 0.0f; rp.boundary = false; rp.s1 = rp.s2 = 0.0f; rp.e1 = rp.e2 = -1; rp.sall = 0.0f;
        const int32_t p = (h >> 4) << 2;
        const int32_t re = 2 * (int32_t)t.E[d - 1] + p + ((((h >> 2) & 3) + 3) & 3);
        const P3 c = vertex_scatter_contrib(rp, old, Eown, PE, F, 0.0f, 0.0f, re, CREASED);
        const P3 c1 = row_ror3<12>(c), c2 = row_ror3<8>(c), c3 = row_ror3<4>(c);
        if (!live || off != 2) return;
        P3 sum = add(zero3(), c);
        sum = add(sum, c1);
        sum = add(sum, c2);
        sum = add(sum, c3);
        *at32(vps, threadIdx.x >> 4) = sum;
        return;
    }
    const P3 F1 = row_ror3<12>(F), F2 = row_ror3<8>(F), F3 = row_ror3<4>(F);
    const P3 E0 = row_ror3<5>(E), E1 = row_ror3<12>(E), E2 = row_ror3<8>(E), E3 = row_ror3<4>(E);
    if (!live || (h & 15) != 2) return;
    const int32_t p = (h - 2) >> 2;
    const int32_t Ep2 = 2 * (int32_t)t.E[d - 1];
    const P3 Fq[4] = {F, F1, F2, F3}, Eq[4] = {E0, E1, E2, E3};
    FanEntry en[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int n1 = (q + 1) & 3;
        en[q].r = h + 4 * q;
        en[q].re = Ep2 + p + ((q + 3) & 3);
        en[q].pe = Ep2 + p + ((n1 + 3) & 3);
        en[q].E = Eq[q];
        en[q].PE = Eq[n1];
        en[q].F = Fq[q];
        en[q].ts = 0.0f;
        en[q].ps = 0.0f;
    }
    Fan f;
    f.kind = 1; f.shift = 0; f.n = 4; f.ri = 0; f.boundary = false;
    f.a0 = h; f.a1 = f.a2 = f.a3 = 0;
    *at32(vps, threadIdx.x >> 4) = fan_vertex_point<GATHER, CREASED, 4>(f, FanPreloaded{en[0], en[1], en[2], en[3]}, old);
}

// uvIDs of children 1, 2, 3 of a halfedge from its own, the next, the
// opposite and the previous halfedge's uvs (:2878-2921; the face average sums
// them in that order, :2904-2913); child 0 keeps the halfedge's own code
__device__ __forceinline__ void child_uv_codes(UV uv, UV nuv, UV ouv, UV puv, int32_t &c1, int32_t &c2, int32_t &c3) {
    UV fuv = uv;
    fuv.u += nuv.u; fuv.v += nuv.v;
    fuv.u += ouv.u; fuv.v += ouv.v;
    fuv.u += puv.u; fuv.v += puv.v;
    fuv.u /= 4.0f;
    fuv.v /= 4.0f;
    c1 = uv_encode_unit(uv_mid(uv, nuv));
    c2 = uv_encode_unit(fuv);
    c3 = uv_encode_unit(uv_mid(uv, puv));
}

template <int GATHER, int CREASED, int TOPO, int UVMODE, int POINTS, int STORE, int NT>
__device__ __forceinline__ void level_fused_block(const FusedArgs &a, const DCage &cg, const LevelTab &tb, int d,
                                                  uint32_t lb) {
    // whole child records written (a library without uvs has nothing else)
    constexpr bool kWhole = TOPO && (UVMODE != 0 || !kHasUv);
    constexpr bool kStaged = STORE == kStoreLds && kWhole;
    __shared__ int4 stage[kStaged ? kStageInt4 * NT : 1];
    __shared__ P3 spts[POINTS ? NT : 1];
    __shared__ LevelTab stab;   // for the owner test (read after the block's barrier)
    if (POINTS && a.owner_mask) lds_fill(stab, tb);
    if (POINTS && a.rare && lb == 0 && threadIdx.x == 0) a.rare[0] = 0;
    const int32_t base = (int32_t)(lb * NT);   // NT threads per block, one tile of NT halfedges
    const int32_t h = base + (int32_t)threadIdx.x;
    const bool live = h < a.H;   // H is a multiple of 4: quads are whole
    const HeView he4{a.he};
    // the block's records from a block-uniform base, lanes at small offsets
    const int4 x = live ? he_load(at32(a.he + base, threadIdx.x), 0) : make_int4(0, 0, 0, 0);
    // this level's edge-representative bits of the 64 halfedges of the wave
    const uint64_t rw = (a.erep_write && d >= 2 && live) ? *at32(a.erep + (tb.R[d] + (base >> 6)), threadIdx.x >> 6) : 0ull;
    const int32_t twin = x.x;
    const bool owner = live && h > twin;   // this halfedge writes its edge point
    // twins inside this block give their side of the edge (or their quad's
    // corners) through LDS, the others' quads come from memory
    const bool twin_here = owner && twin >= base && twin < base + NT;
    // Without topology work the sides go to LDS in the prologue, and the
    // barrier before the point work follows at once.  With it, computing
    // them before the child stores lengthens the path to those stores (the
    // level-7 fused kernel 2.43 -> 2.62 ms), so LDS holds the corner points.
    constexpr bool kSides = POINTS && !TOPO && !UVMODE;
    const bool twin_far = owner && twin >= 0 && !twin_here;

    // the loads of the point work come first: this halfedge's vertex point,
    // its edge's sharpness and, for a twin quad outside the block, the
    // vertex IDs of its two corners away from the shared edge and then their
    // points, so that the latency of the second gather overlaps the topology
    // stores below (the empty asm that reads the IDs keeps their wait inside
    // the branch).  With kSides this halfedge's side of its edge then goes
    // to LDS; else its vertex point, between the two gathers: the wait for
    // the IDs covers it, and no later wait has to count the twin-quad point
    // loads still in flight (memory loads complete in order).
    P3 P = zero3(), g2 = zero3(), g3 = zero3(), v1 = zero3(), F = zero3(), mine = zero3();
    float s = 0.0f;
    if (POINTS) {
        P = a.old[x.z];   // a dead lane's record is all zeros: vertex 0
        if (CREASED) {   // both halves of an edge with kSides, else its owner
            const bool has = (kSides ? live : owner) && x.y < a.C;   // a level's crease array is never empty
            const float cs = a.cr[has ? x.y : 0].sharp;
            s = has ? cs : 0.0f;
        }
        int32_t v2 = 0, v3 = 0;
        if (twin_far) {
            const int32_t tq = twin & ~3, tj = twin & 3;
            v2 = he4[tq + ((tj + 2) & 3)].z;
            v3 = he4[tq + ((tj + 3) & 3)].z;
            asm volatile("" ::"v"(v2), "v"(v3));
        }
        if (kSides) {
            if (twin_far) {
                g2 = gld(a.old + v2); g3 = gld(a.old + v3);
            }
            v1 = P3{quad_relf<kQuadNext>(P.x), quad_relf<kQuadNext>(P.y), quad_relf<kQuadNext>(P.z)};
            F = face_rule(quad_bcast3<0>(P), quad_bcast3<1>(P), quad_bcast3<2>(P), quad_bcast3<3>(P), GATHER);
            mine = edge_side<GATHER, CREASED>(P, v1, F, s);
            spts[threadIdx.x] = mine;
        } else {
            spts[threadIdx.x] = P;
            if (twin_far) {
                g2 = gld(a.old + v2); g3 = gld(a.old + v3);
            }
        }
    }

    // record of the previous halfedge of the quad, by DPP
    const int4 p = quad_rel4<kQuadPrev>(x);

    if (TOPO || UVMODE) {
        int4 c[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] = child_record(k, h, x, p, a.V, a.F, a.E);
        if (UVMODE == 1 || (UVMODE == 3 && a.uv_refine)) {
            // uvs of this, the next, the opposite and the previous halfedge;
            // the face average sums them in that order (:2904-2913)
            const UV uv = uv_decode(x.w);
            const UV nuv = quad_reluv<kQuadNext>(uv), ouv = quad_reluv<kQuadOpp>(uv), puv = quad_reluv<kQuadPrev>(uv);
            c[0].w = x.w;
            child_uv_codes(uv, nuv, ouv, puv, c[1].w, c[2].w, c[3].w);
        }
        if (kStaged) {
            staged_children_store<NT>(stage, c, a.out_he, base, a.H);
        } else if (live) {
            if (kWhole) {
#pragma unroll
                for (int k = 0; k < 4; ++k) he_store(a.out_he, 4 * (int64_t)h + k, c[k]);
            } else if (TOPO) {
#pragma unroll
                for (int k = 0; k < 4; ++k) he_store_ids(a.out_he, 4 * (int64_t)h + k, c[k]);
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) he_store_uv(a.out_he, 4 * (int64_t)h + k, c[k].w);
            }
        }
    }

    if (a.erep_write) {
        // Edge-representative bits of the four children from this
        // halfedge's and its quad predecessor's (CatmullClark.h:1206-1262
        // through the child numbering): child 0 when h > twin(h) and h
        // represents its edge, child 1 when h is the quad's last halfedge,
        // child 2 unless it is the first, child 3 when prev(h) > twin(prev(h))
        // and prev(h) represents its edge.  Level 1 takes its own bits from
        // the cage maps.
        bool rh, rp;
        if (d == 1) {
            rh = live && edge_rep_first(cg, x.y) == h;
            rp = quad_rel<kQuadPrev>(rh ? 1 : 0) != 0;
            const uint64_t m1 = __ballot(rh);
            if ((threadIdx.x & 63u) == 0 && live) *at32(a.erep + (tb.R[1] + (base >> 6)), threadIdx.x >> 6) = m1;
        } else {
            rh = ((rw >> (h & 63)) & 1ull) != 0ull;
            rp = ((rw >> (qprev(h) & 63)) & 1ull) != 0ull;
        }
        const uint32_t nib = live ? ((h > x.x && rh ? 1u : 0u) | ((h & 3) == 3 ? 2u : 0u) | ((h & 3) != 0 ? 4u : 0u) |
                                     (qprev(h) > p.x && rp ? 8u : 0u))
                                  : 0u;
        // children 4h .. 4h+3 are bits 4 (h & 15) .. + 3 of word h >> 4: a
        // DPP row of 16 lanes fills one word
        const int lane = (int)(threadIdx.x & 63u);
        const uint32_t sh = 4u * (uint32_t)(lane & 7);
        const uint32_t lo = row_or((lane & 8) ? 0u : nib << sh);
        const uint32_t hi = row_or((lane & 8) ? nib << sh : 0u);
        if ((lane & 15) == 15 && (h & ~15) < a.H)
            *at32(a.erep + (tb.R[d + 1] + (base >> 4)), threadIdx.x >> 4) = ((uint64_t)hi << 32) | lo;
    }

    if (!POINTS) return;
    if (!kStaged) __syncthreads();
    if (!kSides) {
        v1 = P3{quad_relf<kQuadNext>(P.x), quad_relf<kQuadNext>(P.y), quad_relf<kQuadNext>(P.z)};
        F = face_rule(quad_bcast3<0>(P), quad_bcast3<1>(P), quad_bcast3<2>(P), quad_bcast3<3>(P), GATHER);
        mine = edge_side<GATHER, CREASED>(P, v1, F, s);
    }
    P3 Ev = zero3();
    if (live)
        level_fused_points<GATHER, CREASED, kSides>(a, h, x, P, v1, F, mine, owner, twin_here, base, spts, g2, g3, s,
                                                    Ev);
    if (a.owner_mask) {
        // the face vertices born at this level are finished here: their
        // representatives (children 2 of a quad's first halfedge, h % 16 = 2)
        // leave no bit for the vertex kernel
        const bool face_here = d >= 2;
        if (face_here) face_vertex_in_row<GATHER, CREASED>(a, tb, d, h, base, live, P, F, Ev);
        // one bit per halfedge: is it the representative of its vertex?  Last,
        // so that the cage lookups of a few lanes overlap the wave's stores
        if (a.owner_write) {
            const bool skip = face_here && (h & 15) == 2;
            const uint64_t m =
                __ballot(live && !skip && owner_test(h, x.z, d, cg, stab, a.erep_read ? a.erep : nullptr));
            if ((threadIdx.x & 63u) == 0 && live) *at32(a.owner_mask + (base >> 6), threadIdx.x >> 6) = m;
        }
    }
}


// Vertex points with owners compacted per block: a block scans
// 256 * SCAN halfedges, collects the representatives in LDS, then every
// thread evaluates owners, so all lanes do ring work.
template <int GATHER, int CREASED, int SCAN, class LP>
__device__ __forceinline__ void level_vertices_compact_block(const FusedArgs &a, const DCage &c, const LevelTab &t,
                                                             const LP &lp, int d, uint32_t lb) {
    __shared__ int32_t owners[256 * SCAN];
    __shared__ int32_t nown;
    __shared__ LevelTab st;
    __shared__ LevelPtrs slp;
    lds_fill(st, t);
    lds_fill(slp, lp);
    const int32_t base = (int32_t)(lb * 256 * SCAN);
    const HeView he4{a.he};
    if (threadIdx.x == 0) nown = 0;
    __syncthreads();
    int32_t vs[SCAN];
#pragma unroll
    for (int k = 0; k < SCAN; ++k) {
        const int32_t h = base + k * 256 + (int32_t)threadIdx.x;
        vs[k] = h < a.H ? he4[h].z : -1;
    }
#pragma unroll
    for (int k = 0; k < SCAN; ++k) {
        const int32_t h = base + k * 256 + (int32_t)threadIdx.x;
        if (vs[k] >= 0 && owner_test(h, vs[k], d, c, st, a.erep_read ? a.erep : nullptr)) owners[atomicAdd(&nown, 1)] = h;
    }
    __syncthreads();
    const int32_t count = nown;
    for (int32_t i = (int32_t)threadIdx.x; i < count; i += 256) {
        const int32_t h = owners[i];
        const int32_t v = he4[h].z;
        Fan f;
        owner_fan(h, v, d, c, st, slp.he, f);
        fan_vertex_store<GATHER, CREASED>(a, c, st, d, f, v);
    }
}

// Owners per wave of the masked vertex kernel: kOwnerWords mask words of 64
// halfedges, about a quarter of which represent a vertex.
constexpr int kOwnerWords = 4;

// Calls fn(h) once per bit set in mask words [w0, w0 + GW) of `words`, one
// halfedge per lane in ascending order, in rounds of 64.  The wave lists the
// marked halfedges in `list` (its own 64 GW slots of LDS): lane l of word w
// writes w * 64 + l at the count of set bits before it (mbcnt).
template <int GW, class Fn>
__device__ __forceinline__ void mask_group(const uint64_t *mask, int32_t w0, int32_t words, int32_t *list,
                                           const Fn &fn) {
    const int lane = (int)(threadIdx.x & 63u);
    int32_t total = 0;
#pragma unroll
    for (int j = 0; j < GW; ++j) {
        const int32_t w = w0 + j;
        const uint64_t m = w < words ? mask[w] : 0ull;   // the same word in every lane
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)m);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(m >> 32));
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
        if (((lane < 32 ? lo >> lane : hi >> (lane - 32)) & 1u) != 0u) list[total + (int32_t)rank] = w * 64 + lane;
        total += __popc(lo) + __popc(hi);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int32_t i = lane; i < total; i += 64) fn(list[i]);
}

// A marked owner whose ring follows from the child numbering: its vertex
// point; any other goes to the rare list (no general fan code here, so the
// kernel keeps fewer registers and runs more waves).
template <int GATHER, int CREASED>
struct RingVertexFn {
    const FusedArgs &a;
    const DCage &c;
    const LevelTab &t;
    const LevelPtrs &lp;
    int d;
    __device__ void operator()(int32_t h) const {
        Fan f;
        int32_t v, r[4], e[4];
        if (!marked_ring_fan(h, d, c, t, lp.he, f, v) || !structural_ring(f, t, d, v, r, e, c)) {
            const int32_t slot = atomicAdd(a.rare, 1);
            if (slot < a.rare_cap) a.rare[1 + slot] = h;
            return;
        }
        const FanPreloaded pre = closed_fan(a, r, e);
        a.vp[v] = fan_vertex_point<GATHER, CREASED, 4>(f, pre, a.old[v]);
    }
};

// vertex points of the owners marked by the fused kernel; `wave` is the
// index of this wave among the level's waves (kOwnerWords words each)
template <int GATHER, int CREASED, class LP>
__device__ __forceinline__ void level_vertices_masked(const FusedArgs &a, const DCage &c, const LevelTab &t,
                                                      const LP &lp, int d, int32_t wave) {
    __shared__ LevelTab st;
    __shared__ LevelPtrs slp;
    __shared__ int32_t list[4][64 * kOwnerWords];   // four waves per workgroup
    lds_fill(st, t);
    lds_fill(slp, lp);
    __syncthreads();
    const int32_t words = (a.H + 63) >> 6;
    mask_group<kOwnerWords>(a.owner_mask, wave * kOwnerWords, words, list[threadIdx.x >> 6],
                            RingVertexFn<GATHER, CREASED>{a, c, st, slp, d});
}

// the vertices the ring-only masked kernel listed, one thread each
template <int GATHER, int CREASED, class LP>
__device__ __forceinline__ void level_vertices_rare(const FusedArgs &a, const DCage &c, const LevelTab &t,
                                                    const LP &lp, int d, int32_t i) {
    __shared__ LevelTab st;
    __shared__ LevelPtrs slp;
    lds_fill(st, t);
    lds_fill(slp, lp);
    __syncthreads();
    const int32_t n = a.rare[0] < a.rare_cap ? a.rare[0] : a.rare_cap;
    if (i >= n) return;
    const int32_t h = a.rare[1 + i];
    Fan f;
    const int32_t v = marked_fan(h, d, c, st, slp.he, f);
    fan_vertex_store<GATHER, CREASED, 1>(a, c, st, d, f, v);
}

}  // namespace cch
