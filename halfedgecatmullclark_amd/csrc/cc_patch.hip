// Patch-local fused level kernel (levels d >= 1).
//
// The 256 halfedges of a block are the level-d descendants of one halfedge
// four levels up: an 8x8 patch of quads (for d >= 4; smaller levels give a
// run of whole quads).  One thread per halfedge:
//   1. loads its record and vertex point and stages both in LDS;
//   2. writes its four children (topology and uvIDs, CatmullClark.h
//      :2749-2796, :2878-2921) through an LDS transpose;
//   3. computes its quad's face point by DPP (:2004-2052) and, as the larger
//      halfedge of its pair, its edge's point (:2065-2241) with the twin's
//      face point from LDS when the twin is in the block;
//   4. as the representative of its vertex (:1336), evaluates the vertex
//      rule (:2254-2615) from the LDS face and edge points when the whole
//      fan and every fan edge's owner lie in the block; otherwise it appends
//      the vertex to a list that k_level_deferred finishes from the face and
//      edge points in memory.
// Results equal those of the per-level kernels bit for bit: every point is
// computed by the same expressions from the same inputs.
#include "cc_fan.h"
#include "cc_fused.h"

namespace cch {

namespace {

constexpr int kPatch = 256;

__device__ __forceinline__ bool in_block(int32_t x, int32_t base) {
    return x >= base && x < base + kPatch;
}

// the halfedge that writes edge(x): the larger of the pair, or x on a boundary
__device__ __forceinline__ int32_t edge_owner(int32_t x, int32_t twin) {
    return (twin < 0 || x > twin) ? x : twin;
}

}  // namespace

template <int GATHER, int CREASED, int TOPO, int UVMODE, int STORE>
__global__ __launch_bounds__(256) void k_level_patch(FusedArgs a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    constexpr bool kStaged = STORE == kStoreLds && TOPO && UVMODE != 0;
    __shared__ int4 stage[kStaged ? 4 * kPatch : 1];
    __shared__ int4 srec[kPatch];
    __shared__ P3 spts[kPatch];
    __shared__ P3 sF[kPatch / 4];
    __shared__ P3 sE[kPatch];
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, a.xcd_remap);
    const int32_t base = (int32_t)(lb * kPatch);
    const int tid = (int)threadIdx.x;
    const int32_t h = base + tid;
    const bool live = h < a.H;   // H is a multiple of 4: quads are whole
    const int j = h & 3;
    const int4 *he4 = reinterpret_cast<const int4 *>(a.he);
    const int4 x = live ? he4[h] : make_int4(0, 0, 0, 0);
    const int32_t twin = x.x;
    const bool eowner = live && h > twin;
    const bool twin_here = eowner && in_block(twin, base);
    const bool twin_far = eowner && twin >= 0 && !twin_here;

    // loads of the point work first: this vertex point and a far twin quad
    const P3 P = live ? a.old[x.z] : zero3();
    const int32_t tq = twin_far ? (twin & ~3) : 0;
    const int4 tr0 = he4[tq], tr1 = he4[tq + 1], tr2 = he4[tq + 2], tr3 = he4[tq + 3];
    srec[tid] = x;
    spts[tid] = P;

    const int4 p = quad_rel4<kQuadPrev>(x);
    if (TOPO || UVMODE) {
        int4 ch[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) ch[k] = child_record(k, h, x, p, a.V, a.F, a.E);
        if (UVMODE == 1) {
            const UV uv = uv_decode(x.w);
            const UV nuv = quad_reluv<kQuadNext>(uv), ouv = quad_reluv<kQuadOpp>(uv), puv = quad_reluv<kQuadPrev>(uv);
            UV fuv = uv;
            fuv.u += nuv.u; fuv.v += nuv.v;
            fuv.u += ouv.u; fuv.v += ouv.v;
            fuv.u += puv.u; fuv.v += puv.v;
            fuv.u /= 4.0f;
            fuv.v /= 4.0f;
            ch[0].w = x.w;
            ch[1].w = uv_encode(uv_mid(uv, nuv));
            ch[2].w = uv_encode(fuv);
            ch[3].w = uv_encode(uv_mid(uv, puv));
        }
        int4 *out4 = reinterpret_cast<int4 *>(a.out_he);
        if (kStaged) {
#pragma unroll
            for (int k = 0; k < 4; ++k) stage[4 * tid + k] = ch[k];
            __syncthreads();
            const int64_t obase = 4 * (int64_t)base;
            const int64_t limit = 4 * (int64_t)a.H;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int64_t o = obase + tid + (int64_t)k * kPatch;
                if (o < limit) out4[o] = stage[tid + k * kPatch];
            }
        } else {
            __syncthreads();
            if (live) {
                if (TOPO && (UVMODE == 1 || UVMODE == 2)) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) out4[4 * (int64_t)h + k] = ch[k];
                } else if (TOPO) {
                    He *o = a.out_he + 4 * (int64_t)h;
#pragma unroll
                    for (int k = 0; k < 4; ++k) { o[k].twin = ch[k].x; o[k].edge = ch[k].y; o[k].vert = ch[k].z; }
                } else {
                    He *o = a.out_he + 4 * (int64_t)h;
#pragma unroll
                    for (int k = 0; k < 4; ++k) o[k].uv = ch[k].w;
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
        sF[tid >> 2] = F;
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
                t0 = spts[tb]; t1 = spts[tb + 1]; t2 = spts[tb + 2]; t3 = spts[tb + 3];
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
        sE[tid] = Ev;
    }
    __syncthreads();

    // vertex point of the vertex this halfedge represents
    if (!live) return;
    Fan f;
    if (!owner_fan(h, x.z, d, c, t, lp.he, f)) return;
    bool local = f.kind == 1 || f.kind == 3;
    int32_t r[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        r[q] = local ? fan_at(f, c, q < f.n ? q : 0) : base;
        const int32_t rp = qprev(r[q]);
        local = local && in_block(r[q], base);
        if (local) {
            local = in_block(edge_owner(r[q], srec[r[q] - base].x), base) &&
                    in_block(edge_owner(rp, srec[rp - base].x), base);
        }
    }
    if (!local) {
        a.defer[atomicAdd(a.defer_count, 1)] = h;
        return;
    }
    FanEntry e[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int32_t rb = r[q] - base, pb = qprev(r[q]) - base;
        const int4 rr = srec[rb], pr = srec[pb];
        e[q].re = rr.y;
        e[q].pe = pr.y;
        e[q].E = sE[edge_owner(r[q], rr.x) - base];
        e[q].PE = sE[edge_owner(qprev(r[q]), pr.x) - base];
        e[q].F = sF[rb >> 2];
        e[q].ts = CREASED ? lsharp(a.cr, a.C, rr.y) : 0.0f;
        e[q].ps = CREASED ? lsharp(a.cr, a.C, pr.y) : 0.0f;
    }
    FanPreloaded pre{e[0], e[1], e[2], e[3]};
    a.vp[x.z] = fan_vertex_point<GATHER, CREASED, 4>(f, pre, P);
}

// vertices left by k_level_patch, from the face and edge points in memory
template <int GATHER, int CREASED>
__global__ __launch_bounds__(256) void k_level_deferred(FusedArgs a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    const int32_t count = *a.defer_count;
    const int4 *he4 = reinterpret_cast<const int4 *>(a.he);
    for (int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x); i < count; i += (int32_t)(gridDim.x * blockDim.x)) {
        const int32_t h = a.defer[i];
        const int32_t v = he4[h].z;
        Fan f;
        owner_fan(h, v, d, c, t, lp.he, f);
        const P3 old = a.old[v];
        if (f.kind == 1 || f.kind == 3) {
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
}

#define CC_PATCH_CASE(G, C, T, U, S)                                                                      \
    if (gather == G && creased == C && topo == T && uvMode == U && store == S) {                         \
        hipLaunchKernelGGL((k_level_patch<G, C, T, U, S>), grid, block, 0, st, a, c, t, lp, d);           \
        return;                                                                                          \
    }

void launch_level_patch(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d, int gather,
                        int creased, int topo, int uvMode, int store, hipStream_t st) {
    const dim3 block(kPatch);
    const dim3 grid((unsigned)((a.H + kPatch - 1) / kPatch));
    gather = gather ? 1 : 0;
    creased = creased ? 1 : 0;
    topo = topo ? 1 : 0;
#define CC_PATCH_S(G, C, T, U) CC_PATCH_CASE(G, C, T, U, kStoreDirect) CC_PATCH_CASE(G, C, T, U, kStoreLds)
#define CC_PATCH_U(G, C, T) CC_PATCH_S(G, C, T, 0) CC_PATCH_S(G, C, T, 1) CC_PATCH_S(G, C, T, 2)
#define CC_PATCH_T(G, C) CC_PATCH_U(G, C, 0) CC_PATCH_U(G, C, 1)
    CC_PATCH_T(0, 0) CC_PATCH_T(0, 1) CC_PATCH_T(1, 0) CC_PATCH_T(1, 1)
}

void launch_level_deferred(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d,
                           int gather, int creased, int64_t max_count, hipStream_t st) {
    // a grid that covers the largest possible list, capped; the kernel loops
    int64_t blocks = (max_count + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    const dim3 grid((unsigned)blocks), block(256);
    if (gather && creased) hipLaunchKernelGGL((k_level_deferred<1, 1>), grid, block, 0, st, a, c, t, lp, d);
    else if (gather) hipLaunchKernelGGL((k_level_deferred<1, 0>), grid, block, 0, st, a, c, t, lp, d);
    else if (creased) hipLaunchKernelGGL((k_level_deferred<0, 1>), grid, block, 0, st, a, c, t, lp, d);
    else hipLaunchKernelGGL((k_level_deferred<0, 0>), grid, block, 0, st, a, c, t, lp, d);
}

}  // namespace cch
