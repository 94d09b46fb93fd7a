// Fused per-level kernels for gfx950 (levels d >= 1, where every face is a
// quad and halfedges 4f..4f+3 form face f).
//
// k_level_fused: one thread per parent halfedge h of level d.  It
//   - writes the four child records 4h..4h+3 of level d+1 (topology and
//     uvIDs, CatmullClark.h:2749-2796 and :2878-2921), staged through LDS so
//     that every store instruction of a wave covers 1 KiB of contiguous HBM;
//   - computes the face point of its quad from the four lanes of the quad
//     (DPP quad permutes, no memory round trip), written by lane 0 of the
//     quad (:2004-2052);
//   - computes the edge point of its edge when it is the larger halfedge of
//     the pair (or a boundary halfedge): its own contribution plus the twin's,
//     whose face point is recomputed from the twin quad (:2065-2241).
// k_level_vertices: one thread per halfedge; the halfedge that is the
//   reference's representative of its vertex (CatmullClark.h:1336) walks the
//   vertex ring over the face and edge points k_level_fused stored and writes
//   the vertex point (:2254-2615).  Scatter rules are evaluated per halfedge
//   of the ring and summed in ring order (deterministic, no atomics).
//
// Blocks are remapped so that each XCD processes one contiguous range of
// the level, keeping neighbouring quads in the same L2.
#include "cc_level.h"

namespace cch {

// NT threads per workgroup, one tile of NT halfedges each; MULTI: each
// workgroup processes `tiles` consecutive tiles in turn (the barrier between
// tiles keeps one tile's LDS reads ahead of the next tile's writes).  The
// loop is a kernel of its own: keeping the arguments live across iterations
// takes about 45 more scalar and 17 more vector registers, which at 1024
// threads leaves room for one workgroup per CU instead of two.
template <int GATHER, int CREASED, int TOPO, int UVMODE, int POINTS, int STORE, int NT, int MULTI>
__global__ __launch_bounds__(NT) void k_level_fused(FusedArgs a, DCage c, LevelTab t, int d, int tiles) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, a.xcd_remap);
    if constexpr (!MULTI) {
        level_fused_block<GATHER, CREASED, TOPO, UVMODE, POINTS, STORE, NT>(a, c, t, d, lb);
    } else {
        for (int k = 0; k < tiles; ++k) {
            if (k) __syncthreads();
            level_fused_block<GATHER, CREASED, TOPO, UVMODE, POINTS, STORE, NT>(a, c, t, d,
                                                                              lb * (uint32_t)tiles + (uint32_t)k);
        }
    }
}

// ring vertices only; the others to the rare list, then k_level_vertices_rare
// (W waves per SIMD: 8 fit without spilling, 56 registers)
template <int GATHER, int CREASED, int W>
__global__ __launch_bounds__(256, W) void k_level_vertices_ring(FusedArgs a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, a.xcd_remap);
    level_vertices_masked<GATHER, CREASED>(a, c, t, lp, d, (int32_t)(lb * 4 + (threadIdx.x >> 6)));
}

template <int GATHER, int CREASED>
__global__ __launch_bounds__(256) void k_level_vertices_rare(FusedArgs a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    level_vertices_rare<GATHER, CREASED>(a, c, t, lp, d, (int32_t)(blockIdx.x * 256 + threadIdx.x));
}

template <int GATHER, int CREASED>
__global__ __launch_bounds__(256) void k_level_vertices(FusedArgs a, DCage c, LevelTab t, int d) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, a.xcd_remap);
    const int32_t h = (int32_t)(lb * blockDim.x + threadIdx.x);
    if (h >= a.H) return;
    const He *he = a.he;
    const int32_t v = he[h].vert;
    if (vertex_rep(c, t, v, d) != h) return;
    const P3 old = a.old[v];
    if (GATHER) {
        // reference walk from the representative (:2376-2490)
        if (!CREASED) {
            P3 acc = add(mul(a.fp[h >> 2], -1.0f), mul(a.ep[he[h].edge], 4.0f));
            float n = 1.0f;
            int32_t it;
            for (it = he[qprev(h)].twin; it >= 0 && it != h && n < (float)kWalkCap; it = he[qprev(it)].twin) {
                acc = add(acc, mul(a.fp[it >> 2], -1.0f));
                acc = add(acc, mul(a.ep[he[it].edge], 4.0f));
                n = n + 1.0f;
            }
            P3 sm = add(mul(acc, 1.0f / (n * n)), mul(old, 1.0f - 3.0f / n));
            a.vp[v] = lerp(old, sm, it != h ? 0.0f : 1.0f);
            return;
        }
        const int32_t p = qprev(h);
        const He xh = he[h], xp = he[p];
        const float ts = lsharp(a.cr, a.C, xh.edge), ps = lsharp(a.cr, a.C, xp.edge);
        const float cw = sgn(ts), pcw = sgn(ps);
        const P3 E = a.ep[xh.edge], PE = a.ep[xp.edge], PF = a.fp[p >> 2];
        P3 sm = add(mul(PF, -1.0f), mul(PE, 4.0f));
        P3 cr = add(zero3(), mul(PE, pcw));
        float avgS = ps, cc = pcw, n = 1.0f;
        int32_t fw, bw;
        for (fw = xp.twin; fw >= 0 && fw != h && n < (float)kWalkCap; fw = he[fw].twin) {
            const int32_t q = qprev(fw);
            const int32_t qe = he[q].edge;
            const P3 qE = a.ep[qe], qF = a.fp[q >> 2];
            const float qs = lsharp(a.cr, a.C, qe), qw = sgn(qs);
            sm = add(sm, mul(qF, -1.0f));
            sm = add(sm, mul(qE, 4.0f));
            n = n + 1.0f;
            cr = add(cr, mul(qE, qw));
            avgS = avgS + qs;
            cc = cc + qw;
            fw = q;
        }
        for (bw = xh.twin; fw < 0 && bw >= 0 && bw != h && n < (float)kWalkCap; bw = he[bw].twin) {
            const int32_t q = qnext(bw);
            const int32_t qe = he[q].edge;
            const P3 qE = a.ep[qe], qF = a.fp[q >> 2];
            const float qs = lsharp(a.cr, a.C, qe), qw = sgn(qs);
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
        if (cc <= 1.0f) a.vp[v] = sm;
        else if (cc >= 3.0f || n == 2.0f) a.vp[v] = old;
        else a.vp[v] = lerp(old, cr, sat(avgS * 0.5f));
        return;
    }
    // scatter rules: find the first halfedge of the fan (the one whose own
    // edge is on the boundary, if any), gather the ring parameters, then sum
    // the per-halfedge contributions in fan order
    int32_t first = h;
    bool boundary = false;
    {
        int32_t it = h;
        for (int g = 0; g < kWalkCap; ++g) {
            const int32_t tw = he[it].twin;
            if (tw < 0) { boundary = true; break; }
            const int32_t nx = qnext(tw);
            if (nx == h) break;
            it = nx;
        }
        if (boundary) first = it;
    }
    RingParams rp;
    rp.n = 0.0f; rp.cc = 0.0f; rp.boundary = boundary; rp.s1 = rp.s2 = 0.0f; rp.e1 = rp.e2 = -1; rp.sall = 0.0f;
    {
        // edges around the vertex: the own edge of `first`, then prev(h_i)
        int32_t it = first;
        if (CREASED && boundary) {
            const int32_t e0 = he[first].edge;
            const float s0 = lsharp(a.cr, a.C, e0);
            rp.cc += sgn(s0);
            if (s0 != 0.0f) { rp.s1 = s0; rp.e1 = e0; }
            rp.sall += s0;
        }
        for (int g = 0; g < kWalkCap; ++g) {
            rp.n += 1.0f;
            const int32_t q = qprev(it);
            const He xq = he[q];
            if (CREASED) {
                const float sq = lsharp(a.cr, a.C, xq.edge);
                rp.cc += sgn(sq);
                if (sq != 0.0f) {
                    if (rp.e1 < 0) { rp.s1 = sq; rp.e1 = xq.edge; }
                    else if (rp.e2 < 0) { rp.s2 = sq; rp.e2 = xq.edge; }
                }
                rp.sall += sq;
            }
            if (xq.twin < 0 || xq.twin == first) break;
            it = xq.twin;
        }
    }
    P3 acc = zero3();
    int32_t it = first;
    for (int g = 0; g < kWalkCap; ++g) {
        const He xi = he[it];
        const int32_t q = qprev(it);
        const He xq = he[q];
        const P3 E = a.ep[xi.edge], PE = a.ep[xq.edge], Fc = a.fp[it >> 2];
        const float ts = CREASED ? lsharp(a.cr, a.C, xi.edge) : 0.0f;
        const float ps = CREASED ? lsharp(a.cr, a.C, xq.edge) : 0.0f;
        acc = add(acc, vertex_scatter_contrib(rp, old, E, PE, Fc, ts, ps, xi.edge, CREASED));
        if (xq.twin < 0 || xq.twin == first) break;
        it = xq.twin;
    }
    a.vp[v] = acc;
}

template <int GATHER, int CREASED>
__global__ __launch_bounds__(256) void k_level_vertices_fan(FusedArgs a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, a.xcd_remap);
    const int32_t h = (int32_t)(lb * blockDim.x + threadIdx.x);
    if (h >= a.H) return;
    const HeView he4{a.he};
    const int32_t v = he4[h].z;
    Fan f;
    if (!owner_fan(h, v, d, c, t, lp.he, f)) return;
    const P3 old = a.old[v];
    if (f.kind == 1 || f.kind == 3) {
        // fans of at most four halfedges known by arithmetic: load all
        // records, then all points, then evaluate
        const int n = f.n;
        int32_t r[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) r[i] = fan_at(f, c, i < n ? i : 0);
        int4 rec[4], prv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            rec[i] = he4[r[i]];
            prv[i] = he4[qprev(r[i])];
        }
        FanPreloaded pre;
        pre.e0 = fan_entry(a, r[0], rec[0], prv[0]);
        pre.e1 = fan_entry(a, r[1], rec[1], prv[1]);
        pre.e2 = fan_entry(a, r[2], rec[2], prv[2]);
        pre.e3 = fan_entry(a, r[3], rec[3], prv[3]);
        a.vp[v] = fan_vertex_point<GATHER, CREASED, 4>(f, pre, old);
        return;
    }
    FanOnDemand od{&a, &f, &c};
    a.vp[v] = fan_vertex_point<GATHER, CREASED, 0>(f, od, old);
}

template <int GATHER, int CREASED, int SCAN>
__global__ __launch_bounds__(256) void k_level_vertices_compact(FusedArgs a, DCage c, LevelTab t, LevelPtrs lp,
                                                                int d) {
    level_vertices_compact_block<GATHER, CREASED, SCAN>(a, c, t, lp, d, xcd_block(blockIdx.x, gridDim.x, a.xcd_remap));
}

// Vertex points, one thread per vertex of level d with the fan derived from
// the level the vertex was born at (vertex_fan): no ownership test, vertex
// points written in order.
template <int GATHER, int CREASED>
__global__ __launch_bounds__(256) void k_level_vertices_direct(FusedArgs a, DCage c, LevelTab t, LevelPtrs lp, int d) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, a.xcd_remap);
    const int32_t v = (int32_t)(lb * 256 + threadIdx.x);
    if (v >= a.V) return;
    Fan f;
    vertex_fan(v, d, c, t, lp.he, f);
    fan_vertex_store<GATHER, CREASED>(a, c, t, d, f, v);
}

// ---- launcher ----------------------------------------------------------------
template <int G, int C, int T, int U, int P, int S, int NT>
static void launch_fused_tiles(const FusedArgs &a, const DCage &c, const LevelTab &t, int d, int tiles,
                               hipStream_t st) {
    constexpr bool kMulti = S == kStoreLds && P == 1;
    if (!kMulti) tiles = 1;
    const int64_t per = (int64_t)NT * tiles;
    const dim3 grid((unsigned)((a.H + per - 1) / per));
    if constexpr (kMulti) {
        if (tiles > 1) {
            hipLaunchKernelGGL((k_level_fused<G, C, T, U, P, S, NT, 1>), grid, dim3(NT), 0, st, a, c, t, d, tiles);
            return;
        }
    }
    hipLaunchKernelGGL((k_level_fused<G, C, T, U, P, S, NT, 0>), grid, dim3(NT), 0, st, a, c, t, d, 1);
}

// 256 and 512 threads, and several tiles per workgroup, with LDS-staged
// stores and points only; 1024 threads and one tile otherwise
template <int G, int C, int T, int U, int P, int S>
static void launch_fused_nt(const FusedArgs &a, const DCage &c, const LevelTab &t, int d, int nt, int tiles,
                            hipStream_t st) {
    if constexpr (S == kStoreLds && P == 1) {
        if (nt == 256) return launch_fused_tiles<G, C, T, U, P, S, 256>(a, c, t, d, tiles, st);
        if (nt == 512) return launch_fused_tiles<G, C, T, U, P, S, 512>(a, c, t, d, tiles, st);
    }
    launch_fused_tiles<G, C, T, U, P, S, kFusedBlock>(a, c, t, d, tiles, st);
}

#define CC_FUSED_CASE(G, C, T, U, P, S)                                                      \
    if (gather == G && creased == C && topo == T && uvMode == U && points == P && store == S) { \
        launch_fused_nt<G, C, T, U, P, S>(a, c, t, d, nt, tiles, st);                         \
        return;                                                                               \
    }

static void launch_fused_impl(const FusedArgs &a, const DCage &c, const LevelTab &t, int d, int gather, int creased,
                              int topo, int uvMode, int points, int store, int nt, int tiles, hipStream_t st) {
#define CC_FUSED_STORES(G, C, T, U, P) CC_FUSED_CASE(G, C, T, U, P, kStoreDirect) CC_FUSED_CASE(G, C, T, U, P, kStoreLds)
#define CC_FUSED_POINTS(G, C, T, U) CC_FUSED_STORES(G, C, T, U, 0) CC_FUSED_STORES(G, C, T, U, 1)
#define CC_FUSED_UV(G, C, T) CC_FUSED_POINTS(G, C, T, 0) CC_FUSED_POINTS(G, C, T, 1) CC_FUSED_POINTS(G, C, T, 2)
#define CC_FUSED_TOPO(G, C) CC_FUSED_UV(G, C, 0) CC_FUSED_UV(G, C, 1)
    CC_FUSED_TOPO(0, 0) CC_FUSED_TOPO(0, 1) CC_FUSED_TOPO(1, 0) CC_FUSED_TOPO(1, 1)
}

void launch_level_fused(const FusedArgs &a, const DCage &c, const LevelTab &t, int d, int gather, int creased, int topo,
                        int uvMode, int points, int store, int nt, int tiles, hipStream_t st) {
    if (nt != 256 && nt != 512) nt = kFusedBlock;
    if (tiles < 1) tiles = 1;
    launch_fused_impl(a, c, t, d, gather ? 1 : 0, creased ? 1 : 0, topo ? 1 : 0, uvMode, points ? 1 : 0, store, nt,
                      tiles, st);
}

template <int SCAN>
static void launch_compact(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d, int gather,
                           int creased, hipStream_t st) {
    const dim3 grid((unsigned)((a.H + 256 * SCAN - 1) / (256 * SCAN))), block(256);
    if (gather && creased) hipLaunchKernelGGL((k_level_vertices_compact<1, 1, SCAN>), grid, block, 0, st, a, c, t, lp, d);
    else if (gather) hipLaunchKernelGGL((k_level_vertices_compact<1, 0, SCAN>), grid, block, 0, st, a, c, t, lp, d);
    else if (creased) hipLaunchKernelGGL((k_level_vertices_compact<0, 1, SCAN>), grid, block, 0, st, a, c, t, lp, d);
    else hipLaunchKernelGGL((k_level_vertices_compact<0, 0, SCAN>), grid, block, 0, st, a, c, t, lp, d);
}

template <int G, int C>
static void launch_ring(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d, int waves,
                        dim3 grid, hipStream_t st) {
    if (waves == 4) hipLaunchKernelGGL((k_level_vertices_ring<G, C, 4>), grid, dim3(256), 0, st, a, c, t, lp, d);
    else if (waves == 6) hipLaunchKernelGGL((k_level_vertices_ring<G, C, 6>), grid, dim3(256), 0, st, a, c, t, lp, d);
    else hipLaunchKernelGGL((k_level_vertices_ring<G, C, kVertexWaves>), grid, dim3(256), 0, st, a, c, t, lp, d);
}

void launch_level_vertices(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d,
                           int gather, int creased, int fan, int scan, int waves, hipStream_t st) {
    const dim3 block(256);
    const dim3 grid((unsigned)((a.H + 255) / 256));
    if (fan == 3) {
        const dim3 g3((unsigned)((a.V + 255) / 256));
        if (gather && creased) hipLaunchKernelGGL((k_level_vertices_direct<1, 1>), g3, block, 0, st, a, c, t, lp, d);
        else if (gather) hipLaunchKernelGGL((k_level_vertices_direct<1, 0>), g3, block, 0, st, a, c, t, lp, d);
        else if (creased) hipLaunchKernelGGL((k_level_vertices_direct<0, 1>), g3, block, 0, st, a, c, t, lp, d);
        else hipLaunchKernelGGL((k_level_vertices_direct<0, 0>), g3, block, 0, st, a, c, t, lp, d);
        return;
    }
    if (fan == 2 && scan == 0 && a.owner_mask && a.rare) {   // owners marked by the fused kernel
        const int64_t words = ((int64_t)a.H + 63) / 64;
        const int64_t waves = (words + kOwnerWords - 1) / kOwnerWords;
        const dim3 g0((unsigned)((waves + 3) / 4)), gr((unsigned)((a.rare_cap + 255) / 256));
        if (gather && creased) {
            launch_ring<1, 1>(a, c, t, lp, d, waves, g0, st);
            if (a.rare_cap > 0) hipLaunchKernelGGL((k_level_vertices_rare<1, 1>), gr, block, 0, st, a, c, t, lp, d);
        } else if (gather) {
            launch_ring<1, 0>(a, c, t, lp, d, waves, g0, st);
            if (a.rare_cap > 0) hipLaunchKernelGGL((k_level_vertices_rare<1, 0>), gr, block, 0, st, a, c, t, lp, d);
        } else if (creased) {
            launch_ring<0, 1>(a, c, t, lp, d, waves, g0, st);
            if (a.rare_cap > 0) hipLaunchKernelGGL((k_level_vertices_rare<0, 1>), gr, block, 0, st, a, c, t, lp, d);
        } else {
            launch_ring<0, 0>(a, c, t, lp, d, waves, g0, st);
            if (a.rare_cap > 0) hipLaunchKernelGGL((k_level_vertices_rare<0, 0>), gr, block, 0, st, a, c, t, lp, d);
        }
        return;
    }
    if (fan == 2) {
        if (scan == 2) launch_compact<2>(a, c, t, lp, d, gather, creased, st);
        else if (scan == 8) launch_compact<8>(a, c, t, lp, d, gather, creased, st);
        else launch_compact<4>(a, c, t, lp, d, gather, creased, st);
        return;
    }
    if (fan) {
        if (gather && creased) hipLaunchKernelGGL((k_level_vertices_fan<1, 1>), grid, block, 0, st, a, c, t, lp, d);
        else if (gather) hipLaunchKernelGGL((k_level_vertices_fan<1, 0>), grid, block, 0, st, a, c, t, lp, d);
        else if (creased) hipLaunchKernelGGL((k_level_vertices_fan<0, 1>), grid, block, 0, st, a, c, t, lp, d);
        else hipLaunchKernelGGL((k_level_vertices_fan<0, 0>), grid, block, 0, st, a, c, t, lp, d);
        return;
    }
    if (gather && creased) hipLaunchKernelGGL((k_level_vertices<1, 1>), grid, block, 0, st, a, c, t, d);
    else if (gather) hipLaunchKernelGGL((k_level_vertices<1, 0>), grid, block, 0, st, a, c, t, d);
    else if (creased) hipLaunchKernelGGL((k_level_vertices<0, 1>), grid, block, 0, st, a, c, t, d);
    else hipLaunchKernelGGL((k_level_vertices<0, 0>), grid, block, 0, st, a, c, t, d);
}

}  // namespace cch
