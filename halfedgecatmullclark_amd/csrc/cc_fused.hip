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
#include "cc_fan.h"
#include "cc_fused.h"

namespace cch {


template <int GATHER, int CREASED, int TOPO, int UVMODE, int POINTS, int STORE>
__global__ __launch_bounds__(256) void k_level_fused(FusedArgs a) {
    constexpr bool kStaged = STORE == kStoreLds && TOPO && UVMODE != 0;
    __shared__ int4 stage[kStaged ? 4 * 256 : 1];
    __shared__ P3 spts[POINTS ? 256 : 1];
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, a.xcd_remap);
    const int32_t base = (int32_t)(lb * blockDim.x);
    const int32_t h = base + (int32_t)threadIdx.x;
    const bool live = h < a.H;   // H is a multiple of 4: quads are whole
    const int j = h & 3;
    const int4 *he4 = reinterpret_cast<const int4 *>(a.he);
    const int4 x = live ? he4[h] : make_int4(0, 0, 0, 0);
    const int32_t twin = x.x;
    const bool owner = live && h > twin;   // this halfedge writes its edge point
    // twin quads inside this block come from LDS, the others from memory
    const bool twin_here = owner && twin >= base && twin < base + (int32_t)blockDim.x;
    const bool twin_far = owner && twin >= 0 && !twin_here;

    // issue the loads of the point work first: this halfedge's vertex point
    // and the records of a twin quad outside the block
    P3 P = zero3();
    int4 tr0, tr1, tr2, tr3;
    if (POINTS) {
        if (live) P = a.old[x.z];
        const int32_t tq = twin_far ? (twin & ~3) : (h & ~3);
        tr0 = he4[tq]; tr1 = he4[tq + 1]; tr2 = he4[tq + 2]; tr3 = he4[tq + 3];
    }

    // record of the previous halfedge of the quad, by DPP
    const int4 p = quad_rel4<kQuadPrev>(x);

    if (TOPO || UVMODE) {
        int4 c[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] = child_record(k, h, x, p, a.V, a.F, a.E);
        if (UVMODE == 1) {
            // uvs of this, the next, the opposite and the previous halfedge;
            // the face average sums them in that order (:2904-2913)
            const UV uv = uv_decode(x.w);
            const UV nuv = quad_reluv<kQuadNext>(uv), ouv = quad_reluv<kQuadOpp>(uv), puv = quad_reluv<kQuadPrev>(uv);
            UV fuv = uv;
            fuv.u += nuv.u; fuv.v += nuv.v;
            fuv.u += ouv.u; fuv.v += ouv.v;
            fuv.u += puv.u; fuv.v += puv.v;
            fuv.u /= 4.0f;
            fuv.v /= 4.0f;
            c[0].w = x.w;
            c[1].w = uv_encode(uv_mid(uv, nuv));
            c[2].w = uv_encode(fuv);
            c[3].w = uv_encode(uv_mid(uv, puv));
        }
        int4 *out4 = reinterpret_cast<int4 *>(a.out_he);
        if (kStaged) {
#pragma unroll
            for (int k = 0; k < 4; ++k) stage[4 * threadIdx.x + k] = c[k];
            if (POINTS) spts[threadIdx.x] = P;
            __syncthreads();
            const int64_t obase = 4 * (int64_t)base;
            const int64_t limit = 4 * (int64_t)a.H;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int64_t o = obase + threadIdx.x + (int64_t)k * blockDim.x;
                if (o < limit) out4[o] = stage[threadIdx.x + k * blockDim.x];
            }
        } else if (live) {
            if (TOPO && (UVMODE == 1 || UVMODE == 2)) {
#pragma unroll
                for (int k = 0; k < 4; ++k) out4[4 * (int64_t)h + k] = c[k];
            } else if (TOPO) {
                He *o = a.out_he + 4 * (int64_t)h;
#pragma unroll
                for (int k = 0; k < 4; ++k) { o[k].twin = c[k].x; o[k].edge = c[k].y; o[k].vert = c[k].z; }
            } else {
                He *o = a.out_he + 4 * (int64_t)h;
#pragma unroll
                for (int k = 0; k < 4; ++k) o[k].uv = c[k].w;
            }
        }
    }

    if (!POINTS) return;
    if (!kStaged) {
        spts[threadIdx.x] = P;
        __syncthreads();
    }
    const P3 q0 = quad_bcast3<0>(P), q1 = quad_bcast3<1>(P), q2 = quad_bcast3<2>(P), q3 = quad_bcast3<3>(P);
    if (!live) return;
    const P3 F = face_rule(q0, q1, q2, q3, GATHER);
    if (j == 0) a.fp[h >> 2] = F;
    if (!owner) return;
    // this halfedge owns its edge: both contributions
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
            const float w = CREASED ? sat(s) : 1.0f;
            Ev = edge_gather(P, v1, F, Ft, w, CREASED);
        } else {
            const int tj = twin & 3;
            const P3 ct = edge_contrib(pick4(t0, t1, t2, t3, tj), pick4(t0, t1, t2, t3, (tj + 1) & 3), Ft, false, s,
                                       CREASED);
            const P3 ch = edge_contrib(P, v1, F, false, s, CREASED);
            Ev = add(add(zero3(), ct), ch);
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
    const int4 *he4 = reinterpret_cast<const int4 *>(a.he);
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

// Vertex points with owners compacted per block: a block scans
// 256 * SCAN halfedges, collects the representatives in LDS, then every
// thread evaluates owners, so all lanes do ring work.
template <int GATHER, int CREASED, int SCAN>
__global__ __launch_bounds__(256) void k_level_vertices_compact(FusedArgs a, DCage c, LevelTab t, LevelPtrs lp,
                                                                int d) {
    __shared__ int32_t owners[256 * SCAN];
    __shared__ int32_t nown;
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, a.xcd_remap);
    const int32_t base = (int32_t)(lb * 256 * SCAN);
    const int4 *he4 = reinterpret_cast<const int4 *>(a.he);
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
        Fan f;
        if (vs[k] >= 0 && owner_fan(h, vs[k], d, c, t, lp.he, f)) owners[atomicAdd(&nown, 1)] = h;
    }
    __syncthreads();
    const int32_t count = nown;
    for (int32_t i = (int32_t)threadIdx.x; i < count; i += 256) {
        const int32_t h = owners[i];
        const int32_t v = he4[h].z;
        Fan f;
        owner_fan(h, v, d, c, t, lp.he, f);
        const P3 old = a.old[v];
        if (f.kind == 1 || f.kind == 3) {
            const int n = f.n;
            int32_t r[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) r[q] = fan_at(f, c, q < n ? q : 0);
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

// ---- launcher ----------------------------------------------------------------
#define CC_FUSED_CASE(G, C, T, U, P, S)                                                      \
    if (gather == G && creased == C && topo == T && uvMode == U && points == P && store == S) { \
        hipLaunchKernelGGL((k_level_fused<G, C, T, U, P, S>), grid, block, 0, st, a);          \
        return;                                                                               \
    }

static void launch_fused_impl(const FusedArgs &a, int gather, int creased, int topo, int uvMode, int points, int store,
                              dim3 grid, dim3 block, hipStream_t st) {
#define CC_FUSED_STORES(G, C, T, U, P) CC_FUSED_CASE(G, C, T, U, P, kStoreDirect) CC_FUSED_CASE(G, C, T, U, P, kStoreLds)
#define CC_FUSED_POINTS(G, C, T, U) CC_FUSED_STORES(G, C, T, U, 0) CC_FUSED_STORES(G, C, T, U, 1)
#define CC_FUSED_UV(G, C, T) CC_FUSED_POINTS(G, C, T, 0) CC_FUSED_POINTS(G, C, T, 1) CC_FUSED_POINTS(G, C, T, 2)
#define CC_FUSED_TOPO(G, C) CC_FUSED_UV(G, C, 0) CC_FUSED_UV(G, C, 1)
    CC_FUSED_TOPO(0, 0) CC_FUSED_TOPO(0, 1) CC_FUSED_TOPO(1, 0) CC_FUSED_TOPO(1, 1)
}

void launch_level_fused(const FusedArgs &a, int gather, int creased, int topo, int uvMode, int points, int store,
                        hipStream_t st) {
    const dim3 block(256);
    const dim3 grid((unsigned)((a.H + 255) / 256));
    launch_fused_impl(a, gather ? 1 : 0, creased ? 1 : 0, topo ? 1 : 0, uvMode, points ? 1 : 0, store, grid, block, st);
}

void launch_level_vertices(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d,
                           int gather, int creased, int fan, hipStream_t st) {
    const dim3 block(256);
    const dim3 grid((unsigned)((a.H + 255) / 256));
    if (fan == 2) {
        constexpr int kScan = 4;
        const dim3 g2((unsigned)((a.H + 256 * kScan - 1) / (256 * kScan)));
        if (gather && creased) hipLaunchKernelGGL((k_level_vertices_compact<1, 1, kScan>), g2, block, 0, st, a, c, t, lp, d);
        else if (gather) hipLaunchKernelGGL((k_level_vertices_compact<1, 0, kScan>), g2, block, 0, st, a, c, t, lp, d);
        else if (creased) hipLaunchKernelGGL((k_level_vertices_compact<0, 1, kScan>), g2, block, 0, st, a, c, t, lp, d);
        else hipLaunchKernelGGL((k_level_vertices_compact<0, 0, kScan>), g2, block, 0, st, a, c, t, lp, d);
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
