// Batched forward kernels: the cage-step rules, crease refinement and the
// patch-local level step over many cages per launch (see cc_batch.h for the
// job lookup).
#include "cc_batch.h"
#include "cc_level.h"
#include "cc_level0.h"
#include "cc_patch.h"

namespace cch {

namespace {

struct OpCreases {
    __device__ void operator()(const CreaseJob &J, int64_t i) const {
        if (i < J.C) creases_at(J.in, J.C, J.out, (int32_t)i);
    }
};

struct OpCageHalfedges {
    int topo;
    __device__ void operator()(const CageJob &J, int64_t i) const {
        if (i < J.c.H) cage_halfedges_at(J.c, J.he1, topo, J.uv_mode, (int32_t)i);
    }
};

struct OpCageFaces {
    int gather;
    __device__ void operator()(const CageJob &J, int64_t i) const {
        if (i < J.c.F) cage_faces_at(J.c, J.pts1 + J.c.V, gather, (int32_t)i);
    }
};

struct OpCageEdges {
    int gather, creased;
    __device__ void operator()(const CageJob &J, int64_t i) const {
        const P3 *fp = J.pts1 + J.c.V;
        if (i < J.c.E) cage_edges_at(J.c, fp, J.pts1 + J.c.V + J.c.F, gather, creased, (int32_t)i);
    }
};

struct OpCageVerticesGather {
    int creased;
    __device__ void operator()(const CageJob &J, int64_t i) const {
        const P3 *fp = J.pts1 + J.c.V, *ep = fp + J.c.F;
        if (i < J.c.V) cage_vertices_gather_at(J.c, fp, ep, J.pts1, creased, (int32_t)i);
    }
};

struct OpCageVerticesScatter {
    int creased;
    __device__ void operator()(const CageJob &J, int64_t i) const {
        const P3 *fp = J.pts1 + J.c.V, *ep = fp + J.c.F;
        if (i < J.c.V) cage_vertices_scatter_at(J.c, fp, ep, J.pts1, creased, (int32_t)i);
    }
};

struct OpZeroCageVertices {
    __device__ void operator()(const CageJob &J, int64_t i) const {
        if (i < 3 * (int64_t)J.c.V) reinterpret_cast<float *>(J.pts1)[i] = 0.0f;
    }
};

struct OpCopy {
    const float *src;
    float *dst;
    __device__ void operator()(const SpanJob &J, int64_t i) const {
        if (i < J.n) dst[J.dst_off + i] = src[J.src_off + i];
    }
};

// tiles [b*hpt, b*hpt + hpt) of job k's level, b the job-local workgroup
template <int GATHER, int CREASED, int TOPO, int NT>
__global__ __launch_bounds__(NT) void k_batch_patch(const LevelJob *__restrict__ jobs, const int32_t *__restrict__ start,
                                                    int n, int d, uint32_t hpt, int remap, char *arena) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, remap);
    const int k = find_job(start, n, lb);
    const LevelJob &J = jobs[k];
    FusedArgs a;
    DCage c;
    unpack_level(J, arena, a, c);
    const JobLevelPtrs lp{J.lp, arena};
    const uint32_t tiles = (uint32_t)((a.H + NT - 1) / NT);
    const uint32_t b = lb - (uint32_t)start[k];
    for (uint32_t it = 0; it < hpt; ++it) {
        const uint32_t tile = b * hpt + it;
        if (tile >= tiles) break;
        if (it) __syncthreads();   // the previous tile's LDS reads are done
        patch_tile<GATHER, CREASED, TOPO, TOPO ? 3 : 0, NT>(a, c, J.t, lp, d, tile);
    }
}

// children with face and edge points of job k's level (workgroups of NT
// threads; MULTI: each processing `tiles` tiles of NT halfedges in turn, a
// kernel of its own since the loop costs registers)
template <int GATHER, int CREASED, int TOPO, int NT, int MULTI>
__global__ __launch_bounds__(NT) void k_batch_fused(const LevelJob *__restrict__ jobs, const int32_t *__restrict__ start,
                                                    int n, int d, int remap, int tiles, char *arena) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, remap);
    const int k = find_job(start, n, lb);
    const LevelJob &J = jobs[k];
    FusedArgs a;
    DCage c;
    unpack_level(J, arena, a, c);
    constexpr int UV = TOPO ? 3 : 0;
    if constexpr (!MULTI) {
        level_fused_block<GATHER, CREASED, TOPO, UV, 1, kStoreLds, NT>(a, c, J.t, d, lb - (uint32_t)start[k]);
    } else {
        const uint32_t tile0 = (lb - (uint32_t)start[k]) * (uint32_t)tiles;
        for (int q = 0; q < tiles; ++q) {
            if (q) __syncthreads();
            level_fused_block<GATHER, CREASED, TOPO, UV, 1, kStoreLds, NT>(a, c, J.t, d, tile0 + (uint32_t)q);
        }
    }
}

// vertex points of the owners k_batch_fused marked (kOwnerWords mask words
// per wave, four waves per workgroup; W waves per SIMD)
template <int GATHER, int CREASED, int W>
__global__ __launch_bounds__(256, W) void k_batch_vertices_masked(const LevelJob *__restrict__ jobs,
                                                               const int32_t *__restrict__ start, int n, int d,
                                                               int remap, char *arena) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, remap);
    const int k = find_job(start, n, lb);
    const LevelJob &J = jobs[k];
    FusedArgs a;
    DCage c;
    unpack_level(J, arena, a, c);
    const JobLevelPtrs lp{J.lp, arena};
    const int32_t wave = (int32_t)(lb - (uint32_t)start[k]) * 4 + (int32_t)(threadIdx.x >> 6);
    level_vertices_masked<GATHER, CREASED>(a, c, J.t, lp, d, wave);
}

template <int GATHER, int CREASED>
__global__ __launch_bounds__(256) void k_batch_vertices_rare(const LevelJob *__restrict__ jobs,
                                                             const int32_t *__restrict__ start, int n, int d,
                                                             char *arena) {
    const int k = find_job(start, n, blockIdx.x);
    const LevelJob &J = jobs[k];
    FusedArgs a;
    DCage c;
    unpack_level(J, arena, a, c);
    const JobLevelPtrs lp{J.lp, arena};
    level_vertices_rare<GATHER, CREASED>(a, c, J.t, lp, d,
                                         (int32_t)((blockIdx.x - (uint32_t)start[k]) * 256 + threadIdx.x));
}

// vertex points of job k's level (one block per 256 * SCAN halfedges)
template <int GATHER, int CREASED, int SCAN>
__global__ __launch_bounds__(256) void k_batch_vertices(const LevelJob *__restrict__ jobs,
                                                        const int32_t *__restrict__ start, int n, int d, int remap,
                                                        char *arena) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, remap);
    const int k = find_job(start, n, lb);
    const LevelJob &J = jobs[k];
    FusedArgs a;
    DCage c;
    unpack_level(J, arena, a, c);
    const JobLevelPtrs lp{J.lp, arena};
    level_vertices_compact_block<GATHER, CREASED, SCAN>(a, c, J.t, lp, d, lb - (uint32_t)start[k]);
}

template <int G, int C, int SCAN>
void vertices_scan(const BatchTable &t, char *arena, int d, int remap, hipStream_t st) {
    hipLaunchKernelGGL((k_batch_vertices<G, C, SCAN>), dim3(t.blocks), dim3(256), 0, st,
                       static_cast<const LevelJob *>(t.jobs), t.start, (int)t.n, d, remap, arena);
}

template <int G, int C, int W>
void vertices_masked(const BatchTable &t, char *arena, int d, int remap, hipStream_t st) {
    hipLaunchKernelGGL((k_batch_vertices_masked<G, C, W>), dim3(t.blocks), dim3(256), 0, st,
                       static_cast<const LevelJob *>(t.jobs), t.start, (int)t.n, d, remap, arena);
}

template <int G, int C>
void vertices_gc(const BatchTable &t, char *arena, int d, int scan, int remap, int waves, hipStream_t st) {
    if (scan == 0) {
        if (waves == 4) vertices_masked<G, C, 4>(t, arena, d, remap, st);
        else if (waves == 6) vertices_masked<G, C, 6>(t, arena, d, remap, st);
        else vertices_masked<G, C, kVertexWaves>(t, arena, d, remap, st);
        return;
    }
    if (scan == 2) vertices_scan<G, C, 2>(t, arena, d, remap, st);
    else if (scan == 8) vertices_scan<G, C, 8>(t, arena, d, remap, st);
    else vertices_scan<G, C, 4>(t, arena, d, remap, st);
}

template <int G, int C, int T, int NT>
void fused_tiles(const BatchTable &t, char *arena, int d, int remap, int tiles, hipStream_t st) {
    const LevelJob *jobs = static_cast<const LevelJob *>(t.jobs);
    if (tiles > 1)
        hipLaunchKernelGGL((k_batch_fused<G, C, T, NT, 1>), dim3(t.blocks), dim3(NT), 0, st, jobs, t.start, (int)t.n, d,
                           remap, tiles, arena);
    else
        hipLaunchKernelGGL((k_batch_fused<G, C, T, NT, 0>), dim3(t.blocks), dim3(NT), 0, st, jobs, t.start, (int)t.n, d,
                           remap, 1, arena);
}

// (the table's workgroup counts are per nt * tiles halfedges)
template <int G, int C, int T>
void fused_nt(const BatchTable &t, char *arena, int d, int remap, int nt, int tiles, hipStream_t st) {
    if (nt == 256) fused_tiles<G, C, T, 256>(t, arena, d, remap, tiles, st);
    else if (nt == 512) fused_tiles<G, C, T, 512>(t, arena, d, remap, tiles, st);
    else fused_tiles<G, C, T, kFusedBlock>(t, arena, d, remap, tiles, st);
}

template <int G, int C>
void fused_gc(const BatchTable &t, char *arena, int d, int topo, int remap, int nt, int tiles, hipStream_t st) {
    if (topo) fused_nt<G, C, 1>(t, arena, d, remap, nt, tiles, st);
    else fused_nt<G, C, 0>(t, arena, d, remap, nt, tiles, st);
}

// job k's deferred vertices: kDeferGroup mask words per wave
template <int GATHER, int CREASED>
__global__ __launch_bounds__(256) void k_batch_deferred(const LevelJob *__restrict__ jobs,
                                                        const int32_t *__restrict__ start, int n, int d, char *arena) {
    const int k = find_job(start, n, blockIdx.x);
    const LevelJob &J = jobs[k];
    FusedArgs a;
    DCage c;
    unpack_level(J, arena, a, c);
    const JobLevelPtrs lp{J.lp, arena};
    const int32_t words = (a.H + 63) >> 6;
    const int32_t wave = (int32_t)(blockIdx.x - (uint32_t)start[k]) * 4 + (int32_t)(threadIdx.x >> 6);
    deferred_group<GATHER, CREASED>(a, c, J.t, lp, d, wave * kDeferGroup, words);
}

template <int G, int C, int T, int NT>
void patch_nt(const BatchTable &t, char *arena, int d, int hpt, int remap, hipStream_t st) {
    hipLaunchKernelGGL((k_batch_patch<G, C, T, NT>), dim3(t.blocks), dim3(NT), 0, st,
                       static_cast<const LevelJob *>(t.jobs), t.start, (int)t.n, d, (uint32_t)hpt, remap, arena);
}

template <int G, int C, int T>
void patch_t(const BatchTable &t, char *arena, int d, int nt, int hpt, int remap, hipStream_t st) {
    if (nt == 128) patch_nt<G, C, T, 128>(t, arena, d, hpt, remap, st);
    else if (nt == 512) patch_nt<G, C, T, 512>(t, arena, d, hpt, remap, st);
    else patch_nt<G, C, T, 256>(t, arena, d, hpt, remap, st);
}

template <int G, int C>
void patch_gc(const BatchTable &t, char *arena, int d, int topo, int nt, int hpt, int remap, hipStream_t st) {
    if (topo) patch_t<G, C, 1>(t, arena, d, nt, hpt, remap, st);
    else patch_t<G, C, 0>(t, arena, d, nt, hpt, remap, st);
}

}  // namespace

void batch_creases(const BatchTable &t, hipStream_t st) { launch_batch<CreaseJob>(t, OpCreases{}, st); }

void batch_cage_halfedges(const BatchTable &t, int topo, hipStream_t st) {
    launch_batch<CageJob>(t, OpCageHalfedges{topo}, st);
}

void batch_cage_faces(const BatchTable &t, int gather, hipStream_t st) {
    launch_batch<CageJob>(t, OpCageFaces{gather}, st);
}

void batch_cage_edges(const BatchTable &t, int gather, int creased, hipStream_t st) {
    launch_batch<CageJob>(t, OpCageEdges{gather, creased}, st);
}

void batch_cage_vertices(const BatchTable &t, int gather, int creased, hipStream_t st) {
    if (gather) launch_batch<CageJob>(t, OpCageVerticesGather{creased}, st);
    else launch_batch<CageJob>(t, OpCageVerticesScatter{creased}, st);
}

void batch_zero_cage_vertices(const BatchTable &t, hipStream_t st) {
    launch_batch<CageJob>(t, OpZeroCageVertices{}, st);
}

void batch_copy(const BatchTable &t, const float *src, float *dst, hipStream_t st) {
    launch_batch<SpanJob>(t, OpCopy{src, dst}, st);
}

void batch_level_patch(const BatchTable &t, char *arena, int d, int gather, int creased, int topo, int nt, int hpt,
                       int remap, hipStream_t st) {
    if (t.blocks == 0 || t.n == 0) return;
    hpt = hpt < 1 ? 1 : hpt;
    if (gather && creased) patch_gc<1, 1>(t, arena, d, topo, nt, hpt, remap, st);
    else if (gather) patch_gc<1, 0>(t, arena, d, topo, nt, hpt, remap, st);
    else if (creased) patch_gc<0, 1>(t, arena, d, topo, nt, hpt, remap, st);
    else patch_gc<0, 0>(t, arena, d, topo, nt, hpt, remap, st);
}

void batch_level_fused(const BatchTable &t, char *arena, int d, int gather, int creased, int topo, int remap, int nt,
                       int tiles, hipStream_t st) {
    if (t.blocks == 0 || t.n == 0) return;
    if (gather && creased) fused_gc<1, 1>(t, arena, d, topo, remap, nt, tiles, st);
    else if (gather) fused_gc<1, 0>(t, arena, d, topo, remap, nt, tiles, st);
    else if (creased) fused_gc<0, 1>(t, arena, d, topo, remap, nt, tiles, st);
    else fused_gc<0, 0>(t, arena, d, topo, remap, nt, tiles, st);
}

void batch_level_vertices(const BatchTable &t, char *arena, int d, int gather, int creased, int scan, int remap,
                          int waves, hipStream_t st) {
    if (t.blocks == 0 || t.n == 0) return;
    if (gather && creased) vertices_gc<1, 1>(t, arena, d, scan, remap, waves, st);
    else if (gather) vertices_gc<1, 0>(t, arena, d, scan, remap, waves, st);
    else if (creased) vertices_gc<0, 1>(t, arena, d, scan, remap, waves, st);
    else vertices_gc<0, 0>(t, arena, d, scan, remap, waves, st);
}

void batch_level_vertices_rare(const BatchTable &t, char *arena, int d, int gather, int creased, hipStream_t st) {
    if (t.blocks == 0 || t.n == 0) return;
    const dim3 grid(t.blocks), block(256);
    const LevelJob *jobs = static_cast<const LevelJob *>(t.jobs);
    const int n = (int)t.n;
    if (gather && creased) hipLaunchKernelGGL((k_batch_vertices_rare<1, 1>), grid, block, 0, st, jobs, t.start, n, d, arena);
    else if (gather) hipLaunchKernelGGL((k_batch_vertices_rare<1, 0>), grid, block, 0, st, jobs, t.start, n, d, arena);
    else if (creased) hipLaunchKernelGGL((k_batch_vertices_rare<0, 1>), grid, block, 0, st, jobs, t.start, n, d, arena);
    else hipLaunchKernelGGL((k_batch_vertices_rare<0, 0>), grid, block, 0, st, jobs, t.start, n, d, arena);
}

void batch_level_deferred(const BatchTable &t, char *arena, int d, int gather, int creased, hipStream_t st) {
    if (t.blocks == 0 || t.n == 0) return;
    const dim3 grid(t.blocks), block(256);
    const LevelJob *jobs = static_cast<const LevelJob *>(t.jobs);
    const int n = (int)t.n;
    if (gather && creased) hipLaunchKernelGGL((k_batch_deferred<1, 1>), grid, block, 0, st, jobs, t.start, n, d, arena);
    else if (gather) hipLaunchKernelGGL((k_batch_deferred<1, 0>), grid, block, 0, st, jobs, t.start, n, d, arena);
    else if (creased) hipLaunchKernelGGL((k_batch_deferred<0, 1>), grid, block, 0, st, jobs, t.start, n, d, arena);
    else hipLaunchKernelGGL((k_batch_deferred<0, 0>), grid, block, 0, st, jobs, t.start, n, d, arena);
}

}  // namespace cch
