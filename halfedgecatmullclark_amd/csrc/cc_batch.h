// Batched refinement: many cages of uneven size in one launch per kernel kind
// and level.
//
// A launch covers a list of jobs (one per cage taking part) and a prefix sum
// `start` of the workgroups each job needs (start[0] = 0, start[n] = grid).
// A workgroup finds its job by binary search over `start`; the search reads
// the same addresses in every lane, so it runs on the scalar unit.  All
// per-cage offsets are 64-bit: a batch holds billions of halfedges, while
// the IDs inside one cage stay 32-bit as in the reference.
#pragma once

#include "cc_adjoint.h"
#include "cc_common.h"
#include "cc_fused.h"

namespace cch {

// one cage of a batch, its cage step (level 0 -> 1)
struct CageJob {
    DCage c;
    He *he1;        // level-1 halfedges
    P3 *pts1;       // level-1 points: [V0 vertex | F0 face | E0 edge]
    int uv_mode;    // 1 refine uvIDs, 2 write 0 (cage without uvs)
};

// crease refinement of one cage's level d -> d+1
struct CreaseJob {
    const Crease *in;
    Crease *out;
    int32_t C;
};

// one cage's level d -> d+1
struct LevelJob {
    FusedArgs a;
    DCage c;
    LevelTab t;
    LevelPtrs lp;
};

// a range of n floats copied from src + src_off to dst + dst_off, the base
// pointers being launch arguments
struct SpanJob {
    int64_t src_off, dst_off, n;
};

// one cage's adjoint step d+1 -> d (d = 0 is the cage step)
struct AdjJob {
    AdjArgs<float> a;
    DCage c;
    LevelTab t;
    LevelPtrs lp;
};

// largest k in [0, n) with start[k] <= b (jobs without workgroups are skipped)
__device__ __forceinline__ int find_job(const int32_t *__restrict__ start, int n, uint32_t b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((uint32_t)start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One thread per element of each job: op(job, i) with i the job-local index.
template <class Job, class Op>
__global__ __launch_bounds__(256) void k_batch(const Job *__restrict__ jobs, const int32_t *__restrict__ start, int n,
                                               Op op) {
    const int k = find_job(start, n, blockIdx.x);
    const int64_t i = (int64_t)(blockIdx.x - (uint32_t)start[k]) * 256 + threadIdx.x;
    op(jobs[k], i);
}

// device tables of one batched launch
struct BatchTable {
    const void *jobs = nullptr;
    const int32_t *start = nullptr;
    int32_t n = 0;         // jobs
    uint32_t blocks = 0;   // workgroups of the launch
};

template <class Job, class Op>
void launch_batch(const BatchTable &t, Op op, hipStream_t st) {
    if (t.blocks == 0 || t.n == 0) return;
    hipLaunchKernelGGL((k_batch<Job, Op>), dim3(t.blocks), dim3(256), 0, st, static_cast<const Job *>(t.jobs), t.start,
                       (int)t.n, op);
}

// forward launches (cc_batch.hip)
void batch_creases(const BatchTable &t, hipStream_t st);
void batch_cage_halfedges(const BatchTable &t, int topo, hipStream_t st);
void batch_cage_faces(const BatchTable &t, int gather, hipStream_t st);
void batch_cage_edges(const BatchTable &t, int gather, int creased, hipStream_t st);
void batch_cage_vertices(const BatchTable &t, int gather, int creased, hipStream_t st);
void batch_zero_cage_vertices(const BatchTable &t, hipStream_t st);
void batch_copy(const BatchTable &t, const float *src, float *dst, hipStream_t st);
// patch-local level step; `start` counts workgroups of hpt tiles of nt
// halfedges per job
void batch_level_patch(const BatchTable &t, int d, int gather, int creased, int topo, int nt, int hpt, int remap,
                       hipStream_t st);
void batch_level_deferred(const BatchTable &t, int d, int gather, int creased, hipStream_t st);
// fused children + face/edge points (one workgroup per 256 halfedges), then
// vertex points (one workgroup per 256 * scan halfedges; scan 0: the owners
// the fused kernel marked, one workgroup per 1024 mask words)
void batch_level_fused(const BatchTable &t, int d, int gather, int creased, int topo, int remap, hipStream_t st);
void batch_level_vertices(const BatchTable &t, int d, int gather, int creased, int scan, int remap, hipStream_t st);

// adjoint launches (cc_adjoint.hip)
void batch_adj_level(const BatchTable &quads, const BatchTable &halfedges, int d, int creased, hipStream_t st);
void batch_adj_cage(const BatchTable &t, int creased, hipStream_t st);

}  // namespace cch
