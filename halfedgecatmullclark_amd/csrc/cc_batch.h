// Batched refinement: many cages of uneven size in one launch per kernel kind
// and level.
//
// A launch covers a list of jobs (one per cage taking part) and a prefix sum
// `start` of the workgroups each job needs (start[0] = 0, start[n] = grid),
// followed by the job index of every workgroup, so a workgroup finds its job
// with one load (the same address in every lane: a scalar load).  All
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

// One cage's level d -> d+1.  Buffers are byte offsets into the batch's
// arena (one allocation whose address is a kernel argument, so that the
// compiler knows every address derived from it is global memory and emits
// global rather than flat instructions); -1 marks an absent buffer.
struct LevelJob {
    int64_t he, out_he, cr, old, vp, fp, ep, defer_mask, owner_mask, rare, erep;
    int64_t c_he, c_v2h, c_e2h, c_f2h, c_pts, c_uvs, c_cr, c_vro, c_vr, c_vb, c_fro, c_fr, c_voo, c_vo;
    int64_t lp[kMaxDepth + 1];
    int32_t H, V, F, E, C;
    int32_t cV, cU, cH, cE, cF;
    int32_t xcd_remap, uv_refine, rare_cap, erep_read, erep_write;
    LevelTab t;
};

// a range of n floats copied from src + src_off to dst + dst_off, the base
// pointers being launch arguments
struct SpanJob {
    int64_t src_off, dst_off, n;
};

// One cage's adjoint step d+1 -> d (d = 0 is the cage step): he, cr and the
// cage are offsets into the arena, the gradients and weights into the
// adjoint arena; with gin set gV, gF and gE are offsets into the caller's
// gradient of the finest points, whose address is the adjoint arena's first
// 8 bytes (written before each adjoint run).
struct AdjJob {
    int64_t he, cr, gV, gF, gE, gP, W, erep;
    int64_t vp;   // the level's vertex-rule parameters (kept across calls), or -1: in W
    int64_t gFu;  // the face gradients after the quad pass, or -1: in place (gF)
    int64_t rare; // the vertex pass's rare-vertex list (arena offset), or -1
    int32_t rare_cap;
    int32_t gin;
    int64_t c_he, c_v2h, c_e2h, c_f2h, c_pts, c_uvs, c_cr, c_vro, c_vr, c_vb, c_fro, c_fr, c_voo, c_vo;
    int64_t lp[kMaxDepth + 1];
    int32_t H, C, V;
    int32_t cV, cU, cH, cE, cF;
    LevelTab t;
};

template <class T>
__host__ __device__ __forceinline__ T *arena_ptr(char *base, int64_t off) {
    return off < 0 ? nullptr : reinterpret_cast<T *>(base + off);
}

template <class J>
__host__ __device__ __forceinline__ DCage arena_cage(char *base, const J &j) {
    DCage c;
    c.he = arena_ptr<const CageHe>(base, j.c_he);
    c.v2h = arena_ptr<const int32_t>(base, j.c_v2h);
    c.e2h = arena_ptr<const int32_t>(base, j.c_e2h);
    c.f2h = arena_ptr<const int32_t>(base, j.c_f2h);
    c.pts = arena_ptr<const P3>(base, j.c_pts);
    c.uvs = arena_ptr<const UV>(base, j.c_uvs);
    c.cr = arena_ptr<const Crease>(base, j.c_cr);
    c.vring_off = arena_ptr<const int32_t>(base, j.c_vro);
    c.vring = arena_ptr<const int32_t>(base, j.c_vr);
    c.vring_bnd = arena_ptr<const int32_t>(base, j.c_vb);
    c.fring_off = arena_ptr<const int32_t>(base, j.c_fro);
    c.fring = arena_ptr<const int32_t>(base, j.c_fr);
    c.vout_off = arena_ptr<const int32_t>(base, j.c_voo);
    c.vout = arena_ptr<const int32_t>(base, j.c_vo);
    c.V = j.cV; c.U = j.cU; c.H = j.cH; c.E = j.cE; c.F = j.cF;
    return c;
}

// the kernel arguments of one cage's level from its job and the arena
__host__ __device__ __forceinline__ void unpack_level(const LevelJob &J, char *arena, FusedArgs &a, DCage &c) {
    a.he = arena_ptr<const He>(arena, J.he);
    a.out_he = arena_ptr<He>(arena, J.out_he);
    a.cr = arena_ptr<const Crease>(arena, J.cr);
    a.old = arena_ptr<const P3>(arena, J.old);
    a.vp = arena_ptr<P3>(arena, J.vp);
    a.fp = arena_ptr<P3>(arena, J.fp);
    a.ep = arena_ptr<P3>(arena, J.ep);
    a.defer_mask = arena_ptr<uint64_t>(arena, J.defer_mask);
    a.owner_mask = arena_ptr<uint64_t>(arena, J.owner_mask);
    a.rare = arena_ptr<int32_t>(arena, J.rare);
    a.rare_cap = J.rare_cap;
    a.erep = arena_ptr<uint64_t>(arena, J.erep);
    a.erep_read = J.erep_read;
    a.erep_write = J.erep_write;
    a.H = J.H; a.V = J.V; a.F = J.F; a.E = J.E; a.C = J.C;
    a.xcd_remap = J.xcd_remap;
    a.uv_refine = J.uv_refine;
    c = arena_cage(arena, J);
}

__host__ __device__ __forceinline__ void unpack_adj(const AdjJob &J, char *arena, char *adj, AdjArgs<float> &a,
                                                    DCage &c) {
    a.he = arena_ptr<const He>(arena, J.he);
    a.cr = arena_ptr<const Crease>(arena, J.cr);
    a.sd = nullptr;
    a.H = J.H; a.C = J.C; a.V = J.V;
    char *const g = J.gin ? *reinterpret_cast<char *const *>(adj) : adj;
    a.gV = arena_ptr<float>(g, J.gV);
    a.gF = arena_ptr<float>(g, J.gF);
    a.gE = arena_ptr<float>(g, J.gE);
    a.gFu = arena_ptr<float>(adj, J.gFu);
    a.gP = arena_ptr<float>(adj, J.gP);
    a.W = arena_ptr<float>(adj, J.W);
    a.VP = J.vp >= 0 ? arena_ptr<VParam<float>>(adj, J.vp) : reinterpret_cast<VParam<float> *>(a.W);
    a.gEu = a.W ? a.W + 4 * (int64_t)J.V : nullptr;
    a.erep = arena_ptr<const uint64_t>(arena, J.erep);
    a.rare = arena_ptr<int32_t>(arena, J.rare);
    a.rare_cap = J.rare_cap;
    c = arena_cage(arena, J);
}

// The job of workgroup b: a job table's start array holds the first
// workgroup of each of its n jobs and the total (n + 1 entries), then the
// job of every workgroup, so one load replaces a binary search (jobs
// without workgroups own none)
__device__ __forceinline__ int find_job(const int32_t *__restrict__ start, int n, uint32_t b) {
    return start[n + 1 + (int)b];
}

// One thread per element of each job: op(job, i) with i the job-local index
// (workgroups in contiguous ranges per XCD).
template <class Job, class Op>
__global__ __launch_bounds__(256) void k_batch(const Job *__restrict__ jobs, const int32_t *__restrict__ start, int n,
                                               Op op) {
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x, 1);
    const int k = find_job(start, n, lb);
    const int64_t i = (int64_t)(lb - (uint32_t)start[k]) * 256 + threadIdx.x;
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
void batch_level_patch(const BatchTable &t, char *arena, int d, int gather, int creased, int topo, int nt, int hpt,
                       int remap, hipStream_t st);
void batch_level_deferred(const BatchTable &t, char *arena, int d, int gather, int creased, hipStream_t st);
// fused children + face/edge points (workgroups of nt threads, one per
// nt * tiles halfedges), then vertex points (one workgroup per 256 * scan
// halfedges; scan 0: the owners the fused kernel marked, one workgroup per
// 16 mask words, kernel compiled for `waves` waves per SIMD)
void batch_level_fused(const BatchTable &t, char *arena, int d, int gather, int creased, int topo, int remap, int nt,
                       int tiles, hipStream_t st);
void batch_level_vertices(const BatchTable &t, char *arena, int d, int gather, int creased, int scan, int remap,
                          int waves, hipStream_t st);
// after a scan-0 vertex launch: each job's rare list, one thread per entry
// (`start` counts workgroups of 256 list entries)
void batch_level_vertices_rare(const BatchTable &t, char *arena, int d, int gather, int creased, hipStream_t st);

// adjoint launches (cc_adjoint.hip)
// one level's adjoint: vertex parameters (when `params`; else the ones a
// previous call left) and vertex gathers over `quads` (the owner passes scan
// 4 halfedges per thread), the quad pass one thread per halfedge over
// `halfedges`
void batch_adj_level(const BatchTable &quads, const BatchTable &halfedges, const BatchTable &rare, char *arena, char *adj,
                     int d, int creased,
                     int params, hipStream_t st);
void batch_adj_cage(const BatchTable &t, char *arena, char *adj, int creased, hipStream_t st);
// the cage step as ordered gathers (every cage has its ring lists): one
// thread per cage halfedge, edge, face and vertex in turn
void batch_adj_cage_gather(const BatchTable &halfedges, const BatchTable &edges, const BatchTable &faces,
                           const BatchTable &vertices, char *arena, char *adj, int creased, hipStream_t st);

}  // namespace cch
