// Fused per-level kernels (levels d >= 1) and their launchers.
#pragma once

#include "cc_common.h"

namespace cch {

enum { kStoreDirect = 0, kStoreLds = 1 };

// Threads (parent halfedges) per workgroup of the fused level kernel by
// default: a workgroup covers the descendants of one level d-5 face (a 16 x
// 16 patch of quads on a regular grid), and the twin quads of all but the
// patch's border halfedges come from its LDS rather than from memory.  The
// autotuner may pick 256 or 512 per level, and 2 or 4 tiles (of one
// workgroup's size) processed in turn by each workgroup.
constexpr int kFusedBlock = 1024;
// waves per SIMD the masked vertex kernel is compiled for by default (it
// needs 56 registers: 8 fit without spilling); the autotuner may pick 4 or 6
constexpr int kVertexWaves = 8;

// Level-d inputs and level-(d+1) outputs of one fused launch.
struct FusedArgs {
    const He *he;       // level d halfedges
    He *out_he;         // level d+1 halfedges
    const Crease *cr;   // level d creases
    const P3 *old;      // level d points
    P3 *vp, *fp, *ep;   // level d+1 points: vertex | face | edge slices
    int32_t H, V, F, E, C;
    int xcd_remap;
    int uv_refine;      // patch kernel with uvMode 3: 1 refine uvIDs, 0 write 0
    // vertices the patch-local kernel leaves to the deferred kernel: bit l of
    // word w marks halfedge 64w + l as the representative of such a vertex;
    // the patch kernel writes every one of the ceil(H / 64) words
    uint64_t *defer_mask;
    // representatives of the level-d vertices (CatmullClark.h:1336), one bit
    // per halfedge, written by the fused kernel when set, read by the masked
    // vertex kernel
    uint64_t *owner_mask;
    // when set: the masked vertex kernel handles the vertices whose ring
    // follows from the child numbering and appends the others (cage-level
    // and boundary fans) to this list, rare[0] their count (cleared by the
    // fused kernel), rare[1 .. rare_cap] their representatives; the rare
    // kernel then finishes them
    int32_t *rare;
    int32_t rare_cap;
    // Edge-representative bits of levels 1 .. (word offsets LevelTab::R):
    // bit x of level L is set when halfedge x is the reference's
    // representative of its edge (ccs_EdgeToHalfedgeID, CatmullClark.h:1206).
    // erep_write: this launch derives level d+1's bits from level d's (and
    // writes level 1's when d = 1); erep_read: levels 1 .. d-1 hold valid
    // bits, so the owner test reads them instead of descending the edge's
    // ancestry.
    uint64_t *erep = nullptr;
    int erep_read = 0, erep_write = 0;
    // 0: owner_mask already holds this level's bits (a points-only pass over
    // an unchanged topology), the fused kernel leaves them
    int owner_write = 1;
};

// topo: write child twin/edge/vertex IDs; uvMode: 0 leave uvIDs, 1 refine
// them, 2 write 0; points: face and edge points of level d+1.
// (c, t, d: the cage and counts for the owner bits when a.owner_mask is set)
// nt: threads (halfedges) per workgroup, 256, 512 or 1024 with LDS-staged
// stores (direct stores always use 1024); tiles: tiles of nt halfedges each
// workgroup processes in turn, 1, 2 or 4
void launch_level_fused(const FusedArgs &a, const DCage &c, const LevelTab &t, int d, int gather, int creased, int topo,
                        int uvMode, int points, int store, int nt, int tiles, hipStream_t st);
// base pointers of the halfedges of every level (index 1..D)
struct LevelPtrs {
    const He *he[kMaxDepth + 1];
};

// The level pointers of one job of a batch: byte offsets into the batch's
// arena, -1 for an absent level.  Kernels copy them into a LevelPtrs in LDS
// lane by lane straight from the job table (a per-lane indexed local copy
// would live in scratch memory).
struct JobLevelPtrs {
    const int64_t *off;
    char *arena;
};

__device__ __forceinline__ void lds_fill(LevelPtrs &dst, const JobLevelPtrs &src) {
    for (int i = (int)threadIdx.x; i <= kMaxDepth; i += (int)blockDim.x) {
        const int64_t o = src.off[i];
        dst.he[i] = o < 0 ? nullptr : reinterpret_cast<const He *>(src.arena + o);
    }
}

// Patch-local fused kernel: children, face, edge and vertex points of level
// d+1; vertices whose ring leaves the tile are marked in a.defer_mask and
// finished by launch_level_deferred.
// nt: halfedges per tile (128, 256, 512); hpt: tiles per workgroup
void launch_level_patch(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d, int gather,
                        int creased, int topo, int uvMode, int nt, int hpt, hipStream_t st);
void launch_level_deferred(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d,
                           int gather, int creased, hipStream_t st);

// vertex points of level d+1 (after launch_level_fused with points); fan
// selects the structurally enumerated rings over the ring walk
// (scan: halfedges each thread of the compacted kernel tests, 2, 4 or 8;
// 0: the owners marked in a.owner_mask by the fused kernel, with the masked
// kernel compiled for `waves` waves per SIMD, 4, 6 or 8)
void launch_level_vertices(const FusedArgs &a, const DCage &c, const LevelTab &t, const LevelPtrs &lp, int d,
                           int gather, int creased, int fan, int scan, int waves, hipStream_t st);

}  // namespace cch
