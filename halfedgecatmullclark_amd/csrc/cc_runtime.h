// The device mirror of a cc_Subd (the opaque ccd_Subd of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "cc_internal.h"
#include "cc_adjoint.h"
#include "cc_fused.h"
#include "cc_kernels.h"

struct ccd_Subd {
    struct Timed {
        const char *name;
        hipEvent_t start, stop;
    };

    int device = 0;
    int32_t D = 0, V0 = 0, U0 = 0, H0 = 0, E0 = 0, F0 = 0;
    cch::Counts n;
    cch::LevelTab tab;
    // cage (level 0) on the device
    cch::CageHe *c_he = nullptr;
    int32_t *c_v2h = nullptr, *c_e2h = nullptr, *c_f2h = nullptr;
    cch::P3 *c_pts = nullptr;
    cch::UV *c_uvs = nullptr;
    cch::Crease *c_cr = nullptr;
    // the cage's vertex and face ring lists (cch::DCage) in one device
    // buffer; rings_ok false: the kernels walk the cage
    int32_t *rings = nullptr;
    size_t rings_cap = 0;
    size_t o_vro = 0, o_vr = 0, o_vb = 0, o_fro = 0, o_fr = 0, o_voo = 0, o_vo = 0;
    // the masked vertex kernel's list of vertices without a ring from the
    // child numbering (FusedArgs::rare): at most V0 + F0 + (boundary
    // halfedges of the cage) * 2^(D-1) per level
    int32_t *rare = nullptr;
    int32_t rare_cap = 0;
    bool rings_ok = false;
    // no vertex of this cage can reach the rare lists (rings_without_rare):
    // the vertex passes get a capacity of 0 and no rare kernel runs
    bool rare_free = false;
    // levels 1..D, cumulative, laid out as cc_Subd's arrays
    cch::He *he = nullptr;
    cch::Crease *cr = nullptr;
    cch::P3 *pts = nullptr;
    hipStream_t stream = nullptr;
    // adjoint scratch (allocated on first use, per scalar size)
    void *adj_a = nullptr, *adj_b = nullptr, *adj_w = nullptr, *adj_in = nullptr, *adj_out = nullptr,
         *adj_sharp = nullptr;
    // the float adjoint's vertex-rule parameters of levels 1..D-1 (16 bytes
    // per vertex), valid for the creased (1) or uncreased (0) rules of the
    // current topology and creases, -1 when they must be computed again
    // (a topology or crease refinement, a cage upload)
    void *adj_vp = nullptr;
    int adj_vp_key = -1;
    // deferred-vertex masks of the patch kernels, one word per 64 halfedges
    uint64_t *defer_mask = nullptr;
    // owner bits of levels 1..D-1 (FusedArgs::owner_mask; level d from word
    // own_off[d]), and which levels hold the bits of the current cage's
    // topology: points-only passes reuse them
    uint64_t *own_masks = nullptr;
    int64_t own_off[cch::kMaxDepth + 1] = {};
    bool own_valid[cch::kMaxDepth + 1] = {};
    // edge-representative bits of levels 1..D-2 (FusedArgs::erep; the owner
    // test of level d reads levels up to d-1), and which levels hold bits of
    // the current cage's topology
    uint64_t *erep = nullptr;
    bool erep_valid[cch::kMaxDepth + 1] = {};
    size_t adj_scalar = 0;
    // halfedge normals (allocated on first use, H_D points each): the corner
    // normals the vertex normals sum, and the host entry points' output
    cch::P3 *nrm_tmp = nullptr, *nrm_out = nullptr;
    bool timing = false;
    int opt_fused = 1, opt_store = cch::kStoreLds, opt_remap = 1, opt_fan = 2;
    int32_t opt_adj_vertex = 1 << 20;   // CCD_OPT_ADJ_VERTEX
    // per-level kernels of the fused paths (options or ccd_Autotune): kind 1
    // fused children + face/edge points (fb threads per workgroup, ft tiles
    // of fb halfedges per workgroup), then vertex points (scan 0: the
    // representatives the fused kernel marked, kernel compiled for vw waves
    // per SIMD; 2, 4, 8: halfedges each thread of the compacted kernel
    // tests); kind 2 patch-local (nt halfedges per tile, hpt tiles per
    // workgroup) plus deferred vertices
    struct LevelCfg {
        int kind = 1, scan = 0, nt = 256, hpt = 1;
        int fb = cch::kFusedBlock, ft = 1, vw = cch::kVertexWaves;
    };
    LevelCfg cfg[cch::kMaxDepth + 1];
    std::vector<Timed> timed;
    // ccd_Refine's launch sequences captured as graphs (CCB_OPT_GRAPH), keyed
    // by the steps and by the levels whose edge-representative bits were
    // valid at capture; dropped when an option, a level's kernels or the
    // cage change
    std::map<int64_t, hipGraphExec_t> graphs;
    int opt_graph = 1;

    bool alloc(const cc_Mesh *cage, int32_t depth);
    int32_t autotune(int32_t steps, hipStream_t st);
    void free_all();
    bool upload_cage(const cc_Mesh *cage, hipStream_t st);
    cch::DCage cage_view(const cch::P3 *points) const;
    cch::Lvl level_view(int d, const cch::P3 *levelPts, cch::P3 *nextPts) const;
    cch::FusedArgs fused_args(int d, const cch::P3 *points) const;
    cch::LevelPtrs level_ptrs() const;
    void cage_points(int32_t steps, const cch::P3 *cagePts, cch::P3 *out, hipStream_t st);
    bool refine(int32_t steps, int dmax, hipStream_t st);
    bool refine_graph(int32_t steps, hipStream_t st);
    void drop_graphs();
    bool refine_topology(int32_t steps, int dmax, hipStream_t st);
    bool refine_points(int32_t steps, int dmax, const cch::P3 *cagePts, cch::P3 *out, hipStream_t st);
    void level_patch(int d, int32_t steps, int topo, int uvMode, cch::P3 *points, hipStream_t st);
    void level_fused(int d, int32_t steps, int topo, int uvMode, cch::P3 *points, hipStream_t st);
    void begin_timing();
    bool adjoint(const void *gradFinest, void *gradCage, int32_t steps, hipStream_t st);
    void zero(float *p, int64_t count, hipStream_t st);
    template <typename K, typename... Args>
    void launch(const char *name, K kernel, int64_t threads, hipStream_t st, Args... args);
    template <typename Fn>
    void launch_timed(const char *name, int64_t threads, hipStream_t st, Fn &&fn);
    bool normals(cch::P3 *out, bool smooth, hipStream_t st);
};

namespace cch {
// the device mirror of a host cc_Subd (created on first use, freed by
// ccs_Release), or null with the error channel set
ccd_Subd *host_mirror(const cc_Subd *subd);
}  // namespace cch
