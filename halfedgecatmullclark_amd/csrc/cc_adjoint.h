// Adjoint of the point refinement (transpose of the scatter rules).
#pragma once

#include "cc_common.h"
#include "cc_fused.h"

namespace cch {

// Gradients of one level step d -> d+1: gV, gF, gE are the slices of the
// level-(d+1) gradient (the face and edge slices receive the vertex and edge
// rules' shares), gP receives the level-d gradient.
template <typename T>
struct AdjArgs {
    const He *he;
    const Crease *cr;
    const T *sd;   // sharpness of this level in T, or null to read cr
    int32_t H, C, V;   // halfedges, creases and vertices of the level
    T *gV, *gF, *gE;
    T *gP;
    T *W;          // per-halfedge vertex-rule weights (a, b, g, dl): 4 H scalars
    // edge-representative bits of levels 1..d-1 (FusedArgs::erep), or null
    const uint64_t *erep = nullptr;
};

struct AdjPlan {
    const Counts *n;
    int D;
    const He *he;        // cumulative halfedges of levels 1..D
    const Crease *cr;    // cumulative creases of levels 1..D
    DCage cage;
    LevelTab tab;
    LevelPtrs lp;
    int64_t per_vertex_max;   // levels with at most this many vertices: vertex passes one thread per vertex
    const uint64_t *erep;     // edge-representative bits of levels 1..D-2 from the forward pass, or null
};

// dL/d(cage points) from dL/d(level-D points): V_D x 3 in, V_0 x 3 out.
// bufA holds V_D x 3 and bufB V_{D-1} x 3 scalars of scratch, W 4 max(H_{D-1}, H_0).
// When sharp is given (E_0 + Ccum(D-1) scalars) the sharpness of every level
// is refined again in T and used instead of the float crease records.
template <typename T>
void adjoint_levels(const AdjPlan &pl, const T *gradFinest, T *gradCage, T *bufA, T *bufB, T *W, T *sharp,
                    int creased, hipStream_t st);

}  // namespace cch
