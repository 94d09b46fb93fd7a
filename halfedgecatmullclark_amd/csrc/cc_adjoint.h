// Adjoint of the point refinement (transpose of the scatter rules).
#pragma once

#include "cc_common.h"
#include "cc_fused.h"

namespace cch {

// The vertex rule of one level-d vertex, enough to give the weights of any
// halfedge of its ring from that halfedge's own and previous sharpness:
// valence, boundary and rule (0 smooth, 1 corner, 2 crease) packed in nb.
// The smooth and corner rules' weights are the same for every ring
// halfedge, so s1 and s2 hold them already divided out (s1 the weight a of
// the vertex, s2 the weight g of the face; the edge's b is -4 g); the crease
// rule keeps the first two nonzero sharpness values of the ring (s1 on edge
// e1): exactly two ring edges are sharp, so their sum is s1 + s2 and a ring
// edge other than e1 with nonzero sharpness is the other.
template <typename T>
struct VParam {
    T s1, s2;
    int32_t nb, e1;
};

// Gradients of one level step d -> d+1: gV, gF, gE are the slices of the
// level-(d+1) gradient, gP receives the level-d gradient.
template <typename T>
struct AdjArgs {
    const He *he;
    const Crease *cr;
    const T *sd;   // sharpness of this level in T, or null to read cr
    int32_t H, C, V;   // halfedges, creases and vertices of the level
    T *gV, *gF, *gE;
    T *gP;
    // scratch: the cage step keeps per-halfedge vertex-rule weights
    // (a, b, g, dl) here, 4 H scalars; a level d >= 1 keeps the parameters
    // of its vertices (VP, 4 V scalars) followed by the edge gradients
    // after the vertex rule's shares (gEu, 3 E scalars)
    T *W;
    VParam<T> *VP;
    T *gEu;
    // the face gradients after the quad pass (3 F scalars), or null: the
    // quad pass then updates gF in place
    T *gFu = nullptr;
    // edge-representative bits of levels 1..d-1 (FusedArgs::erep), or null
    const uint64_t *erep = nullptr;
    // when set: the vertex pass leaves the vertices whose ring does not
    // follow from the child numbering (cage vertices, cage faces other than
    // quads with consecutive halfedges, boundary edge vertices) to the rare
    // kernel through this list, rare[0] their
    // count (cleared by the quad pass), rare[1 .. rare_cap] their owner
    // halfedges
    int32_t *rare = nullptr;
    int32_t rare_cap = 0;
};

// scalars of adjoint scratch (AdjArgs::W) a cage refined to depth D needs:
// the cage step's weights and every level's parameters and edge gradients
// (face_out: and its face gradients, AdjArgs::gFu), a multiple of 4 so that
// consecutive cages' scratch stays 16-byte aligned
inline int64_t adj_scratch_scalars(const Counts &n, int D, bool face_out = false) {
    int64_t m = 4 * n.H[0];
    for (int d = 1; d < D; ++d) {
        const int64_t need = 4 * n.V[d] + 3 * n.E[d] + (face_out ? 3 * n.F[d] : 0);
        if (need > m) m = need;
    }
    return (m + 3) & ~int64_t(3);
}

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
    // float adjoint: the vertex rules' parameters of levels 1..D-1 (VParam,
    // level d at n.offV[d]), kept across calls for the same topology and
    // creases; vp_valid: they hold this topology's, so the parameter pass is
    // skipped (null: the parameters go to the scratch W, every call)
    void *vp_cache = nullptr;
    bool vp_valid = false;
    // the vertex pass's rare-vertex list (AdjArgs::rare), or null
    int32_t *rare = nullptr;
    int32_t rare_cap = 0;
};

// dL/d(cage points) from dL/d(level-D points): V_D x 3 in (only read),
// V_0 x 3 out.  bufA holds V_{D-2} x 3 scalars of scratch (V_D x 3 when
// D = 1), bufB V_{D-1} x 3, W adj_scratch_scalars(n, D, true).
// When sharp is given (E_0 + Ccum(D-1) scalars) the sharpness of every level
// is refined again in T and used instead of the float crease records.
template <typename T>
void adjoint_levels(const AdjPlan &pl, const T *gradFinest, T *gradCage, T *bufA, T *bufB, T *W, T *sharp,
                    int creased, hipStream_t st);

}  // namespace cch
