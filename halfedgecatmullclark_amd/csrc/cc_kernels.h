// Kernel declarations shared by the launchers.
#pragma once

#include "cc_common.h"

namespace cch {

// Level-d view for the point kernels (d >= 1): inputs of level d and the
// three output slices of level d+1.
struct Lvl {
    const He *he;        // level d halfedges
    const Crease *cr;    // level d creases
    const P3 *old;       // level d points
    P3 *vp, *fp, *ep;    // level d+1 points: vertex | face | edge slices
    int32_t H, V, F, E, C;
};

__global__ void k_cage_halfedges(DCage c, He *out, int topo, int uvs);
__global__ void k_level_halfedges(const He *__restrict__ in, He *__restrict__ out, int32_t H, int32_t V,
                                  int32_t F, int32_t E, int topo, int uvs);
__global__ void k_creases(const Crease *__restrict__ in, int32_t C, Crease *__restrict__ out);
__global__ void k_zero(float *p, int64_t n);
// *slot = p (one lane of one workgroup)
__global__ void k_set_address(const void **slot, const void *p);

__global__ void k_cage_faces(DCage c, P3 *fp, int gather);
__global__ void k_cage_edges(DCage c, const P3 *fp, P3 *ep, int gather, int creased);
__global__ void k_cage_vertices_gather(DCage c, const P3 *fp, const P3 *ep, P3 *vp, int creased);
__global__ void k_cage_vertices_scatter(DCage c, const P3 *fp, const P3 *ep, P3 *vp, int creased);

__global__ void k_faces(Lvl L, int gather);
__global__ void k_edges_scatter(Lvl L, int creased);
__global__ void k_edges_gather(Lvl L, DCage c, LevelTab t, int d, int creased);
__global__ void k_vertices_gather(Lvl L, DCage c, LevelTab t, int d, int creased);
__global__ void k_vertices_scatter(Lvl L, int creased);

}  // namespace cch
