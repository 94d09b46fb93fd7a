// Per-halfedge normals of the finest level, the step after refinement that
// the reference defines only in its GL shaders
// (glsl/CatmullClark_Scatter.glsl:931-960):
//   corner normal of h  normalize(cross(P(next h) - P(h), P(prev h) - P(h)))
//   vertex normal of h  normalize(corner(h) + the corner normals of the
//                       halfedges met walking it = twin(prev(it)) from h),
//                       or corner(h) when the walk reaches a boundary
// with normalize(v) = v * (1 / sqrt(dot(v, v))), correctly rounded, in the
// shaders' operation order (built with -ffp-contract=off).
//
// Two kernels: one thread per halfedge forms its corner normal from its
// quad's points (exchanged by DPP within the quad); the vertex normals then
// walk the ring over the stored corner normals.
#include <stdio.h>

#include "cc_fan.h"
#include "cc_internal.h"
#include "cc_runtime.h"

using namespace cch;

namespace {

__device__ __forceinline__ P3 sub3(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ P3 cross3(P3 a, P3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ P3 normalize3(P3 v) {
    const float dd = v.x * v.x + v.y * v.y + v.z * v.z;
    const float inv = 1.0f / sqrtf(dd);
    return mul(v, inv);
}

// corner normals, one thread per halfedge of the level (H a multiple of 4,
// blocks of 256: the four lanes of a quad are live together; workgroups in
// contiguous ranges per XCD, so the points a vertex's quads share meet in
// one L2)
__global__ __launch_bounds__(256) void k_corner_normals(const He *__restrict__ he, const P3 *__restrict__ pts,
                                                        int32_t H, P3 *__restrict__ out) {
    const int32_t h = (int32_t)(xcd_block(blockIdx.x, gridDim.x, 1) * blockDim.x + threadIdx.x);
    const bool live = h < H;
    const int4 x = live ? he_gld(he, h) : make_int4(0, 0, 0, 0);
    const P3 v0 = gld(pts + x.z);
    const P3 v2 = {quad_relf<kQuadNext>(v0.x), quad_relf<kQuadNext>(v0.y), quad_relf<kQuadNext>(v0.z)};
    const P3 v1 = {quad_relf<kQuadPrev>(v0.x), quad_relf<kQuadPrev>(v0.y), quad_relf<kQuadPrev>(v0.z)};
    if (live) out[h] = normalize3(cross3(sub3(v2, v0), sub3(v1, v0)));
}

// One owner per level-d vertex, decided from the child numbering as the
// structural owner test does, except that an edge vertex born at level >= 2
// is owned by child 1 of the larger parent halfedge of its edge (or the
// parent on a boundary): one parent twin read instead of the edge's
// representative descent.  Any owner gives the same fan as a cyclic list.
__device__ __forceinline__ bool normal_owner(int32_t h, int d, const DCage &c, const He *const *levels) {
    const int k = zero_digits(h, d);
    const int32_t q = h >> (2 * k);
    const int lvl = d - k;
    const int digit = q & 3;
    const int32_t p = q >> 2;
    if (digit == 3) return false;
    if (lvl >= 2) {
        if (digit == 2) return (p & 3) == 0;
        const int32_t tw = levels[lvl - 1][p].twin;
        return tw < 0 || p > tw;
    }
    const CageHe x = c.he[p];
    if (digit == 0) return c.v2h[x.vert] == p;
    if (digit == 2) return c.f2h[x.face] == p;
    return c.e2h[x.edge] == p;
}

// Vertex normals of every halfedge of a vertex's fan, from its owner.  The
// fan lists the ring in the walk order twin(prev(x)), so the walk from
// member i visits members i+1, ..., i-1: its normal is
// normalize(c_i + ((0 + c_{i+1}) + ... + c_{i-1})), the per-halfedge walk's
// own sums.  On a boundary every member's walk ends at the boundary and its
// normal is its corner normal.
struct VertexNormalsFn {
    const P3 *corner;
    P3 *out;
    const DCage &c;
    __device__ void operator()(const Fan &f) const {
        const int n = f.n;
        if (f.boundary) {
            for (int i = 0; i < n; ++i) {
                const int32_t r = fan_at(f, c, i);
                out[r] = corner[r];
            }
            return;
        }
        if (n <= 4) {
            int32_t r[4];
            P3 cn[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) r[q] = fan_at(f, c, q < n ? q : 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) cn[q] = corner[r[q]];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (i >= n) break;
                P3 avg = zero3();
#pragma unroll
                for (int k = 1; k < 4; ++k)
                    if (k < n) avg = add(avg, pick4(cn[0], cn[1], cn[2], cn[3], (i + k) % n));
                out[r[i]] = normalize3(add(cn[i], avg));
            }
            return;
        }
        for (int i = 0; i < n; ++i) {
            const int32_t ri = fan_at(f, c, i);
            P3 avg = zero3();
            for (int k = 1; k < n; ++k) avg = add(avg, corner[fan_at(f, c, (i + k) % n)]);
            out[ri] = normalize3(add(corner[ri], avg));
        }
    }
};

// vertex normals by owner: a block lists the owners among its 1024
// halfedges in LDS (halfedge order), then each thread finishes one vertex
__global__ __launch_bounds__(256) void k_vertex_normals_owners(DCage c, LevelTab t, LevelPtrs lp, int d, int32_t H,
                                                               const P3 *__restrict__ corner, P3 *__restrict__ out) {
    __shared__ int32_t owners[1024];
    __shared__ int32_t nown;
    __shared__ LevelTab st;
    __shared__ LevelPtrs slp;
    lds_fill(st, t);
    lds_fill(slp, lp);
    const int32_t base = (int32_t)(xcd_block(blockIdx.x, gridDim.x, 1) * 1024u);
    if (threadIdx.x == 0) nown = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int32_t h = base + k * 256 + (int32_t)threadIdx.x;
        if (h < H && normal_owner(h, d, c, slp.he)) owners[atomicAdd(&nown, 1)] = h;
    }
    __syncthreads();
    const int32_t count = nown;
    const VertexNormalsFn fn{corner, out, c};
    for (int32_t i = (int32_t)threadIdx.x; i < count; i += 256) {
        Fan f;
        (void)marked_fan(owners[i], d, c, st, slp.he, f);
        fn(f);
    }
}

// vertex normals from the corner normals, one walk per halfedge
__global__ __launch_bounds__(256) void k_vertex_normals(const He *__restrict__ he, const P3 *__restrict__ corner,
                                                        int32_t H, P3 *__restrict__ out) {
    const int32_t h = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (h >= H) return;
    const P3 n = corner[h];
    P3 avg = zero3();
    int32_t it = he[qprev(h)].twin;
    for (int g = 0; it >= 0 && it != h && g < kWalkCap; ++g) {
        avg = add(avg, corner[it]);
        it = he[qprev(it)].twin;
    }
    out[h] = it < 0 ? n : normalize3(add(n, avg));
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// normals of level D from the mirror's level-D halfedges and points
bool ccd_Subd::normals(P3 *out, bool smooth, hipStream_t st) {
    const int64_t H = n.H[D];
    const He *lhe = he + n.offH[D];
    const P3 *lpts = pts + n.offV[D];
    P3 *corner = out;
    if (smooth) {
        if (!nrm_tmp && hip_failed(hipMalloc(&nrm_tmp, sizeof(P3) * (size_t)H), "hipMalloc(corner normals)"))
            return false;
        corner = nrm_tmp;
    }
    hipLaunchKernelGGL(k_corner_normals, dim3(blocks_for(H)), dim3(256), 0, st, lhe, lpts, (int32_t)H, corner);
    if (smooth && D >= 1 && rings_ok) {
        // by owner: each ring's corner normals read once
        hipLaunchKernelGGL(k_vertex_normals_owners, dim3((unsigned)((H + 1023) / 1024)), dim3(256), 0, st,
                           cage_view(nullptr), tab, level_ptrs(), D, (int32_t)H, (const P3 *)corner, out);
    } else if (smooth) {
        hipLaunchKernelGGL(k_vertex_normals, dim3(blocks_for(H)), dim3(256), 0, st, lhe, (const P3 *)corner,
                           (int32_t)H, out);
    }
    return !hip_failed(hipGetLastError(), "halfedge normals");
}

namespace {

// the host entry points: the level-maxDepth halfedges and points of the
// subd's host arrays go to its device mirror, the normals come back
void host_normals(const cc_Subd *subd, cc_VertexPoint *normals, bool smooth, const char *fname) {
    clear_error();
    if (!normals) {
        set_error(11, "null normals");
        fprintf(stderr, "cc: %s: null normals\n", fname);
        return;
    }
    ccd_Subd *s = host_mirror(subd);
    if (!s) {
        fprintf(stderr, "cc: %s: %s\n", fname, ccs_LastErrorString());
        return;
    }
    DeviceGuard g(s->device);
    hipStream_t st = s->stream;
    const Counts &n = s->n;
    const int D = s->D;
    const int64_t H = n.H[D];
    if (!s->nrm_out && hip_failed(hipMalloc(&s->nrm_out, sizeof(P3) * (size_t)H), "hipMalloc(normals)")) return;
    if (hip_failed(hipMemcpyAsync(s->he + n.offH[D], subd->halfedges + n.offH[D], sizeof(He) * (size_t)H,
                                  hipMemcpyHostToDevice, st), "upload halfedges") ||
        hip_failed(hipMemcpyAsync(s->pts + n.offV[D], subd->vertexPoints + n.offV[D], sizeof(P3) * (size_t)n.V[D],
                                  hipMemcpyHostToDevice, st), "upload points"))
        return;
    if (!s->normals(s->nrm_out, smooth, st)) return;
    if (hip_failed(hipMemcpyAsync(normals, s->nrm_out, sizeof(P3) * (size_t)H, hipMemcpyDeviceToHost, st),
                   "download normals") ||
        hip_failed(hipStreamSynchronize(st), fname))
        return;
}

}  // namespace

extern "C" {

void ccs_HalfedgeNormals(const cc_Subd *subd, cc_VertexPoint *normals) {
    host_normals(subd, normals, true, "ccs_HalfedgeNormals");
}

void ccs_HalfedgeNormals_Fast(const cc_Subd *subd, cc_VertexPoint *normals) {
    host_normals(subd, normals, false, "ccs_HalfedgeNormals_Fast");
}

int32_t ccd_HalfedgeNormals(ccd_Subd *s, float *devNormals, int32_t smooth, void *stream) {
    clear_error();
    if (!s || !devNormals) {
        set_error(11, "ccd_HalfedgeNormals: null argument");
        return last_error();
    }
    DeviceGuard g(s->device);
    if (!s->normals(reinterpret_cast<P3 *>(devNormals), smooth != 0, (hipStream_t)stream))
        return last_error() ? last_error() : 1;
    return 0;
}

}  // extern "C"
