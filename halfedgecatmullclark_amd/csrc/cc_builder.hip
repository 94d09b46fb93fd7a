// The cage builder on the GPU: an indexed polygon mesh becomes a halfedge
// cage (cc_Mesh) with the conventions of the reference's converter
// examples/obj_to_ccm.c, every step data-parallel:
//   - halfedge records per face (next, prev, face, vertex, uv), :281-313;
//   - twins: a key v0 + V v1 per halfedge, the reference's bitonic network
//     over a power-of-two table (the same compare-exchange schedule, so
//     equal keys land where the reference puts them), then its binary search
//     for v1 + V v0, :88-190;
//   - edges: one bit per halfedge with h > twin(h), an exclusive scan gives
//     the edge IDs (the rank of max(h, twin)) and edge -> halfedge, :364-402;
//   - vertex -> halfedge: the outgoing boundary halfedge of a boundary
//     vertex, else the largest outgoing halfedge, :417-450;
//   - creases: tags matched to edges by their two vertices (the last tag of
//     an edge wins, as in file order), boundary edges sharpness 16, then the
//     crease neighbours, :197-279, :538-561.
// ccm_LoadObj parses an OBJ file on the host exactly as the converter does
// (:460-570) and builds the cage here.
#include <stdio.h>
#include <string.h>

#include <vector>

#include <hipcub/hipcub.hpp>

#include "cc_internal.h"

using namespace cch;

namespace {

// the reference's twin table entry (16 bytes with padding)
struct TwinEntry {
    int32_t id;
    uint64_t hash;
};

__device__ __forceinline__ uint32_t gid() { return blockIdx.x * blockDim.x + threadIdx.x; }
unsigned blocks_for(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// per face: its halfedges' records (the next of the last is the first)
__global__ void k_face_records(const int32_t *__restrict__ off, const int32_t *__restrict__ fv,
                               const int32_t *__restrict__ fuv, int32_t F, CageHe *__restrict__ he) {
    const int32_t f = (int32_t)gid();
    if (f >= F) return;
    const int32_t b = off[f], e = off[f + 1];
    for (int32_t h = b; h < e; ++h) {
        CageHe x;
        x.twin = -1;
        x.next = h == e - 1 ? b : h + 1;
        x.prev = h == b ? e - 1 : h - 1;
        x.face = f;
        x.edge = -1;
        x.vert = fv[h];
        x.uv = fuv ? fuv[h] : 0;
        he[h] = x;
    }
}

__global__ void k_twin_keys(const CageHe *__restrict__ he, int32_t H, int32_t V, uint32_t N, TwinEntry *t) {
    const uint32_t i = gid();
    if (i >= N) return;
    if (i < (uint32_t)H) {
        const uint64_t v0 = (uint64_t)he[i].vert, v1 = (uint64_t)he[he[i].next].vert;
        t[i].id = (int32_t)i;
        t[i].hash = v0 + (uint64_t)V * v1;
    } else {
        t[i].id = -1;
        t[i].hash = ~0ull;
    }
}

// one compare-exchange of the reference's network (obj_to_ccm.c:114-132):
// pair i of stage (d2, d1)
__device__ __forceinline__ void bitonic_step(TwinEntry *a, uint32_t i, uint32_t il, uint32_t d1, uint32_t d2) {
    const uint32_t mask = 0xFFFFFFFEu * d1;
    const uint32_t i1 = ((il << 1) & mask) | (il & ~(mask >> 1));
    const uint32_t i2 = i1 | d1;
    const TwinEntry t1 = a[i1], t2 = a[i2];
    const bool lt = t1.hash < t2.hash;
    const TwinEntry mn = lt ? t1 : t2, mx = lt ? t2 : t1;
    if ((i & d2) == 0) {
        a[i1] = mn;
        a[i2] = mx;
    } else {
        a[i1] = mx;
        a[i2] = mn;
    }
}

// one stage with d1 >= the block's pair count, in global memory
__global__ void k_bitonic_global(TwinEntry *a, uint32_t pairs, uint32_t d1, uint32_t d2) {
    const uint32_t i = gid();
    if (i < pairs) bitonic_step(a, i, i, d1, d2);
}

constexpr uint32_t kSortPairs = 512;   // pairs per workgroup of the LDS stages

// Stages (d2, d1) with d1 < pairs per block run in LDS: block b owns elements
// [2 b P, 2 b P + 2 P) (P pairs), and the pairs of such a stage stay inside
// it.  Runs every d2 in [d2_lo, d2_hi] (d1 from min(d2, P / 2) down to 1).
__global__ void k_bitonic_lds(TwinEntry *a, uint32_t P, uint32_t d2_lo, uint32_t d2_hi) {
    __shared__ TwinEntry s[2 * kSortPairs];
    const uint32_t base = 2 * blockIdx.x * P;
    for (uint32_t k = threadIdx.x; k < 2 * P; k += blockDim.x) s[k] = a[base + k];
    __syncthreads();
    const uint32_t il = threadIdx.x, i = blockIdx.x * P + il;
    for (uint32_t d2 = d2_lo; d2 <= d2_hi; d2 <<= 1) {
        for (uint32_t d1 = d2 < P ? d2 : P / 2; d1 >= 1; d1 >>= 1) {
            if (il < P) bitonic_step(s, i, il, d1, d2);
            __syncthreads();
        }
    }
    for (uint32_t k = threadIdx.x; k < 2 * P; k += blockDim.x) a[base + k] = s[k];
}

// the reference's search (obj_to_ccm.c:88-107): over `size` entries, but it
// may read entry `size` (one past), as there
__device__ __forceinline__ int32_t twin_search(const TwinEntry *t, int32_t size, uint64_t hash) {
    int32_t lo = 0, hi = size - 1;
    while (lo <= hi) {
        const int32_t c = (lo + hi) / 2;
        if (t[c].hash < hash) lo = c + 1;
        else hi = c - 1;
    }
    return t[lo].hash == hash ? t[lo].id : -1;
}

__global__ void k_twins(CageHe *he, int32_t H, int32_t V, const TwinEntry *__restrict__ t) {
    const int32_t h = (int32_t)gid();
    if (h >= H) return;
    const uint64_t v0 = (uint64_t)he[h].vert, v1 = (uint64_t)he[he[h].next].vert;
    he[h].twin = twin_search(t, H - 1, v1 + (uint64_t)V * v0);
}

__global__ void k_edge_bits(const CageHe *__restrict__ he, int32_t H, int32_t *bits) {
    const int32_t h = (int32_t)gid();
    if (h <= H) bits[h] = h < H && h > he[h].twin ? 1 : 0;
}

// edge IDs (the rank of max(h, twin)), edge -> halfedge, vertex maxima
__global__ void k_edges(CageHe *he, int32_t H, const int32_t *__restrict__ rank, int32_t *e2h, int32_t *vmax,
                        int32_t *vbnd) {
    const int32_t h = (int32_t)gid();
    if (h >= H) return;
    const int32_t t = he[h].twin;
    he[h].edge = rank[h > t ? h : t];
    if (h > t) e2h[rank[h]] = h;
    const int32_t v = he[h].vert;
    atomicMax(vmax + v, h);
    if (t < 0) atomicMax(vbnd + v, h);
}

__global__ void k_fill(int32_t *p, int32_t n, int32_t v) {
    const int32_t i = (int32_t)gid();
    if (i < n) p[i] = v;
}

__global__ void k_vertex_map(const int32_t *__restrict__ vmax, const int32_t *__restrict__ vbnd, int32_t V,
                             int32_t *v2h) {
    const int32_t v = (int32_t)gid();
    if (v < V) v2h[v] = vbnd[v] >= 0 ? vbnd[v] : vmax[v];
}

// crease tags: the edge of halfedge a -> b or b -> a records the largest tag
// index that names it
__global__ void k_crease_tags(const CageHe *__restrict__ he, int32_t H, int32_t V, const TwinEntry *__restrict__ t,
                              const int32_t *__restrict__ tagv, int32_t T, int32_t *winner) {
    const int32_t k = (int32_t)gid();
    if (k >= T) return;
    const int32_t a = tagv[2 * k], b = tagv[2 * k + 1];
    if (a < 0 || a >= V || b < 0 || b >= V) return;
    const int32_t h0 = twin_search(t, H - 1, (uint64_t)a + (uint64_t)V * (uint64_t)b);
    const int32_t h1 = twin_search(t, H - 1, (uint64_t)b + (uint64_t)V * (uint64_t)a);
    if (h0 >= 0) atomicMax(winner + he[h0].edge, k);
    if (h1 >= 0) atomicMax(winner + he[h1].edge, k);
}

__global__ void k_creases(const CageHe *__restrict__ he, const int32_t *__restrict__ e2h, int32_t E,
                          const int32_t *__restrict__ winner, const float *__restrict__ tags, int sharpen,
                          Crease *cr) {
    const int32_t e = (int32_t)gid();
    if (e >= E) return;
    float s = winner[e] >= 0 ? tags[winner[e]] : 0.0f;
    if (sharpen && he[e2h[e]].twin < 0) s = 16.0f;
    cr[e] = Crease{e, e, s};
}

__device__ __forceinline__ int32_t next_vertex_he(const CageHe *he, int32_t h) {
    const int32_t t = he[h].twin;
    return t >= 0 ? he[t].next : -1;
}

// crease neighbours (obj_to_ccm.c:197-253)
__global__ void k_crease_neighbors(const CageHe *__restrict__ he, const int32_t *__restrict__ e2h, int32_t E,
                                   Crease *cr) {
    const int32_t e = (int32_t)gid();
    if (e >= E) return;
    if (!(cr[e].sharp > 0.0f)) return;
    const int32_t h = e2h[e], nx = he[h].next;
    int32_t count = 0, cand = -1, it = next_vertex_he(he, h);
    for (int g = 0; it != h && it >= 0 && g < kWalkCap; ++g, it = next_vertex_he(he, it)) {
        if (cr[he[it].edge].sharp > 0.0f) {
            cand = he[it].edge;
            ++count;
        }
    }
    if (count == 1 && it == h) cr[e].prev = cand;
    count = 0;
    cand = -1;
    if (cr[he[nx].edge].sharp > 0.0f) {
        cand = he[nx].edge;
        ++count;
    }
    it = next_vertex_he(he, nx);
    for (int g = 0; it != nx && it >= 0 && g < kWalkCap; ++g, it = next_vertex_he(he, it)) {
        if (cr[he[it].edge].sharp > 0.0f && he[it].twin != h) {
            cand = he[it].edge;
            ++count;
        }
    }
    if (count == 1 && it == nx) cr[e].next = cand;
}

int32_t round_up_pow2(int32_t x) {
    --x;
    x |= x >> 1; x |= x >> 2; x |= x >> 4; x |= x >> 8; x |= x >> 16;
    return x + 1;
}

// device buffers of one build, freed on scope exit
struct DevBufs {
    std::vector<void *> ptrs;
    template <class T>
    T *get(size_t count) {
        void *p = nullptr;
        if (hipMalloc(&p, count ? count * sizeof(T) : 1) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return static_cast<T *>(p);
    }
    ~DevBufs() {
        for (void *p : ptrs) (void)hipFree(p);
    }
};

cc_Mesh *fail(int32_t code, const char *msg) {
    set_error(code, msg);
    fprintf(stderr, "cc: %s\n", msg);
    return nullptr;
}

// sorts the twin table with the reference's network
void bitonic_sort(TwinEntry *t, uint32_t N, hipStream_t st) {
    if (N < 2) return;
    const uint32_t pairs = N / 2;
    const uint32_t P = pairs < kSortPairs ? pairs : kSortPairs;   // pairs per LDS block
    const unsigned lblocks = pairs / P;
    // every stage with d2 < P fits one block
    hipLaunchKernelGGL(k_bitonic_lds, dim3(lblocks), dim3(P), 0, st, t, P, 1u, (P < 2 ? 1u : P / 2));
    for (uint32_t d2 = P; d2 < N && P >= 1; d2 <<= 1) {
        uint32_t d1 = d2;
        for (; d1 >= P && d1 >= 1; d1 >>= 1)
            hipLaunchKernelGGL(k_bitonic_global, dim3(blocks_for(pairs)), dim3(256), 0, st, t, pairs, d1, d2);
        if (d1 >= 1) hipLaunchKernelGGL(k_bitonic_lds, dim3(lblocks), dim3(P), 0, st, t, P, d2, d2);
    }
}

}  // namespace

extern "C" {

// Builds a cage from polygons given as faceOffsets (faceCount + 1 entries,
// from 0) into faceVertices (0-based vertex indices), with optional uvs and
// per-corner uv indices, crease tags (creaseTagVertices: 2 per tag, 0-based,
// as "t crease 2/1/0 v0 v1 s" lines give them) and boundary edges sharpened
// to 16 when sharpenBoundaries is nonzero.  Host arrays in; the cage comes
// back in host memory (ccm_Release frees it).  Assumes a manifold mesh, as
// the converter's twin search does.
cc_Mesh *ccm_CreateFromFaces(int32_t vertexCount, const cc_VertexPoint *vertexPoints, int32_t faceCount,
                             const int32_t *faceOffsets, const int32_t *faceVertices, int32_t uvCount,
                             const cc_VertexUv *uvs, const int32_t *faceUvs, int32_t creaseTagCount,
                             const int32_t *creaseTagVertices, const float *creaseTagSharpness,
                             int32_t sharpenBoundaries) {
    clear_error();
    const int32_t V = vertexCount, F = faceCount, U = uvCount, T = creaseTagCount;
    if (V < 1 || F < 1 || U < 0 || T < 0 || !vertexPoints || !faceOffsets || !faceVertices ||
        (U > 0 && !uvs) || (T > 0 && (!creaseTagVertices || !creaseTagSharpness)))
        return fail(11, "ccm_CreateFromFaces: invalid counts or null array");
    if (faceOffsets[0] != 0) return fail(12, "ccm_CreateFromFaces: faceOffsets[0] must be 0");
    for (int32_t f = 0; f < F; ++f)
        if (faceOffsets[f + 1] - faceOffsets[f] < 3) return fail(12, "ccm_CreateFromFaces: a face has fewer than 3 vertices");
    const int32_t H = faceOffsets[F];
    if (H >= (1 << 30)) return fail(12, "ccm_CreateFromFaces: too many halfedges");
    for (int32_t h = 0; h < H; ++h) {
        if (faceVertices[h] < 0 || faceVertices[h] >= V) return fail(12, "ccm_CreateFromFaces: vertex index out of range");
        if (faceUvs && faceUvs[h] < 0) return fail(12, "ccm_CreateFromFaces: negative uv index");
    }
    const uint32_t N = (uint32_t)round_up_pow2(H);
    hipStream_t st = nullptr;
    DevBufs b;
    int32_t *d_off = b.get<int32_t>(F + 1), *d_fv = b.get<int32_t>(H);
    int32_t *d_fuv = faceUvs ? b.get<int32_t>(H) : nullptr;
    CageHe *d_he = b.get<CageHe>(H);
    TwinEntry *d_t = b.get<TwinEntry>(N);
    int32_t *d_bits = b.get<int32_t>(H + 1), *d_rank = b.get<int32_t>(H + 1);
    int32_t *d_vmax = b.get<int32_t>(V), *d_vbnd = b.get<int32_t>(V), *d_v2h = b.get<int32_t>(V);
    int32_t *d_tagv = b.get<int32_t>(2 * (size_t)T);
    float *d_tags = b.get<float>(T);
    if (!d_off || !d_fv || (faceUvs && !d_fuv) || !d_he || !d_t || !d_bits || !d_rank || !d_vmax || !d_vbnd ||
        !d_v2h || !d_tagv || !d_tags)
        return fail(2, "ccm_CreateFromFaces: out of device memory");
#define CB_TRY(expr, what) \
    do { if (hip_failed((expr), (what))) return nullptr; } while (0)
    CB_TRY(hipMemcpyAsync(d_off, faceOffsets, sizeof(int32_t) * (F + 1), hipMemcpyHostToDevice, st), "upload faces");
    CB_TRY(hipMemcpyAsync(d_fv, faceVertices, sizeof(int32_t) * (size_t)H, hipMemcpyHostToDevice, st), "upload faces");
    if (faceUvs)
        CB_TRY(hipMemcpyAsync(d_fuv, faceUvs, sizeof(int32_t) * (size_t)H, hipMemcpyHostToDevice, st), "upload uvs");
    if (T > 0) {
        CB_TRY(hipMemcpyAsync(d_tagv, creaseTagVertices, sizeof(int32_t) * 2 * (size_t)T, hipMemcpyHostToDevice, st),
               "upload crease tags");
        CB_TRY(hipMemcpyAsync(d_tags, creaseTagSharpness, sizeof(float) * (size_t)T, hipMemcpyHostToDevice, st),
               "upload crease tags");
    }
    hipLaunchKernelGGL(k_face_records, dim3(blocks_for(F)), dim3(256), 0, st, d_off, d_fv, d_fuv, F, d_he);
    hipLaunchKernelGGL(k_twin_keys, dim3(blocks_for(N)), dim3(256), 0, st, d_he, H, V, N, d_t);
    bitonic_sort(d_t, N, st);
    hipLaunchKernelGGL(k_twins, dim3(blocks_for(H)), dim3(256), 0, st, d_he, H, V, (const TwinEntry *)d_t);
    hipLaunchKernelGGL(k_edge_bits, dim3(blocks_for(H + 1)), dim3(256), 0, st, (const CageHe *)d_he, H, d_bits);
    size_t scan_bytes = 0;
    CB_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, d_bits, d_rank, H + 1, st), "edge scan");
    void *d_scan = b.get<char>(scan_bytes);
    if (!d_scan) return fail(2, "ccm_CreateFromFaces: out of device memory");
    CB_TRY(hipcub::DeviceScan::ExclusiveSum(d_scan, scan_bytes, d_bits, d_rank, H + 1, st), "edge scan");
    int32_t E = 0;
    CB_TRY(hipMemcpyAsync(&E, d_rank + H, sizeof(int32_t), hipMemcpyDeviceToHost, st), "edge count");
    CB_TRY(hipStreamSynchronize(st), "edge count");
    if (E < 1) return fail(12, "ccm_CreateFromFaces: no edges");
    int32_t *d_e2h = b.get<int32_t>(E), *d_winner = b.get<int32_t>(E);
    Crease *d_cr = b.get<Crease>(E);
    if (!d_e2h || !d_winner || !d_cr) return fail(2, "ccm_CreateFromFaces: out of device memory");
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(V)), dim3(256), 0, st, d_vmax, V, -1);
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(V)), dim3(256), 0, st, d_vbnd, V, -1);
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(E)), dim3(256), 0, st, d_winner, E, -1);
    hipLaunchKernelGGL(k_edges, dim3(blocks_for(H)), dim3(256), 0, st, d_he, H, (const int32_t *)d_rank, d_e2h, d_vmax,
                       d_vbnd);
    hipLaunchKernelGGL(k_vertex_map, dim3(blocks_for(V)), dim3(256), 0, st, (const int32_t *)d_vmax,
                       (const int32_t *)d_vbnd, V, d_v2h);
    if (T > 0)
        hipLaunchKernelGGL(k_crease_tags, dim3(blocks_for(T)), dim3(256), 0, st, (const CageHe *)d_he, H, V,
                           (const TwinEntry *)d_t, (const int32_t *)d_tagv, T, d_winner);
    hipLaunchKernelGGL(k_creases, dim3(blocks_for(E)), dim3(256), 0, st, (const CageHe *)d_he, (const int32_t *)d_e2h,
                       E, (const int32_t *)d_winner, (const float *)d_tags, (int)(sharpenBoundaries != 0), d_cr);
    hipLaunchKernelGGL(k_crease_neighbors, dim3(blocks_for(E)), dim3(256), 0, st, (const CageHe *)d_he,
                       (const int32_t *)d_e2h, E, d_cr);
    CB_TRY(hipGetLastError(), "cage builder launch");
    cc_Mesh *m = ccm_Create(V, U, H, E, F);
    if (!m) return nullptr;
    const bool ok =
        !hip_failed(hipMemcpyAsync(m->halfedges, d_he, sizeof(CageHe) * (size_t)H, hipMemcpyDeviceToHost, st),
                    "download cage") &&
        !hip_failed(hipMemcpyAsync(m->edgeToHalfedgeIDs, d_e2h, sizeof(int32_t) * (size_t)E, hipMemcpyDeviceToHost, st),
                    "download cage") &&
        !hip_failed(hipMemcpyAsync(m->vertexToHalfedgeIDs, d_v2h, sizeof(int32_t) * (size_t)V, hipMemcpyDeviceToHost, st),
                    "download cage") &&
        !hip_failed(hipMemcpyAsync(m->creases, d_cr, sizeof(Crease) * (size_t)E, hipMemcpyDeviceToHost, st),
                    "download cage") &&
        !hip_failed(hipStreamSynchronize(st), "ccm_CreateFromFaces");
    if (!ok) {
        ccm_Release(m);
        return nullptr;
    }
#undef CB_TRY
    memcpy(m->faceToHalfedgeIDs, faceOffsets, sizeof(int32_t) * (size_t)F);
    memcpy(m->vertexPoints, vertexPoints, sizeof(cc_VertexPoint) * (size_t)V);
    if (U > 0) memcpy(m->uvs, uvs, sizeof(cc_VertexUv) * (size_t)U);
    return m;
}

}  // extern "C"

// ---- OBJ files (examples/obj_to_ccm.c:460-570) ----------------------------------
namespace {

// one face line: the converter tries its four vertex formats in turn, each
// as long as it matches, and takes a missing uv index as 1
int32_t obj_read_face(const char *str, std::vector<int32_t> *verts, std::vector<int32_t> *uvids, bool *bad) {
    const char *formats[] = {"%d/%*d/%*d%n", "%d//%*d%n", "%d/%*d%n", "%d%n"};
    int32_t count = 0, i = 0, n = 0, v;
    for (int fmt = 0; fmt < 4; ++fmt) {
        while (sscanf(&str[i], formats[fmt], &v, &n) == 1 && n > 0) {
            int32_t vt = 0;
            if (sscanf(&str[i], "%d/%d", &v, &vt) == 1) vt = 1;
            if (v < 1 || vt < 1) *bad = true;
            if (verts) {
                verts->push_back(v - 1);
                uvids->push_back(vt - 1);
            }
            i += n;
            ++count;
        }
    }
    if (count <= 2) *bad = true;
    return count;
}

}  // namespace

extern "C" cc_Mesh *ccm_LoadObj(const char *filename) {
    clear_error();
    FILE *f = filename ? fopen(filename, "r") : nullptr;
    if (!f) return fail(3, "ccm_LoadObj: fopen failed");
    char buf[1024];
    int32_t H = 0, V = 0, U = 0;
    bool bad = false;
    while (fgets(buf, sizeof(buf), f)) {
        if (buf[0] == 'v') {
            if (buf[1] == ' ') ++V;
            else if (buf[1] == 't') ++U;
        } else if (buf[0] == 'f') {
            H += obj_read_face(&buf[1], nullptr, nullptr, &bad);
        }
    }
    if (H <= 0 || V < 3) {
        fclose(f);
        return fail(4, "ccm_LoadObj: invalid OBJ file");
    }
    std::vector<cc_VertexPoint> pts;
    std::vector<cc_VertexUv> uvs;
    std::vector<int32_t> off{0}, fv, fuv, tagv;
    std::vector<float> tags;
    pts.reserve((size_t)V);
    uvs.reserve((size_t)U);
    fv.reserve((size_t)H);
    fuv.reserve((size_t)H);
    rewind(f);
    while (fgets(buf, sizeof(buf), f)) {
        if (buf[0] == 'v') {
            if (buf[1] == ' ') {
                cc_VertexPoint p;
                if (sscanf(&buf[2], "%f %f %f", &p.x, &p.y, &p.z) == 3) pts.push_back(p);
            } else if (buf[1] == 't') {
                cc_VertexUv t;
                if (sscanf(&buf[2], "%f %f", &t.u, &t.v) == 2) uvs.push_back(t);
            }
        } else if (buf[0] == 'f') {
            obj_read_face(&buf[2], &fv, &fuv, &bad);
            off.push_back((int32_t)fv.size());
        }
    }
    // crease tags, in file order
    rewind(f);
    while (fgets(buf, sizeof(buf), f)) {
        int32_t v0, v1;
        float s;
        if (sscanf(buf, "t crease 2/1/0 %i %i %f", &v0, &v1, &s) == 3) {
            tagv.push_back(v0);
            tagv.push_back(v1);
            tags.push_back(s);
        }
    }
    fclose(f);
    if (bad) return fail(5, "ccm_LoadObj: unsupported face (relative index or fewer than 3 vertices)");
    if ((int32_t)fv.size() != H || (int32_t)pts.size() != V || (int32_t)uvs.size() != U)
        return fail(5, "ccm_LoadObj: failed to read OBJ data");
    return ccm_CreateFromFaces(V, pts.data(), (int32_t)off.size() - 1, off.data(), fv.data(), U, uvs.data(), fuv.data(),
                               (int32_t)tags.size(), tagv.data(), tags.data(), 1);
}
