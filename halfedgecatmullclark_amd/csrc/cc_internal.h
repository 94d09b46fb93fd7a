// Library-internal helpers shared by the host and runtime translation units.
#pragma once

#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/CatmullClarkHIP.h"
#include "cc_common.h"

namespace cch {

// thread-local error channel behind ccs_LastError / ccs_LastErrorString
void set_error(int32_t code, const char *msg);
void clear_error();
int32_t last_error();
// reports a HIP failure on the error channel and stderr; true on error
bool hip_failed(hipError_t e, const char *what);

// makes `dev` current for the scope
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// true when every cumulative count the reference evaluates in int32 fits
bool ref_counts_fit_int32(const cc_Mesh *cage, int32_t D, const char **why);

// The cage's vertex and face rings as the device walks them (kWalkCap
// steps at most), in CSR form.  `ok` is false when the lists would exceed
// 4 H + V entries (a cage whose walks do not close): the kernels then walk.
struct CageRings {
    std::vector<int32_t> voff, vring, vbnd, foff, fring, ooff, out;
    bool ok = false;
};

inline CageRings build_cage_rings(const cc_Mesh *m) {
    CageRings r;
    const int32_t V = m->vertexCount, F = m->faceCount, H = m->halfedgeCount;
    const CageHe *he = reinterpret_cast<const CageHe *>(m->halfedges);
    const size_t cap = 4 * (size_t)H + (size_t)V;
    r.voff.assign((size_t)V + 1, 0);
    r.vbnd.assign((size_t)V, 0);
    r.foff.assign((size_t)F + 1, 0);
    auto valid = [&](int32_t h) { return h >= 0 && h < H; };
    for (int32_t v = 0; v < V; ++v) {
        r.voff[v] = (int32_t)r.vring.size();
        const int32_t p = m->vertexToHalfedgeIDs[v];
        if (!valid(p)) continue;
        r.vring.push_back(p);
        int32_t n = 1;
        int32_t g = valid(he[p].prev) ? he[he[p].prev].twin : -1;
        for (; g >= 0 && g != p && n < kWalkCap; ++n) {
            if (!valid(g) || r.vring.size() >= cap) return r;
            r.vring.push_back(g);
            g = valid(he[g].prev) ? he[he[g].prev].twin : -1;
        }
        r.vbnd[v] = g < 0 ? 1 : 0;
    }
    r.voff[V] = (int32_t)r.vring.size();
    for (int32_t f = 0; f < F; ++f) {
        r.foff[f] = (int32_t)r.fring.size();
        const int32_t p = m->faceToHalfedgeIDs[f];
        if (!valid(p)) continue;
        r.fring.push_back(p);
        int32_t n = 1;
        for (int32_t it = he[p].next; it != p && n < kWalkCap; it = he[it].next, ++n) {
            if (!valid(it) || r.fring.size() >= cap) return r;
            r.fring.push_back(it);
        }
    }
    r.foff[F] = (int32_t)r.fring.size();
    // halfedges by vertex, ascending (a counting sort)
    r.ooff.assign((size_t)V + 1, 0);
    for (int32_t h = 0; h < H; ++h) {
        const int32_t v = he[h].vert;
        if (v < 0 || v >= V) return r;
        ++r.ooff[(size_t)v + 1];
    }
    for (int32_t v = 0; v < V; ++v) r.ooff[(size_t)v + 1] += r.ooff[v];
    r.out.assign((size_t)H, 0);
    std::vector<int32_t> fill(r.ooff.begin(), r.ooff.end() - 1);
    for (int32_t h = 0; h < H; ++h) r.out[(size_t)fill[he[h].vert]++] = h;
    r.ok = true;
    return r;
}

// True when every vertex of every level of this cage takes a structural ring
// in the vertex passes (marked_ring_fan + structural_ring), so nothing is
// left for the rare-vertex kernels: no boundary, every cage vertex with a
// closed ring of four whose halfedges R have a twin with next(twin(R)) != R,
// and every cage face a quad whose halfedges are f2h .. f2h+3 in next order.
inline bool rings_without_rare(const cc_Mesh *m, const CageRings &r) {
    const int32_t V = m->vertexCount, F = m->faceCount, H = m->halfedgeCount;
    const CageHe *he = reinterpret_cast<const CageHe *>(m->halfedges);
    for (int32_t h = 0; h < H; ++h)
        if (he[h].twin < 0 || he[h].twin >= H) return false;
    for (int32_t v = 0; v < V; ++v) {
        if (r.voff[v + 1] - r.voff[v] != 4 || r.vbnd[v] != 0) return false;
        for (int32_t i = r.voff[v]; i < r.voff[v + 1]; ++i) {
            const int32_t R = r.vring[i], w = he[R].twin;
            const int32_t nt = he[w].next;
            if (nt < 0 || nt >= H || nt == R) return false;
        }
    }
    for (int32_t f = 0; f < F; ++f) {
        const int32_t p = m->faceToHalfedgeIDs[f];
        if (p < 0 || p + 3 >= H) return false;
        if (he[p].next != p + 1 || he[p + 1].next != p + 2 || he[p + 2].next != p + 3 || he[p + 3].next != p)
            return false;
    }
    return true;
}

// host-memory view of a cage with the device-view type (for the O(depth) maps)
inline DCage host_cage_view(const cc_Mesh *m) {
    DCage c;
    c.vring_off = c.vring = c.vring_bnd = c.fring_off = c.fring = c.vout_off = c.vout = nullptr;
    c.he = reinterpret_cast<const CageHe *>(m->halfedges);
    c.v2h = m->vertexToHalfedgeIDs;
    c.e2h = m->edgeToHalfedgeIDs;
    c.f2h = m->faceToHalfedgeIDs;
    c.pts = reinterpret_cast<const P3 *>(m->vertexPoints);
    c.uvs = reinterpret_cast<const UV *>(m->uvs);
    c.cr = reinterpret_cast<const Crease *>(m->creases);
    c.V = m->vertexCount; c.U = m->uvCount; c.H = m->halfedgeCount; c.E = m->edgeCount; c.F = m->faceCount;
    return c;
}

inline LevelTab level_tab(const Counts &n) {
    LevelTab t;
    for (int d = 0; d <= kMaxDepth; ++d) {
        t.E[d] = (int32_t)n.E[d];
        t.V[d] = (int32_t)n.V[d];
        t.F[d] = (int32_t)n.F[d];
    }
    int64_t words = 0;   // levels past the addressable range are never used
    t.R[0] = 0;
    for (int d = 1; d <= kMaxDepth; ++d) {
        t.R[d] = (int32_t)std::min<int64_t>(words, INT32_MAX);
        words += (n.H[d] + 63) / 64;
    }
    return t;
}

}  // namespace cch
