// Library-internal helpers shared by the host and runtime translation units.
#pragma once

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

// host-memory view of a cage with the device-view type (for the O(depth) maps)
inline DCage host_cage_view(const cc_Mesh *m) {
    DCage c;
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
    return t;
}

}  // namespace cch
