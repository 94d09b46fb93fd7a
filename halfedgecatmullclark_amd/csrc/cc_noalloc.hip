// ccd_RefineNoAlloc: the refinement of ccd_Refine for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  The cage and the level arrays are the
// caller's device arrays; the kernels' scratch (the vertex-owner bits of a
// level, the list of vertices left to the rare-vertex kernel, the
// edge-representative bits of levels 1..D-2) lives in one file-scope device
// array per GPU, held until the process ends.  Nothing here calls hipMalloc,
// hipFree or a library routine that allocates.
//
// The cage's ring lists, which ccd_Create builds on the host, are not
// available: the kernels walk the cage instead (same halfedges, same order,
// so the results are the same bits).
#include <stdio.h>

#include <mutex>

#include "cc_internal.h"
#include "cc_runtime.h"

using namespace cch;

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;   // CCD_NOALLOC_SCRATCH_BYTES
constexpr int kMaxDevices = 64;

// The scratch array.  In host code its name is only a stand-in for the
// symbol: the device address of each GPU's copy comes from
// hipGetSymbolAddress with that device current, and only that address is
// ever handed to a kernel.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// the device address of this GPU's scratch array, looked up once per device
void *scratch_on(int dev) {
    if (dev < 0 || dev >= kMaxDevices) {
        set_error(16, "ccd_RefineNoAlloc: device index out of range");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (hip_failed(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "ccd_RefineNoAlloc: hipGetSymbolAddress(scratch)"))
            return nullptr;
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

size_t round256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

extern "C" int32_t ccd_RefineNoAlloc(const cc_Mesh *cage, const cc_Subd *subd, int32_t steps, void *stream) {
    clear_error();
    if (!cage || !subd) {
        set_error(11, "ccd_RefineNoAlloc: null argument");
        return last_error();
    }
    const int32_t D = subd->maxDepth;
    if (D < 1 || D > kMaxDepth || cage->halfedgeCount < 3 || cage->edgeCount < 1 || cage->faceCount < 1 ||
        cage->vertexCount < 1 || cage->uvCount < 0) {
        set_error(12, "ccd_RefineNoAlloc: invalid cage counts or depth");
        return last_error();
    }
    Counts n;
    n.init(cage->vertexCount, cage->halfedgeCount, cage->edgeCount, cage->faceCount);
    for (int d = 0; d <= D; ++d) {
        if (n.H[d] > INT32_MAX || n.V[d] + n.E[d] + n.F[d] > INT32_MAX || 2 * n.E[d] + n.H[d] > INT32_MAX) {
            set_error(12, "ccd_RefineNoAlloc: IDs at this depth overflow int32");
            return last_error();
        }
    }
    const LevelTab tab = level_tab(n);
    // scratch layout: [rare list | owner bits | edge-representative bits]
    const int64_t rare_cap = (int64_t)cage->vertexCount + cage->faceCount +
                             ((int64_t)cage->halfedgeCount << (D >= 1 ? D - 1 : 0));
    const size_t o_mask = round256(sizeof(int32_t) * (size_t)(rare_cap + 1));
    const size_t o_erep = o_mask + round256(sizeof(uint64_t) * (size_t)(D >= 2 ? (n.H[D - 1] + 63) / 64 : 1));
    const size_t need = o_erep + (D >= 4 ? sizeof(uint64_t) * (size_t)tab.R[D - 1] : 0);
    if (rare_cap >= INT32_MAX || need > kScratchBytes) {
        char buf[256];
        snprintf(buf, sizeof(buf),
                 "ccd_RefineNoAlloc: this cage at depth %d needs %zu bytes of scratch, more than the %zu the "
                 "entry point holds",
                 D, need, kScratchBytes);
        fprintf(stderr, "cc: %s\n", buf);
        set_error(15, buf);
        return last_error();
    }
    if (!cage->halfedges || !cage->vertexToHalfedgeIDs || !cage->edgeToHalfedgeIDs || !cage->faceToHalfedgeIDs ||
        !cage->vertexPoints || !cage->creases || (cage->uvCount > 0 && !cage->uvs) || !subd->halfedges ||
        !subd->creases || !subd->vertexPoints) {
        set_error(11, "ccd_RefineNoAlloc: null cage or level array");
        return last_error();
    }
    int dev = 0;
    if (hip_failed(hipGetDevice(&dev), "ccd_RefineNoAlloc: hipGetDevice")) return last_error();
    unsigned char *base = static_cast<unsigned char *>(scratch_on(dev));
    if (!base) return last_error();

    // a view of the caller's arrays with the library's launch sequence; it
    // owns nothing (free_all is never called on it)
    ccd_Subd s;
    s.device = dev;
    s.D = D;
    s.V0 = cage->vertexCount; s.U0 = cage->uvCount; s.H0 = cage->halfedgeCount;
    s.E0 = cage->edgeCount; s.F0 = cage->faceCount;
    s.n = n;
    s.tab = tab;
    s.c_he = reinterpret_cast<CageHe *>(cage->halfedges);
    s.c_v2h = cage->vertexToHalfedgeIDs;
    s.c_e2h = cage->edgeToHalfedgeIDs;
    s.c_f2h = cage->faceToHalfedgeIDs;
    s.c_pts = reinterpret_cast<P3 *>(cage->vertexPoints);
    s.c_uvs = reinterpret_cast<UV *>(cage->uvs);
    s.c_cr = reinterpret_cast<Crease *>(cage->creases);
    s.rings = nullptr;
    s.rings_ok = false;
    s.he = reinterpret_cast<He *>(subd->halfedges);
    s.cr = reinterpret_cast<Crease *>(subd->creases);
    s.pts = reinterpret_cast<P3 *>(subd->vertexPoints);
    s.rare = reinterpret_cast<int32_t *>(base);
    s.rare_cap = (int32_t)rare_cap;
    s.defer_mask = reinterpret_cast<uint64_t *>(base + o_mask);
    s.erep = D >= 4 ? reinterpret_cast<uint64_t *>(base + o_erep) : nullptr;
    if (!s.refine(steps, D, (hipStream_t)stream)) return last_error() ? last_error() : 1;
    return 0;
}
