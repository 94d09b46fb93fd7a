// Runtime of the C ABI: device mirrors of cc_Subd (ccd_Subd), the per-level
// launch sequence, the drop-in ccs_Refine* entry points with the reference's
// host semantics (CatmullClark.h:2631-2683, :2803-2812, :2928-2939,
// :3033-3042, :3051-3082), the device-resident ccd_* API and the error
// channel.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "cc_internal.h"
#include "cc_kernels.h"
#include "cc_runtime.h"

using namespace cch;

// ---- error channel -----------------------------------------------------------
namespace {
thread_local int32_t t_err = 0;
thread_local char t_msg[512] = "";
}  // namespace

namespace cch {
void set_error(int32_t code, const char *msg) {
    t_err = code;
    snprintf(t_msg, sizeof(t_msg), "%s", msg);
}
void clear_error() {
    t_err = 0;
    t_msg[0] = '\0';
}
}  // namespace cch

extern "C" int32_t ccs_LastError(void) { return t_err; }
extern "C" const char *ccs_LastErrorString(void) { return t_msg; }

namespace cch {
// reports a HIP failure on the error channel and stderr; returns true on error
bool hip_failed(hipError_t e, const char *what) {
    if (e == hipSuccess) return false;
    char buf[400];
    snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    fprintf(stderr, "cc: %s\n", buf);
    set_error(100 + (int32_t)e, buf);
    return true;
}
int32_t last_error() { return t_err; }
}  // namespace cch

namespace {

#define CC_TRY(expr, what) \
    do { if (hip_failed((expr), (what))) return false; } while (0)

constexpr int kBlock = 256;

// What a refinement does with the children's uvIDs: 0 leave them, 1 refine
// them, 2 write 0 (a cage without uvs, when the topology is written).  A
// library built without uvs has no uvID field: always 0.
int uv_mode(int32_t steps, int32_t U0) {
    if (!kHasUv) return 0;
    int m = ((steps & CCD_UVS) && U0 > 0) ? 1 : 0;
    if ((steps & CCD_HALFEDGES) && m == 0 && U0 == 0) m = 2;
    return m;
}

}  // namespace

// ---- device mirror -----------------------------------------------------------
bool ccd_Subd::alloc(const cc_Mesh *cage, int32_t depth) {
    D = depth;
    V0 = cage->vertexCount; U0 = cage->uvCount; H0 = cage->halfedgeCount; E0 = cage->edgeCount; F0 = cage->faceCount;
    n.init(V0, H0, E0, F0);
    tab = level_tab(n);
    CC_TRY(hipGetDevice(&device), "hipGetDevice");
    CC_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    CC_TRY(hipMalloc(&c_he, sizeof(CageHe) * (size_t)H0), "hipMalloc(cage halfedges)");
    CC_TRY(hipMalloc(&c_v2h, sizeof(int32_t) * (size_t)(V0 + 1)), "hipMalloc(cage v2h)");
    CC_TRY(hipMalloc(&c_e2h, sizeof(int32_t) * (size_t)(E0 + 1)), "hipMalloc(cage e2h)");
    CC_TRY(hipMalloc(&c_f2h, sizeof(int32_t) * (size_t)(F0 + 1)), "hipMalloc(cage f2h)");
    CC_TRY(hipMalloc(&c_pts, sizeof(P3) * (size_t)(V0 + 1)), "hipMalloc(cage points)");
    CC_TRY(hipMalloc(&c_uvs, sizeof(UV) * (size_t)(U0 + 1)), "hipMalloc(cage uvs)");
    CC_TRY(hipMalloc(&c_cr, sizeof(Crease) * (size_t)(E0 + 1)), "hipMalloc(cage creases)");
    CC_TRY(hipMalloc(&he, sizeof(He) * (size_t)n.cumH(D)), "hipMalloc(halfedges)");
    CC_TRY(hipMalloc(&cr, sizeof(Crease) * (size_t)n.cumC(D)), "hipMalloc(creases)");
    CC_TRY(hipMalloc(&pts, sizeof(P3) * (size_t)n.cumV(D)), "hipMalloc(points)");
    CC_TRY(hipMalloc(&defer_mask, sizeof(uint64_t) * (size_t)(D >= 2 ? (n.H[D - 1] + 63) / 64 : 1)),
           "hipMalloc(deferred-vertex masks)");
    if (D >= 4) CC_TRY(hipMalloc(&erep, sizeof(uint64_t) * (size_t)tab.R[D - 1]), "hipMalloc(edge representative bits)");
    int64_t words = 0;
    for (int d = 1; d < D; ++d) {
        own_off[d] = words;
        words += (n.H[d] + 63) / 64;
    }
    CC_TRY(hipMalloc(&own_masks, sizeof(uint64_t) * (size_t)(words > 0 ? words : 1)), "hipMalloc(owner bits)");
    return true;
}

void ccd_Subd::free_all() {
    DeviceGuard g(device);
    void *ptrs[] = {c_he, c_v2h, c_e2h, c_f2h, c_pts, c_uvs, c_cr, he, cr, pts, adj_a, adj_b, adj_w, adj_in, adj_out, adj_sharp,
                    adj_vp, defer_mask, own_masks, rings, rare, erep, nrm_tmp, nrm_out};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &t : timed) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    timed.clear();
    drop_graphs();
    if (stream) (void)hipStreamDestroy(stream);
}

void ccd_Subd::drop_graphs() {
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
}

bool ccd_Subd::upload_cage(const cc_Mesh *cage, hipStream_t st) {
    if (cage->vertexCount != V0 || cage->uvCount != U0 || cage->halfedgeCount != H0 ||
        cage->edgeCount != E0 || cage->faceCount != F0) {
        set_error(10, "cage counts differ from the ones the device mirror was created for");
        return false;
    }
    for (bool &v : erep_valid) v = false;   // bits of the previous topology
    for (bool &v : own_valid) v = false;
    adj_vp_key = -1;
    drop_graphs();
    CC_TRY(hipMemcpyAsync(c_he, cage->halfedges, sizeof(CageHe) * (size_t)H0, hipMemcpyHostToDevice, st), "upload cage");
    CC_TRY(hipMemcpyAsync(c_v2h, cage->vertexToHalfedgeIDs, sizeof(int32_t) * (size_t)V0, hipMemcpyHostToDevice, st), "upload cage");
    CC_TRY(hipMemcpyAsync(c_e2h, cage->edgeToHalfedgeIDs, sizeof(int32_t) * (size_t)E0, hipMemcpyHostToDevice, st), "upload cage");
    CC_TRY(hipMemcpyAsync(c_f2h, cage->faceToHalfedgeIDs, sizeof(int32_t) * (size_t)F0, hipMemcpyHostToDevice, st), "upload cage");
    CC_TRY(hipMemcpyAsync(c_pts, cage->vertexPoints, sizeof(P3) * (size_t)V0, hipMemcpyHostToDevice, st), "upload cage");
    if (U0 > 0) CC_TRY(hipMemcpyAsync(c_uvs, cage->uvs, sizeof(UV) * (size_t)U0, hipMemcpyHostToDevice, st), "upload cage");
    CC_TRY(hipMemcpyAsync(c_cr, cage->creases, sizeof(Crease) * (size_t)E0, hipMemcpyHostToDevice, st), "upload cage");
    {   // the rare-vertex list: cage vertices, cage faces and boundary vertices
        int64_t b0 = 0;
        const CageHe *ch = reinterpret_cast<const CageHe *>(cage->halfedges);
        for (int32_t h = 0; h < H0; ++h) b0 += ch[h].twin < 0 ? 1 : 0;
        const int64_t cap = std::min<int64_t>((int64_t)V0 + F0 + (b0 << (D >= 1 ? D - 1 : 0)), INT32_MAX - 1);
        if (cap > rare_cap || !rare) {
            if (rare) (void)hipFree(rare);
            rare = nullptr;
            rare_cap = 0;
            CC_TRY(hipMalloc(&rare, sizeof(int32_t) * (size_t)(cap + 1)), "hipMalloc(rare vertices)");
            rare_cap = (int32_t)cap;
        }
    }
    // ring lists: [vertex offsets | vertex boundary flags | face offsets |
    // vertex rings | face rings], each list padded by one entry
    const CageRings r = build_cage_rings(cage);
    rings_ok = r.ok;
    rare_free = r.ok && rings_without_rare(cage, r);
    if (!r.ok) return true;
    std::vector<int32_t> rings_host;
    auto put = [&](const std::vector<int32_t> &v) {
        const size_t o = rings_host.size();
        rings_host.insert(rings_host.end(), v.begin(), v.end());
        rings_host.push_back(0);
        return o;
    };
    o_vro = put(r.voff);
    o_vb = put(r.vbnd);
    o_fro = put(r.foff);
    o_vr = put(r.vring);
    o_fr = put(r.fring);
    o_voo = put(r.ooff);
    o_vo = put(r.out);
    if (rings_host.size() > rings_cap) {
        if (rings) (void)hipFree(rings);
        rings = nullptr;
        rings_cap = 0;
        CC_TRY(hipMalloc(&rings, sizeof(int32_t) * rings_host.size()), "hipMalloc(cage rings)");
        rings_cap = rings_host.size();
    }
    CC_TRY(hipMemcpyAsync(rings, rings_host.data(), sizeof(int32_t) * rings_host.size(), hipMemcpyHostToDevice, st),
           "upload cage rings");
    CC_TRY(hipStreamSynchronize(st), "upload cage rings");   // rings_host is freed on return
    return true;
}

DCage ccd_Subd::cage_view(const P3 *points) const {
    DCage c;
    c.he = c_he; c.v2h = c_v2h; c.e2h = c_e2h; c.f2h = c_f2h;
    c.pts = points ? points : c_pts;
    c.uvs = c_uvs; c.cr = c_cr;
    if (rings_ok) {
        c.vring_off = rings + o_vro; c.vring = rings + o_vr; c.vring_bnd = rings + o_vb;
        c.fring_off = rings + o_fro; c.fring = rings + o_fr;
        c.vout_off = rings + o_voo; c.vout = rings + o_vo;
    } else {
        c.vring_off = c.vring = c.vring_bnd = c.fring_off = c.fring = c.vout_off = c.vout = nullptr;
    }
    c.V = V0; c.U = U0; c.H = H0; c.E = E0; c.F = F0;
    return c;
}

Lvl ccd_Subd::level_view(int d, const P3 *levelPts, P3 *nextPts) const {
    Lvl L;
    L.he = he + n.offH[d];
    L.cr = cr + n.offC[d];
    L.old = levelPts;
    L.vp = nextPts;
    L.fp = nextPts + n.V[d];
    L.ep = L.fp + n.F[d];
    L.H = (int32_t)n.H[d]; L.V = (int32_t)n.V[d]; L.F = (int32_t)n.F[d]; L.E = (int32_t)n.E[d]; L.C = (int32_t)n.C[d];
    return L;
}

void ccd_Subd::begin_timing() {
    for (auto &t : timed) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    timed.clear();
}

template <typename K, typename... Args>
void ccd_Subd::launch(const char *name, K kernel, int64_t threads, hipStream_t st, Args... args) {
    if (threads <= 0) return;
    const unsigned grid = (unsigned)((threads + kBlock - 1) / kBlock);
    Timed t{name, nullptr, nullptr};
    if (timing) {
        (void)hipEventCreate(&t.start);
        (void)hipEventCreate(&t.stop);
        (void)hipEventRecord(t.start, st);
    }
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, st, args...);
    if (timing) {
        (void)hipEventRecord(t.stop, st);
        timed.push_back(t);
    }
}

template <typename Fn>
void ccd_Subd::launch_timed(const char *name, int64_t threads, hipStream_t st, Fn &&fn) {
    if (threads <= 0) return;
    Timed t{name, nullptr, nullptr};
    if (timing) {
        (void)hipEventCreate(&t.start);
        (void)hipEventCreate(&t.stop);
        (void)hipEventRecord(t.start, st);
    }
    fn();
    if (timing) {
        (void)hipEventRecord(t.stop, st);
        timed.push_back(t);
    }
}

void ccd_Subd::zero(float *p, int64_t count, hipStream_t st) {
    if (count <= 0) return;
    const int64_t blocks = (count + kBlock - 1) / kBlock;
    const unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
    hipLaunchKernelGGL(k_zero, dim3(grid), dim3(kBlock), 0, st, p, count);
}

FusedArgs ccd_Subd::fused_args(int d, const P3 *points) const {
    FusedArgs a;
    a.he = he + n.offH[d];
    a.out_he = he + n.offH[d + 1];
    a.cr = cr + n.offC[d];
    a.old = points + n.offV[d];
    a.vp = const_cast<P3 *>(points) + n.offV[d + 1];
    a.fp = a.vp + n.V[d];
    a.ep = a.fp + n.F[d];
    a.H = (int32_t)n.H[d]; a.V = (int32_t)n.V[d]; a.F = (int32_t)n.F[d]; a.E = (int32_t)n.E[d]; a.C = (int32_t)n.C[d];
    a.xcd_remap = opt_remap;
    a.uv_refine = 0;
    a.defer_mask = nullptr;
    a.owner_mask = nullptr;
    a.rare = nullptr;
    a.rare_cap = 0;
    a.erep = nullptr;
    a.erep_read = a.erep_write = 0;
    return a;
}

LevelPtrs ccd_Subd::level_ptrs() const {
    LevelPtrs lp;
    for (int d = 0; d <= kMaxDepth; ++d) lp.he[d] = (d >= 1 && d <= D) ? he + n.offH[d] : nullptr;
    return lp;
}

// Topology, creases and uvs of levels 1..dmax with one kernel per output
// kind (CatmullClark.h:2803-2812, :2928-2939, :3033-3042).  uvMode: 0 leave
// uvIDs, 1 refine, 2 write zero.
bool ccd_Subd::refine_topology(int32_t steps, int dmax, hipStream_t st) {
    const bool topo = (steps & CCD_HALFEDGES) != 0;
    const int uvMode = uv_mode(steps, U0);
    if (topo || uvMode == 1) {
        const DCage c = cage_view(nullptr);
        launch("cage_halfedges", k_cage_halfedges, H0, st, c, he, (int)topo, uvMode);
        for (int d = 1; d < dmax; ++d) {
            launch("level_halfedges", k_level_halfedges, n.H[d], st, (const He *)(he + n.offH[d]), he + n.offH[d + 1],
                   (int32_t)n.H[d], (int32_t)n.V[d], (int32_t)n.F[d], (int32_t)n.E[d], (int)topo, uvMode);
        }
    }
    if (steps & CCD_CREASES) {
        launch("cage_creases", k_creases, E0, st, (const Crease *)c_cr, E0, cr);
        for (int d = 1; d < dmax; ++d)
            launch("level_creases", k_creases, n.C[d], st, (const Crease *)(cr + n.offC[d]), (int32_t)n.C[d], cr + n.offC[d + 1]);
    }
    return true;
}

// one level of the patch-local kernels: children, face, edge and most vertex
// points, then the vertices whose rings leave their block
void ccd_Subd::level_patch(int d, int32_t steps, int topo, int uvMode, P3 *points, hipStream_t st) {
    const bool gather = (steps & CCD_GATHER) != 0;
    const int creased = (steps & CCD_NOCREASES) ? 0 : 1;
    FusedArgs a = fused_args(d, points);
    a.defer_mask = defer_mask;
    const DCage c = cage_view(nullptr);
    const LevelPtrs lp = level_ptrs();
    launch_timed("level_patch", n.H[d], st,
                 [&] { launch_level_patch(a, c, tab, lp, d, gather, creased, topo, uvMode, cfg[d].nt, cfg[d].hpt, st); });
    launch_timed("level_deferred", n.V[d], st, [&] { launch_level_deferred(a, c, tab, lp, d, gather, creased, st); });
}

// one level of the fused kernels: children with face and edge points, then
// vertex points
void ccd_Subd::level_fused(int d, int32_t steps, int topo, int uvMode, P3 *points, hipStream_t st) {
    const bool gather = (steps & CCD_GATHER) != 0;
    const int creased = (steps & CCD_NOCREASES) ? 0 : 1;
    const bool pts_ = (steps & CCD_POINTS) != 0;
    FusedArgs a = fused_args(d, points);
    const DCage c = cage_view(nullptr);
    // the fused kernel marks the vertex representatives for the masked
    // kernel (the bits depend on the topology only: a points-only pass keeps
    // the bits of an earlier pass over the same cage)
    if (pts_ && opt_fan == 2 && cfg[d].scan == 0) {
        a.owner_mask = own_masks + own_off[d];
        a.owner_write = own_valid[d] && !topo ? 0 : 1;
        a.rare = rare;
        a.rare_cap = rare_free ? 0 : rare_cap;   // 0: the list stays empty, no rare kernel
    }
    // edge-representative bits: derive level d+1's while they are still
    // needed (levels up to D-2), read levels 1..d-1 in the owner test
    if (erep) {
        bool below = true;
        for (int l = 1; l < d; ++l) below = below && erep_valid[l];
        a.erep = erep;
        a.erep_read = d >= 2 && below ? 1 : 0;
        a.erep_write = d + 1 <= D - 2 && (d == 1 || erep_valid[d]) ? 1 : 0;
    }
    if (topo || uvMode == 1 || pts_) {
        launch_timed(pts_ ? (topo ? "level_fused" : "level_points") : "level_fused", n.H[d], st,
                     [&] {
                         launch_level_fused(a, c, tab, d, gather, creased, topo, uvMode, pts_, opt_store, cfg[d].fb,
                                            cfg[d].ft, st);
                     });
        if (a.erep_write) {
            erep_valid[d + 1] = true;
            if (d == 1) erep_valid[1] = true;
        }
        if (a.owner_mask) own_valid[d] = true;
    }
    if (pts_) {
        launch_timed("level_vertices", n.H[d], st, [&] {
            launch_level_vertices(a, c, tab, level_ptrs(), d, gather, creased, opt_fan, cfg[d].scan, cfg[d].vw, st);
        });
    }
}

// the cage step of the points (CatmullClark.h:2633-2636, :2661-2663)
void ccd_Subd::cage_points(int32_t steps, const P3 *cagePts, P3 *out, hipStream_t st) {
    const bool gather = (steps & CCD_GATHER) != 0;
    const int creased = (steps & CCD_NOCREASES) ? 0 : 1;
    const DCage c = cage_view(cagePts);
    P3 *vp = out, *fp = out + V0, *ep = out + V0 + F0;
    launch("cage_faces", k_cage_faces, F0, st, c, fp, (int)gather);
    launch("cage_edges", k_cage_edges, E0, st, c, (const P3 *)fp, ep, (int)gather, creased);
    if (gather) {
        launch("cage_vertices", k_cage_vertices_gather, V0, st, c, (const P3 *)fp, (const P3 *)ep, vp, creased);
    } else {
        launch("cage_vertices", k_cage_vertices_scatter, V0, st, c, (const P3 *)fp, (const P3 *)ep, vp, creased);
    }
}

// Points of levels 1..dmax (CatmullClark.h:2631-2683): the cage step, then
// face, edge and vertex points per level.
bool ccd_Subd::refine_points(int32_t steps, int dmax, const P3 *cagePts, P3 *out, hipStream_t st) {
    const bool gather = (steps & CCD_GATHER) != 0;
    const int creased = (steps & CCD_NOCREASES) ? 0 : 1;
    const DCage c = cage_view(cagePts);
    cage_points(steps, cagePts, out, st);
    for (int d = 1; d < dmax; ++d) {
        if (opt_fused) {
            const int32_t s = CCD_POINTS | (steps & (CCD_GATHER | CCD_NOCREASES));
            if (cfg[d].kind == 2) level_patch(d, s, 0, 0, out, st);
            else level_fused(d, s, 0, 0, out, st);
            continue;
        }
        const Lvl L = level_view(d, out + n.offV[d], out + n.offV[d + 1]);
        launch("faces", k_faces, n.F[d], st, L, (int)gather);
        if (gather) {
            launch("edges", k_edges_gather, n.E[d], st, L, c, tab, d, creased);
            launch("vertices", k_vertices_gather, n.V[d], st, L, c, tab, d, creased);
        } else {
            launch("edges", k_edges_scatter, n.H[d], st, L, creased);
            zero(&L.vp->x, 3 * n.V[d], st);
            launch("vertices", k_vertices_scatter, n.H[d], st, L, creased);
        }
    }
    return true;
}

bool ccd_Subd::refine(int32_t steps, int dmax, hipStream_t st) {
    if (timing) begin_timing();
    if (steps & (CCD_HALFEDGES | CCD_CREASES)) adj_vp_key = -1;   // the adjoint's parameters follow these
    if (!opt_fused) {
        if (!refine_topology(steps, dmax, st)) return false;
        if ((steps & CCD_POINTS) && !refine_points(steps, dmax, nullptr, pts, st)) return false;
        CC_TRY(hipGetLastError(), "kernel launch");
        return true;
    }
    // fused: creases of every level first (the point rules read them), then
    // per level one kernel for children + face/edge points and one for
    // vertex points
    const bool topo = (steps & CCD_HALFEDGES) != 0;
    const bool points = (steps & CCD_POINTS) != 0;
    const int uvMode = uv_mode(steps, U0);
    if (steps & CCD_CREASES) {
        launch("cage_creases", k_creases, E0, st, (const Crease *)c_cr, E0, cr);
        for (int d = 1; d < dmax; ++d)
            launch("level_creases", k_creases, n.C[d], st, (const Crease *)(cr + n.offC[d]), (int32_t)n.C[d], cr + n.offC[d + 1]);
    }
    const DCage c = cage_view(nullptr);
    if (topo || uvMode == 1) launch("cage_halfedges", k_cage_halfedges, H0, st, c, he, (int)topo, uvMode);
    if (points) cage_points(steps, nullptr, pts, st);
    for (int d = 1; d < dmax; ++d) {
        if (points && cfg[d].kind == 2) level_patch(d, steps, topo, uvMode, pts, st);
        else level_fused(d, steps, topo, uvMode, pts, st);
    }
    CC_TRY(hipGetLastError(), "kernel launch");
    return true;
}

// refine(steps, D) as a graph captured on the subd's own stream on first
// use and launched on st.  The host-side choices a capture bakes in (the
// level kernels, options, which levels' edge-representative bits to read or
// write) are part of the key or drop the graphs when they change.
bool ccd_Subd::refine_graph(int32_t steps, hipStream_t st) {
    int64_t key = (uint32_t)steps;
    for (int l = 1; l <= kMaxDepth; ++l)
        if (erep_valid[l]) key |= (int64_t)1 << (32 + l);
    for (int l = 1; l <= kMaxDepth; ++l)
        if (own_valid[l]) key |= (int64_t)1 << (48 + l);
    auto it = graphs.find(key);
    if (it == graphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t exec = nullptr;
        CC_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        const bool ok = refine(steps, D, stream);
        const hipError_t e = hipStreamEndCapture(stream, &g);
        if (!ok) {
            if (g) (void)hipGraphDestroy(g);
            return false;
        }
        CC_TRY(e, "hipStreamEndCapture");
        const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        CC_TRY(ei, "hipGraphInstantiate");
        it = graphs.emplace(key, exec).first;
    }
    CC_TRY(hipGraphLaunch(it->second, st), "hipGraphLaunch");
    return true;
}

// Adjoint of the point refinement for the current topology and creases.
bool ccd_Subd::adjoint(const void *gradFinest, void *gradCage, int32_t steps, hipStream_t st) {
    const size_t scalar = (steps & CCD_FP64) ? sizeof(double) : sizeof(float);
    if (adj_scalar != scalar) {
        if (adj_a) (void)hipFree(adj_a);
        if (adj_b) (void)hipFree(adj_b);
        if (adj_w) (void)hipFree(adj_w);
        adj_a = adj_b = adj_w = nullptr;
        adj_scalar = 0;
        // level gradients: D-1 in adj_b, then D-2, D-3, ... alternating with
        // adj_a (a cage refined once copies the finest gradient to adj_a)
        CC_TRY(hipMalloc(&adj_a, 3 * scalar * (size_t)(D >= 2 ? n.V[D - 2] : n.V[D])), "hipMalloc(adjoint scratch)");
        CC_TRY(hipMalloc(&adj_b, 3 * scalar * (size_t)(D >= 1 ? n.V[D - 1] : 1)), "hipMalloc(adjoint scratch)");
        CC_TRY(hipMalloc(&adj_w, scalar * (size_t)adj_scratch_scalars(n, D, true)), "hipMalloc(adjoint scratch)");
        adj_scalar = scalar;
    }
    AdjPlan pl;
    pl.n = &n;
    pl.D = D;
    pl.he = he;
    pl.cr = cr;
    pl.cage = cage_view(nullptr);
    pl.tab = tab;
    pl.lp = level_ptrs();
    pl.per_vertex_max = opt_adj_vertex;
    pl.rare = rare;   // the forward's rare-vertex list, free between passes
    pl.rare_cap = rare_free ? 0 : rare_cap;
    bool kept = erep != nullptr;
    for (int l = 1; l <= D - 2; ++l) kept = kept && erep_valid[l];
    pl.erep = kept ? erep : nullptr;
    const int creased = (steps & CCD_NOCREASES) ? 0 : 1;
    if (scalar == sizeof(double)) {
        // the double adjoint refines sharpness in double too, as the double
        // build of the reference does
        if (!adj_sharp)
            CC_TRY(hipMalloc(&adj_sharp, sizeof(double) * (size_t)(E0 + n.cumC(D))), "hipMalloc(adjoint sharpness)");
        adjoint_levels<double>(pl, (const double *)gradFinest, (double *)gradCage, (double *)adj_a, (double *)adj_b,
                               (double *)adj_w, (double *)adj_sharp, creased, st);
    } else {
        if (!adj_vp && D >= 2)
            CC_TRY(hipMalloc(&adj_vp, sizeof(VParam<float>) * (size_t)n.cumV(D - 1)), "hipMalloc(adjoint parameters)");
        pl.vp_cache = adj_vp;
        pl.vp_valid = adj_vp_key == creased;
        adjoint_levels<float>(pl, (const float *)gradFinest, (float *)gradCage, (float *)adj_a, (float *)adj_b,
                              (float *)adj_w, nullptr, creased, st);
        adj_vp_key = creased;
    }
    CC_TRY(hipGetLastError(), "adjoint launch");
    return true;
}

// ---- drop-in host API ----------------------------------------------------------
namespace {

std::mutex g_mu;
std::unordered_map<const cc_Subd *, ccd_Subd *> g_mirrors;

// The level arrays ccs_Create allocates are page-locked when the runtime
// allows it, so that every refinement's copy back to them runs at the DMA
// rate instead of through a pageable staging copy; otherwise, and for
// arrays the caller allocated, plain malloc memory.
std::mutex g_pin_mu;
std::unordered_map<void *, bool> g_pinned;

void *host_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess && p) {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        g_pinned[p] = true;
        return p;
    }
    (void)hipGetLastError();   // no device or no pinned memory: pageable instead
    return malloc(bytes);
}

void host_free(void *p) {
    if (!p) return;
    bool pinned = false;
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        pinned = g_pinned.erase(p) > 0;
    }
    if (pinned) (void)hipHostFree(p);
    else free(p);
}

ccd_Subd *create_mirror(const cc_Mesh *cage, int32_t depth) {
    ccd_Subd *s = new (std::nothrow) ccd_Subd();
    if (!s) { set_error(2, "out of host memory"); return nullptr; }
    if (!s->alloc(cage, depth)) {
        s->free_all();
        delete s;
        return nullptr;
    }
    return s;
}

ccd_Subd *mirror_for(const cc_Subd *subd) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_mirrors.find(subd);
    const cc_Mesh *c = subd->cage;
    if (it != g_mirrors.end()) {
        ccd_Subd *s = it->second;
        if (s->D == subd->maxDepth && s->V0 == c->vertexCount && s->U0 == c->uvCount && s->H0 == c->halfedgeCount &&
            s->E0 == c->edgeCount && s->F0 == c->faceCount)
            return s;
        s->free_all();
        delete s;
        g_mirrors.erase(it);
    }
    ccd_Subd *s = create_mirror(c, subd->maxDepth);
    if (s) g_mirrors[subd] = s;
    return s;
}

void drop_mirror(const cc_Subd *subd) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_mirrors.find(subd);
    if (it != g_mirrors.end()) {
        it->second->free_all();
        delete it->second;
        g_mirrors.erase(it);
    }
}

// Run `steps` for levels 1..dmax of a host cc_Subd on the GPU.  The data a
// step reads from the subd (the topology and creases that a points-only call
// consumes, the other field of a partially rewritten halfedge record) is
// uploaded first, so the result equals what the reference computes from the
// current host arrays.
bool host_refine(cc_Subd *subd, int32_t steps, int dmax, const char *fname) {
    clear_error();
    if (!subd || !subd->cage) {
        set_error(11, "null subd or cage");
        fprintf(stderr, "cc: %s: null subd or cage\n", fname);
        return false;
    }
    const char *why = nullptr;
    if (!ref_counts_fit_int32(subd->cage, subd->maxDepth, &why)) {
        char buf[256];
        snprintf(buf, sizeof(buf), "%s: %s", fname, why);
        fprintf(stderr, "cc: %s\n", buf);
        set_error(12, buf);
        return false;
    }
    ccd_Subd *s = mirror_for(subd);
    if (!s) return false;
    DeviceGuard g(s->device);
    hipStream_t st = s->stream;
    const cc_Mesh *cage = subd->cage;
    if (!s->upload_cage(cage, st)) return false;
    const Counts &n = s->n;
    const bool uvActive = kHasUv && (steps & CCD_UVS) && s->U0 > 0;
    const bool writesHe = (steps & CCD_HALFEDGES) || uvActive;
    const bool partialHe = (steps & CCD_HALFEDGES) ? (kHasUv && !uvActive && s->U0 > 0) : uvActive;
    const bool pointsNeedTopo = (steps & CCD_POINTS) && !(steps & CCD_HALFEDGES);
    if (partialHe || pointsNeedTopo) {
        const int64_t count = partialHe ? n.cumH(dmax) : n.cumH(dmax - 1);
        CC_TRY(hipMemcpyAsync(s->he, subd->halfedges, sizeof(He) * (size_t)count, hipMemcpyHostToDevice, st), "upload halfedges");
    }
    if ((steps & CCD_POINTS) && !(steps & CCD_CREASES) && dmax > 1) {
        CC_TRY(hipMemcpyAsync(s->cr, subd->creases, sizeof(Crease) * (size_t)n.cumC(dmax - 1), hipMemcpyHostToDevice, st), "upload creases");
    }
    if (!s->refine(steps, dmax, st)) return false;
    if (writesHe)
        CC_TRY(hipMemcpyAsync(subd->halfedges, s->he, sizeof(He) * (size_t)n.cumH(dmax), hipMemcpyDeviceToHost, st), "download halfedges");
    if (steps & CCD_CREASES)
        CC_TRY(hipMemcpyAsync(subd->creases, s->cr, sizeof(Crease) * (size_t)n.cumC(dmax), hipMemcpyDeviceToHost, st), "download creases");
    if (steps & CCD_POINTS)
        CC_TRY(hipMemcpyAsync(subd->vertexPoints, s->pts, sizeof(P3) * (size_t)n.cumV(dmax), hipMemcpyDeviceToHost, st), "download points");
    CC_TRY(hipStreamSynchronize(st), "refine");
    return true;
}

}  // namespace

namespace cch {
ccd_Subd *host_mirror(const cc_Subd *subd) {
    if (!subd || !subd->cage) {
        set_error(11, "null subd or cage");
        return nullptr;
    }
    const char *why = nullptr;
    if (!ref_counts_fit_int32(subd->cage, subd->maxDepth, &why)) {
        set_error(12, why);
        return nullptr;
    }
    return mirror_for(subd);
}
}  // namespace cch

extern "C" {

cc_Subd *ccs_Create(const cc_Mesh *cage, int32_t maxDepth) {
    clear_error();
    const char *why = nullptr;
    if (!cage || !ref_counts_fit_int32(cage, maxDepth, &why)) {
        char buf[256];
        snprintf(buf, sizeof(buf), "ccs_Create: %s", cage ? why : "null cage");
        fprintf(stderr, "cc: %s\n", buf);
        set_error(12, buf);
        return NULL;
    }
    Counts n;
    n.init(cage->vertexCount, cage->halfedgeCount, cage->edgeCount, cage->faceCount);
    cc_Subd *s = (cc_Subd *)malloc(sizeof(cc_Subd));
    if (!s) { set_error(2, "ccs_Create: out of host memory"); return NULL; }
    s->maxDepth = maxDepth;
    s->cage = cage;
    s->halfedges = (cc_Halfedge_SemiRegular *)host_alloc(sizeof(cc_Halfedge_SemiRegular) * (size_t)n.cumH(maxDepth));
    s->creases = (cc_Crease *)host_alloc(sizeof(cc_Crease) * (size_t)n.cumC(maxDepth) + 1);
    s->vertexPoints = (cc_VertexPoint *)host_alloc(sizeof(cc_VertexPoint) * (size_t)n.cumV(maxDepth));
    if (!s->halfedges || !s->creases || !s->vertexPoints) {
        host_free(s->halfedges); host_free(s->creases); host_free(s->vertexPoints); free(s);
        set_error(2, "ccs_Create: out of host memory");
        fprintf(stderr, "cc: ccs_Create: out of host memory\n");
        return NULL;
    }
    return s;
}

void ccs_Release(cc_Subd *subd) {
    if (!subd) return;
    drop_mirror(subd);
    host_free(subd->halfedges);
    host_free(subd->creases);
    host_free(subd->vertexPoints);
    free(subd);
}

#define CC_HOST_ENTRY(name, steps, dmax) \
    void name(cc_Subd *subd) { (void)host_refine(subd, (steps), (dmax), #name); }

CC_HOST_ENTRY(ccs_Refine_Scatter, CCD_ALL, subd ? subd->maxDepth : 1)
CC_HOST_ENTRY(ccs_Refine_Gather, CCD_ALL | CCD_GATHER, subd ? subd->maxDepth : 1)
CC_HOST_ENTRY(ccs_Refine_NoCreases_Scatter, CCD_ALL | CCD_NOCREASES, subd ? subd->maxDepth : 1)
CC_HOST_ENTRY(ccs_Refine_NoCreases_Gather, CCD_ALL | CCD_GATHER | CCD_NOCREASES, subd ? subd->maxDepth : 1)
CC_HOST_ENTRY(ccs_RefineVertexPoints_Scatter, CCD_POINTS, subd ? subd->maxDepth : 1)
CC_HOST_ENTRY(ccs_RefineVertexPoints_Gather, CCD_POINTS | CCD_GATHER, subd ? subd->maxDepth : 1)
CC_HOST_ENTRY(ccs_RefineVertexPoints_NoCreases_Scatter, CCD_POINTS | CCD_NOCREASES, subd ? subd->maxDepth : 1)
CC_HOST_ENTRY(ccs_RefineVertexPoints_NoCreases_Gather, CCD_POINTS | CCD_GATHER | CCD_NOCREASES, subd ? subd->maxDepth : 1)
CC_HOST_ENTRY(ccs_RefineHalfedges, CCD_HALFEDGES, subd ? subd->maxDepth : 1)
CC_HOST_ENTRY(ccs_RefineCreases, CCD_CREASES, subd ? subd->maxDepth : 1)
CC_HOST_ENTRY(ccs_RefineCageHalfedges, CCD_HALFEDGES, 1)
CC_HOST_ENTRY(ccs_RefineCageCreases, CCD_CREASES, 1)
#ifndef CC_DISABLE_UV   // absent from the reference's CC_DISABLE_UV build
CC_HOST_ENTRY(ccs_RefineCageVertexUvs, CCD_UVS, 1)
#endif
CC_HOST_ENTRY(ccs_RefineCageVertexPoints_Scatter, CCD_POINTS, 1)
CC_HOST_ENTRY(ccs_RefineCageVertexPoints_Gather, CCD_POINTS | CCD_GATHER, 1)

// The adjoint on a host cc_Subd: the topology and creases it reads are the
// subd's host arrays (levels 1..D-1), uploaded with the cage and gradient.
static void host_adjoint(const cc_Subd *subd, const void *dL_dFinest, void *dL_dCage, size_t scalar) {
    clear_error();
    if (!subd || !subd->cage || !dL_dFinest || !dL_dCage) {
        set_error(11, "ccs_RefineVertexPoints_Adjoint: null argument");
        fprintf(stderr, "cc: ccs_RefineVertexPoints_Adjoint: null argument\n");
        return;
    }
    const char *why = nullptr;
    if (!ref_counts_fit_int32(subd->cage, subd->maxDepth, &why)) {
        set_error(12, why);
        return;
    }
    ccd_Subd *s = mirror_for(subd);
    if (!s) return;
    DeviceGuard g(s->device);
    hipStream_t st = s->stream;
    const Counts &n = s->n;
    const int D = s->D;
    if (!s->upload_cage(subd->cage, st)) return;
    // gradient staging, sized for doubles so both precisions fit
    if (!s->adj_in && hip_failed(hipMalloc(&s->adj_in, 3 * sizeof(double) * (size_t)n.V[D]), "hipMalloc(adjoint input)")) return;
    if (!s->adj_out && hip_failed(hipMalloc(&s->adj_out, 3 * sizeof(double) * (size_t)n.V[0]), "hipMalloc(adjoint output)")) return;
    if (D > 1 && (hip_failed(hipMemcpyAsync(s->he, subd->halfedges, sizeof(He) * (size_t)n.cumH(D - 1), hipMemcpyHostToDevice, st), "upload halfedges") ||
                  hip_failed(hipMemcpyAsync(s->cr, subd->creases, sizeof(Crease) * (size_t)n.cumC(D - 1), hipMemcpyHostToDevice, st), "upload creases")))
        return;
    if (hip_failed(hipMemcpyAsync(s->adj_in, dL_dFinest, 3 * scalar * (size_t)n.V[D], hipMemcpyHostToDevice, st), "upload gradient"))
        return;
    if (!s->adjoint(s->adj_in, s->adj_out, scalar == sizeof(double) ? CCD_FP64 : 0, st)) return;
    if (hip_failed(hipMemcpyAsync(dL_dCage, s->adj_out, 3 * scalar * (size_t)n.V[0], hipMemcpyDeviceToHost, st), "download gradient") ||
        hip_failed(hipStreamSynchronize(st), "ccs_RefineVertexPoints_Adjoint"))
        return;
}

void ccs_RefineVertexPoints_Adjoint(const cc_Subd *subd, const cc_VertexPoint *dL_dFinest, cc_VertexPoint *dL_dCage) {
    host_adjoint(subd, dL_dFinest, dL_dCage, sizeof(float));
}

void ccs_RefineVertexPoints_Adjoint_f64(const cc_Subd *subd, const double *dL_dFinest, double *dL_dCage) {
    host_adjoint(subd, dL_dFinest, dL_dCage, sizeof(double));
}

#ifndef CC_DISABLE_UV   // absent from the reference's CC_DISABLE_UV build (CatmullClark.h:204-206)
// no-op without cage uvs, like the reference (CatmullClark.h:2930)
void ccs_RefineVertexUvs(cc_Subd *subd) {
    if (subd && subd->cage && subd->cage->uvCount <= 0) { clear_error(); return; }
    (void)host_refine(subd, CCD_UVS, subd ? subd->maxDepth : 1, "ccs_RefineVertexUvs");
}
#endif

// ---- device-resident API -------------------------------------------------------
ccd_Subd *ccd_Create(const cc_Mesh *cage, int32_t maxDepth) {
    clear_error();
    if (!cage || maxDepth < 1 || maxDepth > kMaxDepth || cage->halfedgeCount < 3 || cage->edgeCount < 1 ||
        cage->faceCount < 1 || cage->vertexCount < 1) {
        set_error(12, "ccd_Create: invalid cage or depth");
        return nullptr;
    }
    Counts n;
    n.init(cage->vertexCount, cage->halfedgeCount, cage->edgeCount, cage->faceCount);
    for (int d = 0; d <= maxDepth; ++d) {
        if (n.H[d] > INT32_MAX || n.V[d] + n.E[d] + n.F[d] > INT32_MAX || 2 * n.E[d] + n.H[d] > INT32_MAX) {
            set_error(12, "ccd_Create: IDs at this depth overflow int32");
            return nullptr;
        }
    }
    ccd_Subd *s = create_mirror(cage, maxDepth);
    if (!s) return nullptr;
    DeviceGuard g(s->device);
    if (!s->upload_cage(cage, s->stream) || hip_failed(hipStreamSynchronize(s->stream), "ccd_Create")) {
        s->free_all();
        delete s;
        return nullptr;
    }
    return s;
}

void ccd_Release(ccd_Subd *s) {
    if (!s) return;
    s->free_all();
    delete s;
}

int32_t ccd_MaxDepth(const ccd_Subd *s) { return s ? s->D : 0; }

int64_t ccd_Count(const ccd_Subd *s, int32_t what, int32_t d) {
    if (!s || d < 0 || d > kMaxDepth) return -1;
    switch (what) {
    case 0: return s->n.H[d];
    case 1: return s->n.C[d];
    case 2: return s->n.F[d];
    case 3: return s->n.E[d];
    case 4: return s->n.V[d];
    case 5: return s->n.cumH(d);
    case 6: return s->n.cumC(d);
    case 7: return s->n.cumV(d);
    default: return -1;
    }
}

void *ccd_DevicePointer(const ccd_Subd *s, int32_t which) {
    if (!s) return nullptr;
    switch (which) {
    case CCD_PTR_HALFEDGES: return s->he;
    case CCD_PTR_CREASES: return s->cr;
    case CCD_PTR_POINTS: return s->pts;
    case CCD_PTR_CAGE_POINTS: return s->c_pts;
    default: return nullptr;
    }
}

int32_t ccd_UploadCagePoints(ccd_Subd *s, const cc_VertexPoint *hostPoints) {
    clear_error();
    if (!s || !hostPoints) { set_error(11, "ccd_UploadCagePoints: null argument"); return t_err; }
    DeviceGuard g(s->device);
    // work queued on any stream may still read the cage points
    if (hip_failed(hipDeviceSynchronize(), "ccd_UploadCagePoints") ||
        hip_failed(hipMemcpyAsync(s->c_pts, hostPoints, sizeof(P3) * (size_t)s->V0, hipMemcpyHostToDevice, s->stream),
                   "ccd_UploadCagePoints") ||
        hip_failed(hipStreamSynchronize(s->stream), "ccd_UploadCagePoints"))
        return t_err;
    return 0;
}

int32_t ccd_Refine(ccd_Subd *s, int32_t steps, void *stream) {
    clear_error();
    if (!s) { set_error(11, "ccd_Refine: null subd"); return t_err; }
    DeviceGuard g(s->device);
    hipStream_t st = (hipStream_t)stream;
    // a caller capturing its own stream gets the launches in its capture
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (st && hip_failed(hipStreamIsCapturing(st, &cap), "ccd_Refine")) return t_err;
    const bool replay = s->opt_graph && !s->timing && cap == hipStreamCaptureStatusNone;
    const bool ok = replay ? s->refine_graph(steps, st) : s->refine(steps, s->D, st);
    return ok ? 0 : (t_err ? t_err : 1);
}

int32_t ccd_RefinePoints(ccd_Subd *s, const float *devCagePoints, float *devPoints, int32_t steps, void *stream) {
    clear_error();
    if (!s) { set_error(11, "ccd_RefinePoints: null subd"); return t_err; }
    DeviceGuard g(s->device);
    hipStream_t st = (hipStream_t)stream;
    if (s->timing) s->begin_timing();
    P3 *out = devPoints ? reinterpret_cast<P3 *>(devPoints) : s->pts;
    if (!s->refine_points(steps | CCD_POINTS, s->D, reinterpret_cast<const P3 *>(devCagePoints), out, st)) return t_err ? t_err : 1;
    if (hip_failed(hipGetLastError(), "ccd_RefinePoints")) return t_err;
    return 0;
}

int32_t ccd_RefinePointsAdjoint(ccd_Subd *s, const float *devGradFinest, float *devGradCage, int32_t steps,
                                void *stream) {
    clear_error();
    if (!s || !devGradFinest || !devGradCage) { set_error(11, "ccd_RefinePointsAdjoint: null argument"); return t_err; }
    DeviceGuard g(s->device);
    hipStream_t st = (hipStream_t)stream;
    if (!s->adjoint(devGradFinest, devGradCage, steps, st)) return t_err ? t_err : 1;
    return 0;
}

int32_t ccd_Download(const ccd_Subd *s, cc_Subd *dst, int32_t steps) {
    clear_error();
    if (!s || !dst) { set_error(11, "ccd_Download: null argument"); return t_err; }
    if (dst->maxDepth != s->D) { set_error(13, "ccd_Download: depth mismatch"); return t_err; }
    DeviceGuard g(s->device);
    // the refinement may have been queued on any stream
    if (hip_failed(hipDeviceSynchronize(), "ccd_Download")) return t_err;
    hipStream_t st = s->stream;
    if ((steps & (CCD_HALFEDGES | CCD_UVS)) &&
        hip_failed(hipMemcpyAsync(dst->halfedges, s->he, sizeof(He) * (size_t)s->n.cumH(s->D), hipMemcpyDeviceToHost, st), "download"))
        return t_err;
    if ((steps & CCD_CREASES) &&
        hip_failed(hipMemcpyAsync(dst->creases, s->cr, sizeof(Crease) * (size_t)s->n.cumC(s->D), hipMemcpyDeviceToHost, st), "download"))
        return t_err;
    if ((steps & CCD_POINTS) &&
        hip_failed(hipMemcpyAsync(dst->vertexPoints, s->pts, sizeof(P3) * (size_t)s->n.cumV(s->D), hipMemcpyDeviceToHost, st), "download"))
        return t_err;
    if (hip_failed(hipStreamSynchronize(st), "download")) return t_err;
    return 0;
}

int32_t ccd_Synchronize(ccd_Subd *s) {
    clear_error();
    if (!s) { set_error(11, "ccd_Synchronize: null subd"); return t_err; }
    DeviceGuard g(s->device);
    if (hip_failed(hipDeviceSynchronize(), "ccd_Synchronize")) return t_err;
    return 0;
}

int32_t ccd_SetOption(ccd_Subd *s, int32_t key, int32_t value) {
    if (!s) return 11;
    DeviceGuard g(s->device);
    s->drop_graphs();   // captured with the previous choices
    switch (key) {
    case CCB_OPT_GRAPH: s->opt_graph = value != 0; return 0;
    case CCD_OPT_FUSED:
        s->opt_fused = value < 0 ? 0 : (value > 2 ? 2 : value);
        if (s->opt_fused)
            for (auto &c : s->cfg) c.kind = s->opt_fused;
        return 0;
    case CCD_OPT_STORE: s->opt_store = value ? kStoreLds : kStoreDirect; return 0;
    case CCD_OPT_XCD_REMAP: s->opt_remap = value != 0; return 0;
    case CCD_OPT_FAN: s->opt_fan = value < 0 ? 0 : (value > 3 ? 3 : value); return 0;
    case CCD_OPT_PATCH_TILE:
        if (value != 128 && value != 256 && value != 512) return 14;
        for (auto &c : s->cfg) c.nt = value;
        return 0;
    case CCD_OPT_PATCH_HPT:
        if (value < 1 || value > 64) return 14;
        for (auto &c : s->cfg) c.hpt = value;
        return 0;
    case CCD_OPT_SCAN:
        if (value != 0 && value != 2 && value != 4 && value != 8) return 14;
        for (auto &c : s->cfg) c.scan = value;
        return 0;
    case CCD_OPT_ADJ_VERTEX:
        if (value < 0) return 14;
        s->opt_adj_vertex = value;
        return 0;
    case CCD_OPT_FUSED_BLOCK:
        if (value != 256 && value != 512 && value != 1024) return 14;
        for (auto &c : s->cfg) c.fb = value;
        return 0;
    case CCD_OPT_FUSED_TILES:
        if (value != 1 && value != 2 && value != 4) return 14;
        for (auto &c : s->cfg) c.ft = value;
        return 0;
    case CCD_OPT_VERTEX_WAVES:
        if (value != 4 && value != 6 && value != 8) return 14;
        for (auto &c : s->cfg) c.vw = value;
        return 0;
    default: return 14;
    }
}

// Times candidate kernel configurations on this subd with HIP events and
// keeps, per level, the fastest, in three rounds:
//   1. the fused kernel's geometry: 256, 512 or 1024 threads per workgroup
//      times 1, 2 or 4 tiles per workgroup (masked vertex kernel);
//   2. from each level's pick: the masked vertex kernel compiled for 4 or 6
//      waves per SIMD, or the compacted vertex kernel testing 2, 4 or 8
//      halfedges per thread;
//   3. the patch-local kernel with tiles of 128, 256 or 512 halfedges.
// Each candidate runs the given steps three times (the first discarded); a
// level's time is that of its fused (or patch) and vertex kernels.
int32_t ccd_Subd::autotune(int32_t steps, hipStream_t st) {
    if (!(steps & CCD_POINTS) || opt_fused == 0 || D < 2) return 0;
    drop_graphs();
    float best[kMaxDepth + 1];
    LevelCfg pick[kMaxDepth + 1];
    for (int d = 0; d <= kMaxDepth; ++d) { best[d] = 1e30f; pick[d] = cfg[d]; }
    const bool was_timing = timing;
    timing = true;
    // level times of one candidate (cfg set per level), false on error
    auto measure = [&](float (&ms)[kMaxDepth + 1]) -> bool {
        for (auto &x : ms) x = 0.0f;
        for (int rep = 0; rep < 3; ++rep) {
            begin_timing();
            if (!refine(steps, D, st)) return false;
            if (hip_failed(hipStreamSynchronize(st), "autotune")) return false;
            if (rep == 0) continue;  // warm-up
            // level d's kernels are the timed launches between its first
            // kernel and the next level's; they come in level order
            int d = 0;
            for (auto &t : timed) {
                const bool first = !strcmp(t.name, "level_fused") || !strcmp(t.name, "level_points") ||
                                   !strcmp(t.name, "level_patch");
                const bool level_kernel = first || !strcmp(t.name, "level_vertices") || !strcmp(t.name, "level_deferred");
                if (!level_kernel) continue;
                if (first) ++d;
                float x = 0.0f;
                (void)hipEventElapsedTime(&x, t.start, t.stop);
                if (d >= 1 && d <= kMaxDepth) ms[d] += x;
            }
        }
        return true;
    };
    auto keep = [&](const float (&ms)[kMaxDepth + 1]) {
        for (int d = 1; d < D; ++d)
            if (ms[d] < best[d]) { best[d] = ms[d]; pick[d] = cfg[d]; }
    };
    float ms[kMaxDepth + 1];
    bool ok = true;
    for (int fb : {256, 512, 1024}) {
        for (int ft : {1, 2, 4}) {
            for (auto &c : cfg) { c = LevelCfg{}; c.fb = fb; c.ft = ft; }
            if (!(ok = measure(ms))) break;
            keep(ms);
        }
        if (!ok) break;
    }
    LevelCfg geometry[kMaxDepth + 1];
    for (int d = 0; d <= kMaxDepth; ++d) geometry[d] = pick[d];
    const int variants[][2] = {{0, 4}, {0, 6}, {2, kVertexWaves}, {4, kVertexWaves}, {8, kVertexWaves}};
    for (const auto &v : variants) {
        if (!ok) break;
        for (int d = 0; d <= kMaxDepth; ++d) { cfg[d] = geometry[d]; cfg[d].scan = v[0]; cfg[d].vw = v[1]; }
        if ((ok = measure(ms))) keep(ms);
    }
    for (int nt : {128, 256, 512}) {
        if (!ok) break;
        for (auto &c : cfg) { c = LevelCfg{}; c.kind = 2; c.scan = 4; c.nt = nt; }
        if ((ok = measure(ms))) keep(ms);
    }
    begin_timing();
    timing = was_timing;
    for (int d = 0; d <= kMaxDepth; ++d) cfg[d] = pick[d];
    return ok ? 0 : (last_error() ? last_error() : 1);
}

int32_t ccd_TileOption(const ccd_Subd *s, int32_t key, int32_t level) {
    if (!s || level < 0 || level > kMaxDepth) return -1;
    switch (key) {
    case CCD_OPT_FUSED: return s->opt_fused ? s->cfg[level].kind : 0;
    case CCD_OPT_PATCH_TILE: return s->cfg[level].nt;
    case CCD_OPT_PATCH_HPT: return s->cfg[level].hpt;
    case CCD_OPT_SCAN: return s->cfg[level].scan;
    case CCD_OPT_FUSED_BLOCK: return s->cfg[level].fb;
    case CCD_OPT_FUSED_TILES: return s->cfg[level].ft;
    case CCD_OPT_VERTEX_WAVES: return s->cfg[level].vw;
    default: return -1;
    }
}

int32_t ccd_Autotune(ccd_Subd *s, int32_t steps, void *stream) {
    if (!s) return 11;
    DeviceGuard g(s->device);
    return s->autotune(steps, (hipStream_t)stream);
}

int32_t ccd_SetTiming(ccd_Subd *s, int32_t enable) {
    if (!s) return 11;
    s->timing = enable != 0;
    return 0;
}

int32_t ccd_TimingCount(const ccd_Subd *s) { return s ? (int32_t)s->timed.size() : 0; }

const char *ccd_TimingName(const ccd_Subd *s, int32_t i) {
    return (s && i >= 0 && i < (int32_t)s->timed.size()) ? s->timed[i].name : "";
}

float ccd_TimingMs(const ccd_Subd *s, int32_t i) {
    if (!s || i < 0 || i >= (int32_t)s->timed.size()) return -1.0f;
    float ms = -1.0f;
    (void)hipEventSynchronize(s->timed[i].stop);
    if (hipEventElapsedTime(&ms, s->timed[i].start, s->timed[i].stop) != hipSuccess) return -1.0f;
    return ms;
}

}  // extern "C"
