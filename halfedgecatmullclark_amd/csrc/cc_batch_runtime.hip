// Runtime of the batched C ABI (ccb_*): many cages, each with its own depth,
// refined together.  Every cage keeps the cc_Subd layout of its levels
// 1..D_i inside three shared device arrays (halfedges, creases, points) at
// 64-bit offsets, so one cage's slice downloads into a cc_Subd as is.  Each
// kernel kind runs once per level over all cages taking part (job tables in
// cc_batch.h); the whole sequence is captured once per set of steps into a
// hipGraph and replayed.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "cc_batch.h"
#include "cc_internal.h"
#include "cc_kernels.h"
#include "cc_level.h"
#include "cc_patch.h"

using namespace cch;

struct ccb_Batch {
    struct Cage {
        int32_t D = 0, V0 = 0, U0 = 0, H0 = 0, E0 = 0, F0 = 0;
        Counts n;
        LevelTab tab;
        // element offsets: cage arrays, then levels 1..D, then per-cage scratch
        int64_t o_he0 = 0, o_v0 = 0, o_u0 = 0, o_e0 = 0, o_f0 = 0;
        // element offsets of the cage's ring lists in `rings` (-1: walk the cage)
        int64_t o_vro = -1, o_vr = -1, o_vb = -1, o_fro = -1, o_fr = -1, o_voo = -1, o_vo = -1;
        int64_t o_he = 0, o_cr = 0, o_pts = 0, o_fin = 0, o_def = 0, o_adjB = 0, o_adjW = 0, o_rare = 0;
        int64_t o_adjVP = 0;   // VParam offset of levels 1..D-1's vertex-rule parameters
        int32_t rare_cap = 0;   // rare-vertex list entries (FusedArgs::rare), the count first
        int64_t o_erep = -1;    // word offset of the edge-representative bits (depth >= 4)
    };
    struct Timed {
        const char *name;
        int level;
        hipEvent_t start, stop;
    };

    int device = 0;
    hipStream_t stream = nullptr;
    int32_t N = 0, Dmax = 0;
    std::vector<Cage> cages;
    int64_t tot_he0 = 0, tot_v0 = 0, tot_u0 = 0, tot_e0 = 0, tot_f0 = 0;
    std::vector<int32_t> rings_host;   // every cage's ring lists, copied into `rings`
    int32_t *rings = nullptr;
    int64_t tot_he = 0, tot_cr = 0, tot_pts = 0, tot_fin = 0, tot_def = 0, tot_adjB = 0, tot_adjW = 0, tot_rare = 0;
    int64_t tot_adjVP = 0;
    int32_t *rare = nullptr;
    int64_t tot_erep = 0;
    uint64_t *erep = nullptr;   // edge-representative bits of levels 1..D-2, per cage (FusedArgs::erep)
    // levels 1..erep_kept hold bits written by a forward pass already queued
    // (they depend on the topology only); the adjoint tables were built when
    // it was erep_tables
    int erep_kept = 0, erep_tables = 0;
    // device memory: every array of the forward pass is carved out of one
    // arena, the adjoint's out of a second one (job tables hold byte offsets
    // into them)
    char *arena = nullptr, *adj_arena = nullptr;
    CageHe *c_he = nullptr;
    int32_t *c_v2h = nullptr, *c_e2h = nullptr, *c_f2h = nullptr;
    P3 *c_pts = nullptr;
    UV *c_uvs = nullptr;
    Crease *c_cr = nullptr;
    He *he = nullptr;
    Crease *cr = nullptr;
    P3 *pts = nullptr;
    uint64_t *defer_mask = nullptr;   // deferred-vertex masks, per cage one word per 64 halfedges
    float *adjA = nullptr, *adjB = nullptr, *adjW = nullptr, *adjCage = nullptr;
    // every cage's vertex-rule parameters of levels 1..D-1, kept across
    // adjoint calls (a batch's topology and creases never change); valid for
    // the creased (1) or uncreased (0) rules, -1 before the first call
    VParam<float> *adjVP = nullptr;
    int adj_vp_key = -1;
    void *tables = nullptr;
    size_t tables_cap = 0;
    // launch tables
    BatchTable t_cage_he, t_cage_f, t_cage_e, t_cage_v, t_cage_h3, t_cage_zero, t_fin;
    BatchTable t_cr[kMaxDepth + 1], t_patch[kMaxDepth + 1], t_def[kMaxDepth + 1], t_adj[kMaxDepth + 1];
    BatchTable t_rare[kMaxDepth + 1];   // rare-vertex lists after the scan-0 vertex kernel
    BatchTable t_fused[kMaxDepth + 1], t_vert[kMaxDepth + 1], t_adjq[kMaxDepth + 1];
    BatchTable t_adjr[kMaxDepth + 1];   // the adjoint vertex pass's rare-vertex lists
    BatchTable t_adjc_e, t_adjc_f, t_adjc_v;   // cage-step adjoint gathers: per edge, face, vertex
    bool all_rings = true;                     // every cage has its ring lists
    // the adjoint's finest steps read the caller's gradient (its address in
    // the adjoint arena's first bytes); false when a cage is refined once:
    // its cage step updates the finest gradient in place, so every call
    // copies the gradient to adjA
    bool gin_ok = true;
    // per-level kernels: kind 1 fused (fb threads per workgroup, ft tiles
    // of fb halfedges per workgroup) + vertex points (scan 0: the masked
    // kernel compiled for vw waves per SIMD; else the compacted kernel with
    // scan halfedges tested per thread), kind 2 patch-local (tiles of nt,
    // hpt tiles per workgroup) + deferred vertices
    struct LevelCfg {
        int kind = 1, scan = 0, nt = 256, hpt = 1;
        int fb = kFusedBlock, ft = 1, vw = kVertexWaves;
    };
    LevelCfg cfg[kMaxDepth + 1];
    int remap = 1, use_graph = 1;
    bool timing = false;
    std::vector<Timed> timed;
    std::map<int32_t, hipGraphExec_t> graphs;   // forward: steps; adjoint: -1 - creased - 2 params
    bool topo_done = false, creases_done = false;

    bool create(const cc_Mesh *const *meshes, const int32_t *depths, int32_t count);
    void free_all();
    bool build_tables();
    void drop_graphs();
    bool alloc_adjoint();
    void record_forward(int32_t steps, hipStream_t st);
    // flags: 1 creased rules, 2 run the parameter pass
    void record_adjoint(int flags, hipStream_t st);
    bool run(int32_t key, bool adjoint, int32_t steps, hipStream_t st);
    int32_t autotune(int32_t steps, hipStream_t st);
    void begin_timing();
    template <typename Fn>
    void timed_launch(const char *name, int level, hipStream_t st, Fn &&fn);
    DCage cage_view(int i) const;
};

namespace {

#define CC_TRY(expr, what) \
    do { if (hip_failed((expr), (what))) return false; } while (0)

int32_t fail(int32_t code, const char *msg) {
    set_error(code, msg);
    return code;
}

// builds the job list and workgroup prefix sum of one launch into a blob
struct TableBuilder {
    std::vector<char> blob;
    struct Pending {
        BatchTable *t;
        size_t jobs, start;
    };
    std::vector<Pending> pending;

    size_t put(const void *p, size_t bytes) {
        const size_t off = (blob.size() + 255) & ~size_t(255);
        blob.resize(off + bytes);
        if (bytes) memcpy(blob.data() + off, p, bytes);
        return off;
    }
    template <class Job>
    void add(BatchTable &t, const std::vector<Job> &jobs, const std::vector<int64_t> &blocks) {
        int64_t total = 0;
        for (size_t k = 0; k < jobs.size(); ++k) total += blocks[k];
        // first workgroup of each job, the total, then each workgroup's job
        std::vector<int32_t> start(jobs.size() + 1 + (size_t)total, 0);
        int32_t *owner = start.data() + jobs.size() + 1;
        total = 0;
        for (size_t k = 0; k < jobs.size(); ++k) {
            start[k] = (int32_t)total;
            for (int64_t b = 0; b < blocks[k]; ++b) owner[total + b] = (int32_t)k;
            total += blocks[k];
        }
        start[jobs.size()] = (int32_t)total;
        t.n = (int32_t)jobs.size();
        t.blocks = (uint32_t)total;
        pending.push_back({&t, put(jobs.data(), sizeof(Job) * jobs.size()), put(start.data(), sizeof(int32_t) * start.size())});
    }
    void resolve(void *base) {
        for (auto &p : pending) {
            p.t->jobs = static_cast<char *>(base) + p.jobs;
            p.t->start = reinterpret_cast<const int32_t *>(static_cast<char *>(base) + p.start);
        }
    }
};

int64_t blocks_of(int64_t threads, int64_t per = 256) { return (threads + per - 1) / per; }

// consecutive 256-byte aligned ranges of one allocation
struct Carve {
    size_t size = 0;
    size_t take(size_t bytes) {
        const size_t off = size;
        size = (off + bytes + 255) & ~size_t(255);
        return off;
    }
};

}  // namespace

DCage ccb_Batch::cage_view(int i) const {
    const Cage &g = cages[i];
    DCage c;
    c.he = c_he + g.o_he0;
    c.v2h = c_v2h + g.o_v0;
    c.e2h = c_e2h + g.o_e0;
    c.f2h = c_f2h + g.o_f0;
    c.pts = c_pts + g.o_v0;
    c.uvs = c_uvs + g.o_u0;
    c.cr = c_cr + g.o_e0;
    auto ring = [&](int64_t o) -> const int32_t * { return o < 0 ? nullptr : rings + o; };
    c.vring_off = ring(g.o_vro); c.vring = ring(g.o_vr); c.vring_bnd = ring(g.o_vb);
    c.fring_off = ring(g.o_fro); c.fring = ring(g.o_fr);
    c.vout_off = ring(g.o_voo); c.vout = ring(g.o_vo);
    c.V = g.V0; c.U = g.U0; c.H = g.H0; c.E = g.E0; c.F = g.F0;
    return c;
}

bool ccb_Batch::create(const cc_Mesh *const *meshes, const int32_t *depths, int32_t count) {
    N = count;
    cages.resize(count);
    for (int i = 0; i < count; ++i) {
        const cc_Mesh *m = meshes[i];
        Cage &g = cages[i];
        g.D = depths[i];
        g.V0 = m->vertexCount; g.U0 = m->uvCount; g.H0 = m->halfedgeCount; g.E0 = m->edgeCount; g.F0 = m->faceCount;
        g.n.init(g.V0, g.H0, g.E0, g.F0);
        g.tab = level_tab(g.n);
        g.o_he0 = tot_he0; tot_he0 += g.H0;
        g.o_v0 = tot_v0; tot_v0 += g.V0;
        g.o_u0 = tot_u0; tot_u0 += g.U0;
        g.o_e0 = tot_e0; tot_e0 += g.E0;
        g.o_f0 = tot_f0; tot_f0 += g.F0;
        const CageRings r = build_cage_rings(m);
        all_rings = all_rings && r.ok;
        if (r.ok) {
            auto put = [&](const std::vector<int32_t> &v) {
                const int64_t o = (int64_t)rings_host.size();
                rings_host.insert(rings_host.end(), v.begin(), v.end());
                rings_host.push_back(0);
                return o;
            };
            g.o_vro = put(r.voff);
            g.o_vb = put(r.vbnd);
            g.o_fro = put(r.foff);
            g.o_vr = put(r.vring);
            g.o_fr = put(r.fring);
            g.o_voo = put(r.ooff);
            g.o_vo = put(r.out);
        }
        g.o_he = tot_he; tot_he += g.n.cumH(g.D);
        g.o_cr = tot_cr; tot_cr += g.n.cumC(g.D);
        g.o_pts = tot_pts; tot_pts += g.n.cumV(g.D);
        g.o_fin = tot_fin; tot_fin += g.n.V[g.D];
        g.o_def = tot_def; tot_def += g.D >= 2 ? (g.n.H[g.D - 1] + 63) / 64 : 1;
        {
            int64_t b0 = 0;
            const CageHe *ch = reinterpret_cast<const CageHe *>(m->halfedges);
            for (int32_t h = 0; h < g.H0; ++h) b0 += ch[h].twin < 0 ? 1 : 0;
            g.rare_cap = (int32_t)std::min<int64_t>((int64_t)g.V0 + g.F0 + (b0 << (g.D - 1)), INT32_MAX - 1);
            g.o_rare = tot_rare;
            tot_rare += g.rare_cap + 1;
        }
        if (g.D >= 4) {
            g.o_erep = tot_erep;
            tot_erep += g.tab.R[g.D - 1];
        }
        g.o_adjB = tot_adjB; tot_adjB += g.n.V[g.D - 1];
        g.o_adjW = tot_adjW; tot_adjW += adj_scratch_scalars(g.n, g.D, true);
        if (g.D == 1) gin_ok = false;
        g.o_adjVP = tot_adjVP; tot_adjVP += g.n.cumV(g.D - 1);
        Dmax = std::max(Dmax, g.D);
    }
    CC_TRY(hipGetDevice(&device), "hipGetDevice");
    CC_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    {
        Carve cv;
        const size_t o_che = cv.take(sizeof(CageHe) * (size_t)tot_he0);
        const size_t o_cv2h = cv.take(sizeof(int32_t) * (size_t)tot_v0);
        const size_t o_ce2h = cv.take(sizeof(int32_t) * (size_t)tot_e0);
        const size_t o_cf2h = cv.take(sizeof(int32_t) * (size_t)tot_f0);
        const size_t o_cpts = cv.take(sizeof(P3) * (size_t)tot_v0);
        const size_t o_cuvs = cv.take(sizeof(UV) * (size_t)(tot_u0 + 1));
        const size_t o_ccr = cv.take(sizeof(Crease) * (size_t)tot_e0);
        const size_t o_rings = cv.take(sizeof(int32_t) * (rings_host.size() + 1));
        const size_t o_he = cv.take(sizeof(He) * (size_t)tot_he);
        const size_t o_cr = cv.take(sizeof(Crease) * (size_t)tot_cr);
        const size_t o_pts = cv.take(sizeof(P3) * (size_t)tot_pts);
        const size_t o_def = cv.take(sizeof(uint64_t) * (size_t)tot_def);
        const size_t o_rare = cv.take(sizeof(int32_t) * (size_t)tot_rare);
        const size_t o_erep = cv.take(sizeof(uint64_t) * (size_t)(tot_erep + 1));
        CC_TRY(hipMalloc(&arena, cv.size), "hipMalloc(batch arena)");
        c_he = reinterpret_cast<CageHe *>(arena + o_che);
        c_v2h = reinterpret_cast<int32_t *>(arena + o_cv2h);
        c_e2h = reinterpret_cast<int32_t *>(arena + o_ce2h);
        c_f2h = reinterpret_cast<int32_t *>(arena + o_cf2h);
        c_pts = reinterpret_cast<P3 *>(arena + o_cpts);
        c_uvs = reinterpret_cast<UV *>(arena + o_cuvs);
        c_cr = reinterpret_cast<Crease *>(arena + o_ccr);
        rings = reinterpret_cast<int32_t *>(arena + o_rings);
        he = reinterpret_cast<He *>(arena + o_he);
        cr = reinterpret_cast<Crease *>(arena + o_cr);
        pts = reinterpret_cast<P3 *>(arena + o_pts);
        defer_mask = reinterpret_cast<uint64_t *>(arena + o_def);
        rare = reinterpret_cast<int32_t *>(arena + o_rare);
        erep = reinterpret_cast<uint64_t *>(arena + o_erep);
    }
    for (int i = 0; i < count; ++i) {
        const cc_Mesh *m = meshes[i];
        const Cage &g = cages[i];
        CC_TRY(hipMemcpyAsync(c_he + g.o_he0, m->halfedges, sizeof(CageHe) * g.H0, hipMemcpyHostToDevice, stream), "upload");
        CC_TRY(hipMemcpyAsync(c_v2h + g.o_v0, m->vertexToHalfedgeIDs, sizeof(int32_t) * g.V0, hipMemcpyHostToDevice, stream), "upload");
        CC_TRY(hipMemcpyAsync(c_e2h + g.o_e0, m->edgeToHalfedgeIDs, sizeof(int32_t) * g.E0, hipMemcpyHostToDevice, stream), "upload");
        CC_TRY(hipMemcpyAsync(c_f2h + g.o_f0, m->faceToHalfedgeIDs, sizeof(int32_t) * g.F0, hipMemcpyHostToDevice, stream), "upload");
        CC_TRY(hipMemcpyAsync(c_pts + g.o_v0, m->vertexPoints, sizeof(P3) * g.V0, hipMemcpyHostToDevice, stream), "upload");
        if (g.U0 > 0)
            CC_TRY(hipMemcpyAsync(c_uvs + g.o_u0, m->uvs, sizeof(UV) * g.U0, hipMemcpyHostToDevice, stream), "upload");
        CC_TRY(hipMemcpyAsync(c_cr + g.o_e0, m->creases, sizeof(Crease) * g.E0, hipMemcpyHostToDevice, stream), "upload");
    }
    if (!rings_host.empty())
        CC_TRY(hipMemcpyAsync(rings, rings_host.data(), sizeof(int32_t) * rings_host.size(), hipMemcpyHostToDevice,
                              stream), "upload cage rings");
    CC_TRY(hipStreamSynchronize(stream), "batch upload");
    return build_tables();
}

void ccb_Batch::free_all() {
    DeviceGuard g(device);
    drop_graphs();
    begin_timing();
    void *ptrs[] = {arena, adj_arena, tables};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
}

// A graph exec may still be running on any stream the caller used: wait for
// the device before destroying it.
void ccb_Batch::drop_graphs() {
    if (graphs.empty()) return;
    (void)hipDeviceSynchronize();
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
}

bool ccb_Batch::alloc_adjoint() {
    if (adj_arena) return true;
    Carve cv;
    (void)cv.take(sizeof(void *));   // offset 0: the address of the caller's finest gradient
    const size_t oA = cv.take(sizeof(float) * 3 * (size_t)tot_fin);
    const size_t oB = cv.take(sizeof(float) * 3 * (size_t)tot_adjB);
    const size_t oW = cv.take(sizeof(float) * (size_t)tot_adjW);
    const size_t oC = cv.take(sizeof(float) * 3 * (size_t)tot_v0);
    const size_t oP = cv.take(sizeof(VParam<float>) * (size_t)tot_adjVP);
    CC_TRY(hipMalloc(&adj_arena, cv.size), "hipMalloc(batch adjoint arena)");
    adjVP = reinterpret_cast<VParam<float> *>(adj_arena + oP);
    adjA = reinterpret_cast<float *>(adj_arena + oA);
    adjB = reinterpret_cast<float *>(adj_arena + oB);
    adjW = reinterpret_cast<float *>(adj_arena + oW);
    adjCage = reinterpret_cast<float *>(adj_arena + oC);
    return build_tables();
}

// Job tables of every launch, for the current tile options.  The tables are
// rewritten in place, so launches already queued on any stream (the batch's
// own or the caller's) must have finished first.
bool ccb_Batch::build_tables() {
    CC_TRY(hipDeviceSynchronize(), "batch tables: wait for queued launches");
    TableBuilder tb;
    erep_tables = erep_kept;
    auto off = [&](const void *p) -> int64_t { return static_cast<const char *>(p) - arena; };
    auto aoff = [&](const void *p) -> int64_t { return static_cast<const char *>(p) - adj_arena; };
    auto cage_offsets = [&](int i, auto &j) {
        const Cage &g = cages[i];
        j.c_he = off(c_he + g.o_he0);
        j.c_v2h = off(c_v2h + g.o_v0);
        j.c_e2h = off(c_e2h + g.o_e0);
        j.c_f2h = off(c_f2h + g.o_f0);
        j.c_pts = off(c_pts + g.o_v0);
        j.c_uvs = off(c_uvs + g.o_u0);
        j.c_cr = off(c_cr + g.o_e0);
        auto roff = [&](int64_t o) -> int64_t { return o < 0 ? -1 : off(rings + o); };
        j.c_vro = roff(g.o_vro); j.c_vr = roff(g.o_vr); j.c_vb = roff(g.o_vb);
        j.c_fro = roff(g.o_fro); j.c_fr = roff(g.o_fr);
        j.c_voo = roff(g.o_voo); j.c_vo = roff(g.o_vo);
        j.cV = g.V0; j.cU = g.U0; j.cH = g.H0; j.cE = g.E0; j.cF = g.F0;
    };
    {   // cage step
        std::vector<CageJob> jobs(N);
        std::vector<int64_t> bh(N), bf(N), be(N), bv(N), bz(N);
        for (int i = 0; i < N; ++i) {
            const Cage &g = cages[i];
            jobs[i].c = cage_view(i);
            jobs[i].he1 = he + g.o_he;
            jobs[i].pts1 = pts + g.o_pts;
            jobs[i].uv_mode = !kHasUv ? 0 : (g.U0 > 0 ? 1 : 2);
            bh[i] = blocks_of(g.H0);
            bf[i] = blocks_of(g.F0);
            be[i] = blocks_of(g.E0);
            bv[i] = blocks_of(g.V0);
            bz[i] = blocks_of(3 * (int64_t)g.V0);
        }
        tb.add(t_cage_he, jobs, bh);
        tb.add(t_cage_f, jobs, bf);
        tb.add(t_cage_e, jobs, be);
        tb.add(t_cage_v, jobs, bv);
        tb.add(t_cage_h3, jobs, bh);
        tb.add(t_cage_zero, jobs, bz);
    }
    {   // finest points of every cage into one contiguous array
        std::vector<SpanJob> jobs(N);
        std::vector<int64_t> b(N);
        for (int i = 0; i < N; ++i) {
            const Cage &g = cages[i];
            jobs[i] = {3 * (g.o_pts + g.n.offV[g.D]), 3 * g.o_fin, 3 * g.n.V[g.D]};
            b[i] = blocks_of(jobs[i].n);
        }
        tb.add(t_fin, jobs, b);
    }
    for (int d = 0; d < Dmax; ++d) {   // creases d -> d+1
        std::vector<CreaseJob> jobs;
        std::vector<int64_t> b;
        for (int i = 0; i < N; ++i) {
            const Cage &g = cages[i];
            if (g.D <= d) continue;
            CreaseJob j;
            j.in = d == 0 ? c_cr + g.o_e0 : cr + g.o_cr + g.n.offC[d];
            j.out = cr + g.o_cr + g.n.offC[d + 1];
            j.C = (int32_t)g.n.C[d];
            jobs.push_back(j);
            b.push_back(blocks_of(j.C));
        }
        tb.add(t_cr[d], jobs, b);
    }
    for (int d = 1; d < Dmax; ++d) {   // patch-local level steps d -> d+1
        bool fused_below = true;        // levels 1..d-1 run the fused kernel
        for (int l = 1; l < d; ++l) fused_below = fused_below && cfg[l].kind == 1;
        const bool fused_upto = fused_below && cfg[d].kind == 1;
        std::vector<LevelJob> jobs;
        std::vector<int64_t> bp, bd, bf, bv, br;
        for (int i = 0; i < N; ++i) {
            const Cage &g = cages[i];
            if (g.D <= d) continue;
            He *hb = he + g.o_he;
            P3 *pb = pts + g.o_pts;
            P3 *vp = pb + g.n.offV[d + 1];
            LevelJob j;
            j.he = off(hb + g.n.offH[d]);
            j.out_he = off(hb + g.n.offH[d + 1]);
            j.cr = off(cr + g.o_cr + g.n.offC[d]);
            j.old = off(pb + g.n.offV[d]);
            j.vp = off(vp);
            j.fp = off(vp + g.n.V[d]);
            j.ep = off(vp + g.n.V[d] + g.n.F[d]);
            j.defer_mask = off(defer_mask + g.o_def);
            j.owner_mask = cfg[d].kind == 1 && cfg[d].scan == 0 ? j.defer_mask : -1;
            j.rare = j.owner_mask >= 0 ? off(rare + g.o_rare) : -1;
            j.rare_cap = j.owner_mask >= 0 ? g.rare_cap : 0;
            // edge-representative bits: the forward pass runs the fused
            // kernel on every level up to d, so levels 1..d are written in
            // the same pass before they are read
            j.erep = g.o_erep >= 0 ? off(erep + g.o_erep) : -1;
            j.erep_write = j.erep >= 0 && d + 1 <= g.D - 2 && fused_upto ? 1 : 0;
            j.erep_read = j.erep >= 0 && d >= 2 && fused_below ? 1 : 0;
            j.H = (int32_t)g.n.H[d]; j.V = (int32_t)g.n.V[d]; j.F = (int32_t)g.n.F[d];
            j.E = (int32_t)g.n.E[d]; j.C = (int32_t)g.n.C[d];
            j.xcd_remap = remap;
            j.uv_refine = kHasUv && g.U0 > 0 ? 1 : 0;
            cage_offsets(i, j);
            j.t = g.tab;
            for (int k = 0; k <= kMaxDepth; ++k) j.lp[k] = (k >= 1 && k <= g.D) ? off(hb + g.n.offH[k]) : -1;
            jobs.push_back(j);
            const int64_t tiles = blocks_of(g.n.H[d], cfg[d].nt);
            bp.push_back((tiles + cfg[d].hpt - 1) / cfg[d].hpt);
            bd.push_back(blocks_of(blocks_of(blocks_of(g.n.H[d], 64), kDeferGroup), 4));
            bf.push_back(blocks_of(g.n.H[d], (int64_t)cfg[d].fb * cfg[d].ft));
            br.push_back(j.rare_cap > 0 ? blocks_of(j.rare_cap) : 0);
            bv.push_back(cfg[d].scan == 0 ? blocks_of(blocks_of(blocks_of(g.n.H[d], 64), kOwnerWords), 4)
                                          : blocks_of(g.n.H[d], 256 * cfg[d].scan));
        }
        tb.add(t_patch[d], jobs, bp);
        tb.add(t_def[d], jobs, bd);
        tb.add(t_fused[d], jobs, bf);
        tb.add(t_vert[d], jobs, bv);
        tb.add(t_rare[d], jobs, br);
    }
    if (adj_arena) {   // adjoint steps d+1 -> d, d = 0 the cage step
        for (int d = 0; d < Dmax; ++d) {
            // levels 1..d-1's edge-representative bits, when a forward pass
            // has written them
            const bool kept = d >= 2 && d - 1 <= erep_kept;
            std::vector<AdjJob> jobs;
            std::vector<int64_t> b, bq, br, be, bf, bv;
            for (int i = 0; i < N; ++i) {
                const Cage &g = cages[i];
                if (g.D <= d) continue;
                // gradient of level k lives in A (finest layout) when D-k is
                // even, else in B
                auto buf = [&](int k) { return ((g.D - k) & 1) ? adjB + 3 * g.o_adjB : adjA + 3 * g.o_fin; };
                He *hb = he + g.o_he;
                AdjJob j;
                float *G = buf(d + 1);
                if (d == 0) {
                    j.he = -1;
                    j.cr = -1;
                    j.H = g.H0;
                    j.C = g.E0;
                    j.V = g.V0;
                    j.gP = aoff(adjCage + 3 * g.o_v0);
                } else {
                    j.he = off(hb + g.n.offH[d]);
                    j.cr = off(cr + g.o_cr + g.n.offC[d]);
                    j.H = (int32_t)g.n.H[d];
                    j.C = (int32_t)g.n.C[d];
                    j.V = (int32_t)g.n.V[d];
                    j.gP = aoff(buf(d));
                }
                j.W = aoff(adjW + g.o_adjW);
                j.vp = d >= 1 ? aoff(adjVP + g.o_adjVP + g.n.offV[d]) : -1;
                j.erep = g.o_erep >= 0 && kept ? off(erep + g.o_erep) : -1;
                j.gin = gin_ok && d == g.D - 1 ? 1 : 0;
                if (j.gin) {   // byte offsets into the caller's gradient
                    const int64_t o = (int64_t)sizeof(float) * 3 * g.o_fin;
                    j.gV = o;
                    j.gF = o + (int64_t)sizeof(float) * 3 * g.n.V[d];
                    j.gE = o + (int64_t)sizeof(float) * 3 * (g.n.V[d] + g.n.F[d]);
                } else {
                    j.gV = aoff(G);
                    j.gF = aoff(G + 3 * g.n.V[d]);
                    j.gE = aoff(G + 3 * (g.n.V[d] + g.n.F[d]));
                }
                j.gFu = d >= 1 ? aoff(adjW + g.o_adjW + 4 * g.n.V[d] + 3 * g.n.E[d]) : -1;
                // the forward's rare-vertex list, free between passes
                j.rare = d >= 1 ? off(rare + g.o_rare) : -1;
                j.rare_cap = d >= 1 ? g.rare_cap : 0;
                cage_offsets(i, j);
                j.t = g.tab;
                for (int k = 0; k <= kMaxDepth; ++k) j.lp[k] = (k >= 1 && k <= g.D) ? off(hb + g.n.offH[k]) : -1;
                jobs.push_back(j);
                b.push_back(blocks_of(j.H));
                bq.push_back(blocks_of(j.H, 1024));
                br.push_back(j.rare_cap > 0 ? blocks_of(j.rare_cap) : 0);
                be.push_back(blocks_of(g.E0));
                bf.push_back(blocks_of(g.F0));
                bv.push_back(blocks_of(g.V0));
            }
            tb.add(t_adj[d], jobs, b);
            tb.add(t_adjq[d], jobs, bq);
            tb.add(t_adjr[d], jobs, br);
            if (d == 0) {
                tb.add(t_adjc_e, jobs, be);
                tb.add(t_adjc_f, jobs, bf);
                tb.add(t_adjc_v, jobs, bv);
            }
        }
    }
    if (tb.blob.size() > tables_cap) {
        if (tables) (void)hipFree(tables);
        tables = nullptr;
        tables_cap = 0;
        CC_TRY(hipMalloc(&tables, tb.blob.size()), "hipMalloc(batch tables)");
        tables_cap = tb.blob.size();
    }
    CC_TRY(hipMemcpy(tables, tb.blob.data(), tb.blob.size(), hipMemcpyHostToDevice), "upload batch tables");
    tb.resolve(tables);
    drop_graphs();
    return true;
}

void ccb_Batch::begin_timing() {
    for (auto &t : timed) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    timed.clear();
}

template <typename Fn>
void ccb_Batch::timed_launch(const char *name, int level, hipStream_t st, Fn &&fn) {
    Timed t{name, level, nullptr, nullptr};
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

// CatmullClark.h:3051-3082 per cage: creases of every level, cage halfedges
// and points, then per level one patch-local launch and one for the
// deferred vertices.
void ccb_Batch::record_forward(int32_t steps, hipStream_t st) {
    const int gather = (steps & CCD_GATHER) ? 1 : 0;
    const int creased = (steps & CCD_NOCREASES) ? 0 : 1;
    const int topo = (steps & CCD_HALFEDGES) ? 1 : 0;
    if (steps & CCD_CREASES)
        for (int d = 0; d < Dmax; ++d) timed_launch("creases", d, st, [&] { batch_creases(t_cr[d], st); });
    if (topo) timed_launch("cage_halfedges", 0, st, [&] { batch_cage_halfedges(t_cage_he, 1, st); });
    timed_launch("cage_points", 0, st, [&] {
        batch_cage_faces(t_cage_f, gather, st);
        batch_cage_edges(t_cage_e, gather, creased, st);
        batch_cage_vertices(t_cage_v, gather, creased, st);
    });
    for (int d = 1; d < Dmax; ++d) {
        if (cfg[d].kind == 2) {
            timed_launch("level_patch", d, st, [&] {
                batch_level_patch(t_patch[d], arena, d, gather, creased, topo, cfg[d].nt, cfg[d].hpt, remap, st);
            });
            timed_launch("level_deferred", d, st,
                         [&] { batch_level_deferred(t_def[d], arena, d, gather, creased, st); });
        } else {
            timed_launch("level_fused", d, st,
                         [&] {
                             batch_level_fused(t_fused[d], arena, d, gather, creased, topo, remap, cfg[d].fb,
                                               cfg[d].ft, st);
                         });
            timed_launch("level_vertices", d, st, [&] {
                batch_level_vertices(t_vert[d], arena, d, gather, creased, cfg[d].scan, remap, cfg[d].vw, st);
                if (cfg[d].scan == 0) batch_level_vertices_rare(t_rare[d], arena, d, gather, creased, st);
            });
        }
    }
}

// adjoint of record_forward's points, finest level first (gradients of the
// finest points in adjA, cage gradients into adjCage)
void ccb_Batch::record_adjoint(int flags, hipStream_t st) {
    const int creased = flags & 1, params = (flags >> 1) & 1;
    const int64_t nc = 3 * tot_v0;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<int64_t>((nc + 255) / 256, 65536)), dim3(256), 0, st, adjCage, nc);
    for (int d = Dmax - 1; d >= 1; --d)
        timed_launch("adj_level", d, st,
                     [&] { batch_adj_level(t_adjq[d], t_adj[d], t_adjr[d], arena, adj_arena, d, creased, params, st); });
    if (all_rings)
        timed_launch("adj_cage", 0, st, [&] {
            batch_adj_cage_gather(t_adj[0], t_adjc_e, t_adjc_f, t_adjc_v, arena, adj_arena, creased, st);
        });
    else
        timed_launch("adj_cage", 0, st, [&] { batch_adj_cage(t_adj[0], arena, adj_arena, creased, st); });
}

bool ccb_Batch::run(int32_t key, bool adjoint, int32_t steps, hipStream_t st) {
    if (adjoint && erep_kept > erep_tables && !build_tables()) return false;   // now with the bits
    if (!adjoint) {
        // the fused kernel on levels 1..p writes the bits of levels 1..p+1
        int p = 0;
        while (p + 1 < Dmax && cfg[p + 1].kind == 1) ++p;
        if (p >= 1) erep_kept = std::max(erep_kept, p + 1);
    }
    // a caller capturing its own stream gets the launches in its capture
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (st) CC_TRY(hipStreamIsCapturing(st, &cap), "hipStreamIsCapturing");
    if (!use_graph || timing || cap != hipStreamCaptureStatusNone) {
        if (timing) begin_timing();   // the timings are those of the last run
        if (adjoint) record_adjoint(steps, st);
        else record_forward(steps, st);
        CC_TRY(hipGetLastError(), "batch launch");
        return true;
    }
    auto it = graphs.find(key);
    if (it == graphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t exec = nullptr;
        CC_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        if (adjoint) record_adjoint(steps, stream);
        else record_forward(steps, stream);
        CC_TRY(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        CC_TRY(e, "hipGraphInstantiate");
        it = graphs.emplace(key, exec).first;
    }
    CC_TRY(hipGraphLaunch(it->second, st), "hipGraphLaunch");
    return true;
}

// Times candidate kernel configurations on this batch with HIP events and
// keeps, per level, the fastest, in three rounds:
//   1. the fused kernel's geometry: 256, 512 or 1024 threads per workgroup
//      times 1, 2 or 4 tiles per workgroup (masked vertex kernel);
//   2. from each level's pick: the masked vertex kernel compiled for 4 or 6
//      waves per SIMD, or the compacted vertex kernel testing 2, 4 or 8
//      halfedges per thread;
//   3. the patch-local kernel with tiles of 128, 256 or 512 halfedges.
// Each candidate runs the given steps three times (the first discarded); a
// level's time is that of its fused (or patch) and vertex kernels.
int32_t ccb_Batch::autotune(int32_t steps, hipStream_t st) {
    float best[kMaxDepth + 1];
    LevelCfg pick[kMaxDepth + 1];
    for (int d = 0; d <= kMaxDepth; ++d) { best[d] = 1e30f; pick[d] = cfg[d]; }
    const bool was = timing;
    timing = true;
    // level times of the candidate in cfg, false on error
    auto measure = [&](float (&ms)[kMaxDepth + 1]) -> bool {
        for (auto &x : ms) x = 0.0f;
        if (!build_tables()) return false;
        for (int rep = 0; rep < 3; ++rep) {
            begin_timing();
            record_forward(steps, st);
            if (hip_failed(hipStreamSynchronize(st), "batch autotune")) return false;
            if (rep == 0) continue;
            for (auto &x : timed) {
                if (strncmp(x.name, "level_", 6) != 0) continue;
                float v = 0.0f;
                (void)hipEventElapsedTime(&v, x.start, x.stop);
                ms[x.level] += v;
            }
        }
        return true;
    };
    auto keep = [&](const float (&ms)[kMaxDepth + 1]) {
        for (int d = 1; d < Dmax; ++d)
            if (ms[d] < best[d]) { best[d] = ms[d]; pick[d] = cfg[d]; }
    };
    float ms[kMaxDepth + 1];
    bool ok = true;
    for (int fb : {256, 512, 1024}) {
        for (int ft : {1, 2, 4}) {
            if (!ok) break;
            for (auto &c : cfg) { c = LevelCfg{}; c.fb = fb; c.ft = ft; }
            if ((ok = measure(ms))) keep(ms);
        }
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
    timing = was;
    for (int d = 0; d <= kMaxDepth; ++d) cfg[d] = pick[d];
    if (!ok) return last_error() ? last_error() : 1;
    return build_tables() ? 0 : last_error();
}

// ---- C ABI ---------------------------------------------------------------------
extern "C" {

ccb_Batch *ccb_Create(const cc_Mesh *const *cages, const int32_t *depths, int32_t count) {
    clear_error();
    if (!cages || !depths || count < 1) {
        set_error(12, "ccb_Create: empty batch");
        return nullptr;
    }
    for (int32_t i = 0; i < count; ++i) {
        const cc_Mesh *m = cages[i];
        const int32_t D = depths[i];
        if (!m || D < 1 || D > kMaxDepth || m->halfedgeCount < 3 || m->edgeCount < 1 || m->faceCount < 1 ||
            m->vertexCount < 1) {
            set_error(12, "ccb_Create: invalid cage or depth");
            return nullptr;
        }
        Counts n;
        n.init(m->vertexCount, m->halfedgeCount, m->edgeCount, m->faceCount);
        for (int d = 0; d <= D; ++d) {
            if (n.H[d] > INT32_MAX || n.V[d] + n.E[d] + n.F[d] > INT32_MAX || 2 * n.E[d] + n.H[d] > INT32_MAX) {
                set_error(12, "ccb_Create: IDs of a cage overflow int32 at its depth");
                return nullptr;
            }
        }
    }
    ccb_Batch *b = new ccb_Batch();
    if (!b->create(cages, depths, count)) {
        b->free_all();
        delete b;
        return nullptr;
    }
    return b;
}

void ccb_Release(ccb_Batch *b) {
    if (!b) return;
    b->free_all();
    delete b;
}

int64_t ccb_Count(const ccb_Batch *b, int32_t what, int32_t cage) {
    if (!b || cage < -1 || cage >= b->N) return -1;
    if (cage < 0) {
        switch (what) {
        case CCB_COUNT_HALFEDGES: return b->tot_he;
        case CCB_COUNT_CREASES: return b->tot_cr;
        case CCB_COUNT_POINTS: return b->tot_pts;
        case CCB_COUNT_CAGE_POINTS: return b->tot_v0;
        case CCB_COUNT_FINEST_POINTS: return b->tot_fin;
        case CCB_COUNT_DEPTH: return b->Dmax;
        case CCB_COUNT_CAGES: return b->N;
        default: return -1;
        }
    }
    const ccb_Batch::Cage &g = b->cages[cage];
    switch (what) {
    case CCB_COUNT_HALFEDGES: return g.n.cumH(g.D);
    case CCB_COUNT_CREASES: return g.n.cumC(g.D);
    case CCB_COUNT_POINTS: return g.n.cumV(g.D);
    case CCB_COUNT_CAGE_POINTS: return g.V0;
    case CCB_COUNT_FINEST_POINTS: return g.n.V[g.D];
    case CCB_COUNT_DEPTH: return g.D;
    case CCB_COUNT_CAGES: return 1;
    default: return -1;
    }
}

int64_t ccb_Offset(const ccb_Batch *b, int32_t what, int32_t cage) {
    if (!b || cage < 0 || cage >= b->N) return -1;
    const ccb_Batch::Cage &g = b->cages[cage];
    switch (what) {
    case CCB_COUNT_HALFEDGES: return g.o_he;
    case CCB_COUNT_CREASES: return g.o_cr;
    case CCB_COUNT_POINTS: return g.o_pts;
    case CCB_COUNT_CAGE_POINTS: return g.o_v0;
    case CCB_COUNT_FINEST_POINTS: return g.o_fin;
    default: return -1;
    }
}

void *ccb_DevicePointer(const ccb_Batch *b, int32_t which) {
    if (!b) return nullptr;
    switch (which) {
    case CCD_PTR_HALFEDGES: return b->he;
    case CCD_PTR_CREASES: return b->cr;
    case CCD_PTR_POINTS: return b->pts;
    case CCD_PTR_CAGE_POINTS: return b->c_pts;
    default: return nullptr;
    }
}

int32_t ccb_Refine(ccb_Batch *b, int32_t steps, void *stream) {
    clear_error();
    if (!b) return fail(11, "ccb_Refine: null batch");
    if (!(steps & CCD_POINTS)) return fail(15, "ccb_Refine: a batched refinement computes points; add CCD_POINTS");
    if (!(steps & CCD_HALFEDGES) && !b->topo_done) return fail(16, "ccb_Refine: halfedges not refined yet");
    const bool creased = !(steps & CCD_NOCREASES);
    if (creased && !(steps & CCD_CREASES) && !b->creases_done) return fail(16, "ccb_Refine: creases not refined yet");
    steps &= CCD_HALFEDGES | CCD_CREASES | CCD_POINTS | CCD_GATHER | CCD_NOCREASES;
    if (steps & CCD_HALFEDGES) steps |= CCD_UVS;
    DeviceGuard g(b->device);
    hipStream_t st = (hipStream_t)stream;
    if (!b->run(steps, false, steps, st)) return last_error();
    if (steps & CCD_HALFEDGES) b->topo_done = true;
    if (steps & CCD_CREASES) b->creases_done = true;
    return 0;
}

int32_t ccb_RefinePoints(ccb_Batch *b, const float *devCagePoints, int32_t steps, void *stream) {
    clear_error();
    if (!b) return fail(11, "ccb_RefinePoints: null batch");
    if (!b->topo_done) return fail(16, "ccb_RefinePoints: halfedges not refined yet");
    const bool creased = !(steps & CCD_NOCREASES);
    if (creased && !b->creases_done) return fail(16, "ccb_RefinePoints: creases not refined yet");
    DeviceGuard g(b->device);
    hipStream_t st = (hipStream_t)stream;
    if (devCagePoints &&
        hip_failed(hipMemcpyAsync(b->c_pts, devCagePoints, sizeof(P3) * (size_t)b->tot_v0, hipMemcpyDeviceToDevice, st),
                   "ccb_RefinePoints"))
        return last_error();
    const int32_t s = CCD_POINTS | (steps & (CCD_GATHER | CCD_NOCREASES));
    return b->run(s, false, s, st) ? 0 : last_error();
}

int32_t ccb_FinestPoints(ccb_Batch *b, float *devOut, void *stream) {
    clear_error();
    if (!b || !devOut) return fail(11, "ccb_FinestPoints: null argument");
    DeviceGuard g(b->device);
    hipStream_t st = (hipStream_t)stream;
    batch_copy(b->t_fin, reinterpret_cast<const float *>(b->pts), devOut, st);
    return hip_failed(hipGetLastError(), "ccb_FinestPoints") ? last_error() : 0;
}

int32_t ccb_RefinePointsAdjoint(ccb_Batch *b, const float *devGradFinest, float *devGradCage, int32_t steps,
                                void *stream) {
    clear_error();
    if (!b || !devGradFinest || !devGradCage) return fail(11, "ccb_RefinePointsAdjoint: null argument");
    if (!b->topo_done) return fail(16, "ccb_RefinePointsAdjoint: halfedges not refined yet");
    const int creased = (steps & CCD_NOCREASES) ? 0 : 1;
    if (creased && !b->creases_done) return fail(16, "ccb_RefinePointsAdjoint: creases not refined yet");
    DeviceGuard g(b->device);
    if (!b->alloc_adjoint()) return last_error();
    hipStream_t st = (hipStream_t)stream;
    if (b->gin_ok) {   // the finest steps read the caller's gradient
        hipLaunchKernelGGL(k_set_address, dim3(1), dim3(64), 0, st, reinterpret_cast<const void **>(b->adj_arena),
                           devGradFinest);
        if (hip_failed(hipGetLastError(), "ccb_RefinePointsAdjoint")) return last_error();
    } else if (hip_failed(hipMemcpyAsync(b->adjA, devGradFinest, sizeof(float) * 3 * (size_t)b->tot_fin,
                                         hipMemcpyDeviceToDevice, st), "ccb_RefinePointsAdjoint")) {
        return last_error();
    }
    // the parameter pass runs on the first call for these rules only
    const int params = b->adj_vp_key != creased ? 1 : 0;
    if (!b->run(-1 - creased - 2 * params, true, creased | (params << 1), st)) return last_error();
    b->adj_vp_key = creased;
    if (hip_failed(hipMemcpyAsync(devGradCage, b->adjCage, sizeof(float) * 3 * (size_t)b->tot_v0,
                                  hipMemcpyDeviceToDevice, st), "ccb_RefinePointsAdjoint"))
        return last_error();
    return 0;
}

int32_t ccb_Download(const ccb_Batch *b, int32_t cage, cc_Subd *dst, int32_t steps) {
    clear_error();
    if (!b || !dst || cage < 0 || cage >= b->N) return fail(11, "ccb_Download: bad argument");
    const ccb_Batch::Cage &g = b->cages[cage];
    if (dst->maxDepth != g.D) return fail(13, "ccb_Download: depth mismatch");
    DeviceGuard gd(b->device);
    // the refinement may have been queued on any stream
    if (hip_failed(hipDeviceSynchronize(), "ccb_Download")) return last_error();
    hipStream_t st = b->stream;
    if ((steps & (CCD_HALFEDGES | CCD_UVS)) &&
        hip_failed(hipMemcpyAsync(dst->halfedges, b->he + g.o_he, sizeof(He) * (size_t)g.n.cumH(g.D),
                                  hipMemcpyDeviceToHost, st), "ccb_Download"))
        return last_error();
    if ((steps & CCD_CREASES) &&
        hip_failed(hipMemcpyAsync(dst->creases, b->cr + g.o_cr, sizeof(Crease) * (size_t)g.n.cumC(g.D),
                                  hipMemcpyDeviceToHost, st), "ccb_Download"))
        return last_error();
    if ((steps & CCD_POINTS) &&
        hip_failed(hipMemcpyAsync(dst->vertexPoints, b->pts + g.o_pts, sizeof(P3) * (size_t)g.n.cumV(g.D),
                                  hipMemcpyDeviceToHost, st), "ccb_Download"))
        return last_error();
    return hip_failed(hipStreamSynchronize(st), "ccb_Download") ? last_error() : 0;
}

int32_t ccb_Synchronize(ccb_Batch *b) {
    clear_error();
    if (!b) return fail(11, "ccb_Synchronize: null batch");
    DeviceGuard g(b->device);
    return hip_failed(hipDeviceSynchronize(), "ccb_Synchronize") ? last_error() : 0;
}

int32_t ccb_SetOption(ccb_Batch *b, int32_t key, int32_t value) {
    if (!b) return 11;
    DeviceGuard g(b->device);
    switch (key) {
    case CCB_OPT_GRAPH: b->use_graph = value != 0; return 0;
    case CCD_OPT_XCD_REMAP: b->remap = value != 0; b->drop_graphs(); return 0;
    case CCD_OPT_FUSED:
        if (value != 1 && value != 2) return 14;
        for (auto &c : b->cfg) c.kind = value;
        return b->build_tables() ? 0 : last_error();
    case CCD_OPT_PATCH_TILE:
        if (value != 128 && value != 256 && value != 512) return 14;
        for (auto &c : b->cfg) c.nt = value;
        return b->build_tables() ? 0 : last_error();
    case CCD_OPT_PATCH_HPT:
        if (value < 1 || value > 64) return 14;
        for (auto &c : b->cfg) c.hpt = value;
        return b->build_tables() ? 0 : last_error();
    case CCD_OPT_SCAN:
        if (value != 0 && value != 2 && value != 4 && value != 8) return 14;
        for (auto &c : b->cfg) c.scan = value;
        return b->build_tables() ? 0 : last_error();
    case CCD_OPT_FUSED_BLOCK:
        if (value != 256 && value != 512 && value != 1024) return 14;
        for (auto &c : b->cfg) c.fb = value;
        return b->build_tables() ? 0 : last_error();
    case CCD_OPT_FUSED_TILES:
        if (value != 1 && value != 2 && value != 4) return 14;
        for (auto &c : b->cfg) c.ft = value;
        return b->build_tables() ? 0 : last_error();
    case CCD_OPT_VERTEX_WAVES:
        if (value != 4 && value != 6 && value != 8) return 14;
        for (auto &c : b->cfg) c.vw = value;
        b->drop_graphs();
        return 0;
    default: return 14;
    }
}

int32_t ccb_Autotune(ccb_Batch *b, int32_t steps, void *stream) {
    clear_error();
    if (!b) return fail(11, "ccb_Autotune: null batch");
    if (!(steps & CCD_POINTS)) return fail(15, "ccb_Autotune: add CCD_POINTS");
    DeviceGuard g(b->device);
    hipStream_t st = (hipStream_t)stream;
    steps &= CCD_HALFEDGES | CCD_CREASES | CCD_POINTS | CCD_GATHER | CCD_NOCREASES;
    if (!(steps & CCD_HALFEDGES) && !b->topo_done) return fail(16, "ccb_Autotune: halfedges not refined yet");
    const int32_t r = b->autotune(steps, st);
    if (r == 0 && (steps & CCD_HALFEDGES)) b->topo_done = true;
    if (r == 0 && (steps & CCD_CREASES)) b->creases_done = true;
    return r;
}

int32_t ccb_TileOption(const ccb_Batch *b, int32_t key, int32_t level) {
    if (!b || level < 0 || level > kMaxDepth) return -1;
    switch (key) {
    case CCD_OPT_FUSED: return b->cfg[level].kind;
    case CCD_OPT_PATCH_TILE: return b->cfg[level].nt;
    case CCD_OPT_PATCH_HPT: return b->cfg[level].hpt;
    case CCD_OPT_SCAN: return b->cfg[level].scan;
    case CCD_OPT_FUSED_BLOCK: return b->cfg[level].fb;
    case CCD_OPT_FUSED_TILES: return b->cfg[level].ft;
    case CCD_OPT_VERTEX_WAVES: return b->cfg[level].vw;
    default: return -1;
    }
}

int32_t ccb_SetTiming(ccb_Batch *b, int32_t enable) {
    if (!b) return 11;
    b->timing = enable != 0;
    if (b->timing) b->begin_timing();
    return 0;
}

int32_t ccb_TimingCount(const ccb_Batch *b) { return b ? (int32_t)b->timed.size() : 0; }

const char *ccb_TimingName(const ccb_Batch *b, int32_t i) {
    return (b && i >= 0 && i < (int32_t)b->timed.size()) ? b->timed[i].name : "";
}

int32_t ccb_TimingLevel(const ccb_Batch *b, int32_t i) {
    return (b && i >= 0 && i < (int32_t)b->timed.size()) ? b->timed[i].level : -1;
}

float ccb_TimingMs(const ccb_Batch *b, int32_t i) {
    if (!b || i < 0 || i >= (int32_t)b->timed.size()) return -1.0f;
    float ms = -1.0f;
    if (hipEventElapsedTime(&ms, b->timed[i].start, b->timed[i].stop) != hipSuccess) return -1.0f;
    return ms;
}

}  // extern "C"
