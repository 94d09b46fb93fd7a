// Host side of the C ABI: cage/subd counts, O(1) accessors, O(depth) maps,
// lifecycle and the .ccm file format.  Names, signatures and results follow
// CatmullClark.h; counts are computed exactly in 64-bit and returned as the
// reference's int32 values (identical wherever the reference is defined).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>

#include "../../include/CatmullClarkHIP.h"
#include "cc_common.h"
#include "cc_internal.h"

using namespace cch;

namespace {

Counts counts_of(const cc_Mesh *m) {
    Counts n;
    n.init(m->vertexCount, m->halfedgeCount, m->edgeCount, m->faceCount);
    return n;
}

bool depth_ok(int32_t d) { return d >= 0 && d <= kMaxDepth; }

}  // namespace

namespace cch {

bool ref_counts_fit_int32(const cc_Mesh *cage, int32_t D, const char **why) {
    // The reference evaluates its cumulative counts in int32 (CatmullClark.h
    // :830-936); a depth whose intermediate products overflow is undefined
    // there, so it is refused here.
    const int64_t lim = INT32_MAX;
    const Counts n = counts_of(cage);
    if (D < 1 || D > kMaxDepth) { *why = "depth must be in [1, 15]"; return false; }
    if (cage->vertexCount < 0 || cage->halfedgeCount < 3 || cage->edgeCount < 1 || cage->faceCount < 1) {
        *why = "invalid cage counts";
        return false;
    }
    const int64_t H1 = 4 * (int64_t)cage->halfedgeCount, E1 = 2 * (int64_t)cage->edgeCount + cage->halfedgeCount;
    const int64_t A = (1LL << D) - 1;
    if (H1 * ((1LL << (2 * D)) - 1) > lim) { *why = "cumulative halfedge count overflows int32"; return false; }
    if (A * (6 * E1 + A * H1 - H1) > lim) { *why = "cumulative edge count overflows int32"; return false; }
    if (n.cumV(D) > lim || n.cumC(D) > lim) { *why = "cumulative vertex/crease count overflows int32"; return false; }
    for (int d = 0; d <= D; ++d) {
        if (n.V[d] + n.E[d] + n.F[d] > lim || 2 * n.E[d] + n.H[d] > lim) {
            *why = "per-level IDs overflow int32";
            return false;
        }
    }
    return true;
}

}  // namespace cch

extern "C" {

// ---- cage counts -----------------------------------------------------------
int32_t ccm_FaceCount(const cc_Mesh *m) { return m->faceCount; }
int32_t ccm_EdgeCount(const cc_Mesh *m) { return m->edgeCount; }
int32_t ccm_HalfedgeCount(const cc_Mesh *m) { return m->halfedgeCount; }
int32_t ccm_CreaseCount(const cc_Mesh *m) { return m->edgeCount; }
int32_t ccm_VertexCount(const cc_Mesh *m) { return m->vertexCount; }
int32_t ccm_UvCount(const cc_Mesh *m) { return m->uvCount; }

int32_t ccm_HalfedgeCountAtDepth(const cc_Mesh *c, int32_t d) { return depth_ok(d) ? (int32_t)counts_of(c).H[d] : 0; }
int32_t ccm_CreaseCountAtDepth(const cc_Mesh *c, int32_t d) { return depth_ok(d) ? (int32_t)counts_of(c).C[d] : 0; }
int32_t ccm_FaceCountAtDepth(const cc_Mesh *c, int32_t d) { return depth_ok(d) ? (int32_t)counts_of(c).F[d] : 0; }
int32_t ccm_FaceCountAtDepth_Fast(const cc_Mesh *c, int32_t d) { return ccm_FaceCountAtDepth(c, d); }
int32_t ccm_EdgeCountAtDepth(const cc_Mesh *c, int32_t d) { return depth_ok(d) ? (int32_t)counts_of(c).E[d] : 0; }
int32_t ccm_EdgeCountAtDepth_Fast(const cc_Mesh *c, int32_t d) { return ccm_EdgeCountAtDepth(c, d); }
int32_t ccm_VertexCountAtDepth(const cc_Mesh *c, int32_t d) { return depth_ok(d) ? (int32_t)counts_of(c).V[d] : 0; }
int32_t ccm_VertexCountAtDepth_Fast(const cc_Mesh *c, int32_t d) { return ccm_VertexCountAtDepth(c, d); }

// ---- cage accessors (CatmullClark.h:592-757) --------------------------------
int32_t ccm_HalfedgeTwinID(const cc_Mesh *m, int32_t h) { return m->halfedges[h].twinID; }
int32_t ccm_HalfedgeNextID(const cc_Mesh *m, int32_t h) { return m->halfedges[h].nextID; }
int32_t ccm_HalfedgePrevID(const cc_Mesh *m, int32_t h) { return m->halfedges[h].prevID; }
int32_t ccm_HalfedgeFaceID(const cc_Mesh *m, int32_t h) { return m->halfedges[h].faceID; }
int32_t ccm_HalfedgeEdgeID(const cc_Mesh *m, int32_t h) { return m->halfedges[h].edgeID; }
int32_t ccm_HalfedgeVertexID(const cc_Mesh *m, int32_t h) { return m->halfedges[h].vertexID; }
int32_t ccm_HalfedgeUvID(const cc_Mesh *m, int32_t h) { return m->halfedges[h].uvID; }
float ccm_CreaseSharpness(const cc_Mesh *m, int32_t e) { return m->creases[e].sharpness; }
int32_t ccm_CreaseNextID(const cc_Mesh *m, int32_t e) { return m->creases[e].nextID; }
int32_t ccm_CreasePrevID(const cc_Mesh *m, int32_t e) { return m->creases[e].prevID; }
float ccm_HalfedgeSharpness(const cc_Mesh *m, int32_t h) { return ccm_CreaseSharpness(m, ccm_HalfedgeEdgeID(m, h)); }
cc_VertexPoint ccm_VertexPoint(const cc_Mesh *m, int32_t v) { return m->vertexPoints[v]; }
cc_VertexUv ccm_Uv(const cc_Mesh *m, int32_t t) { return m->uvs[t]; }
cc_VertexPoint ccm_HalfedgeVertexPoint(const cc_Mesh *m, int32_t h) { return ccm_VertexPoint(m, ccm_HalfedgeVertexID(m, h)); }
cc_VertexUv ccm_HalfedgeVertexUv(const cc_Mesh *m, int32_t h) { return ccm_Uv(m, ccm_HalfedgeUvID(m, h)); }
int32_t ccm_HalfedgeNextID_Quad(int32_t h) { return qnext(h); }
int32_t ccm_HalfedgePrevID_Quad(int32_t h) { return qprev(h); }
int32_t ccm_HalfedgeFaceID_Quad(int32_t h) { return h >> 2; }
int32_t ccm_VertexToHalfedgeID(const cc_Mesh *m, int32_t v) { return m->vertexToHalfedgeIDs[v]; }
int32_t ccm_EdgeToHalfedgeID(const cc_Mesh *m, int32_t e) { return m->edgeToHalfedgeIDs[e]; }
int32_t ccm_FaceToHalfedgeID(const cc_Mesh *m, int32_t f) { return m->faceToHalfedgeIDs[f]; }
int32_t ccm_FaceToHalfedgeID_Quad(int32_t f) { return f << 2; }
int32_t ccm_NextVertexHalfedgeID(const cc_Mesh *m, int32_t h) {
    const int32_t t = ccm_HalfedgeTwinID(m, h);
    return t >= 0 ? ccm_HalfedgeNextID(m, t) : -1;
}
int32_t ccm_PrevVertexHalfedgeID(const cc_Mesh *m, int32_t h) { return ccm_HalfedgeTwinID(m, ccm_HalfedgePrevID(m, h)); }

// ---- cage lifecycle (CatmullClark.h:763-808) --------------------------------
cc_Mesh *ccm_Create(int32_t V, int32_t U, int32_t H, int32_t E, int32_t F) {
    if (V < 0 || U < 0 || H < 0 || E < 0 || F < 0) {
        set_error(1, "ccm_Create: negative count");
        return NULL;
    }
    cc_Mesh *m = (cc_Mesh *)calloc(1, sizeof(cc_Mesh));
    if (!m) { set_error(2, "ccm_Create: out of host memory"); return NULL; }
    m->vertexCount = V; m->uvCount = U; m->halfedgeCount = H; m->edgeCount = E; m->faceCount = F;
    m->vertexToHalfedgeIDs = (int32_t *)malloc(sizeof(int32_t) * (size_t)V + 1);
    m->edgeToHalfedgeIDs = (int32_t *)malloc(sizeof(int32_t) * (size_t)E + 1);
    m->faceToHalfedgeIDs = (int32_t *)malloc(sizeof(int32_t) * (size_t)F + 1);
    m->halfedges = (cc_Halfedge *)malloc(sizeof(cc_Halfedge) * (size_t)H + 1);
    m->creases = (cc_Crease *)malloc(sizeof(cc_Crease) * (size_t)E + 1);
    m->vertexPoints = (cc_VertexPoint *)malloc(sizeof(cc_VertexPoint) * (size_t)V + 1);
    m->uvs = (cc_VertexUv *)malloc(sizeof(cc_VertexUv) * (size_t)U + 1);
    if (!m->vertexToHalfedgeIDs || !m->edgeToHalfedgeIDs || !m->faceToHalfedgeIDs || !m->halfedges ||
        !m->creases || !m->vertexPoints || !m->uvs) {
        ccm_Release(m);
        set_error(2, "ccm_Create: out of host memory");
        return NULL;
    }
    clear_error();
    return m;
}

void ccm_Release(cc_Mesh *m) {
    if (!m) return;
    free(m->vertexToHalfedgeIDs);
    free(m->faceToHalfedgeIDs);
    free(m->edgeToHalfedgeIDs);
    free(m->halfedges);
    free(m->creases);
    free(m->vertexPoints);
    free(m->uvs);
    free(m);
}

// ---- .ccm I/O (CatmullClark.h:3085-3269) -------------------------------------
// 32-byte header: the 8 bytes "cc_Mesh1", five int32 counts (vertex, uv,
// halfedge, edge, face) and 4 padding bytes; then the arrays v2h, e2h, f2h,
// points, uvs, creases (one per edge) and the 28-byte halfedges.
struct CcmHeader {
    char magic[8];
    int32_t vertexCount, uvCount, halfedgeCount, edgeCount, faceCount;
    int32_t pad;
};
static_assert(sizeof(CcmHeader) == 32, "ccm header");
static const char kMagic[8] = {'c', 'c', '_', 'M', 'e', 's', 'h', '1'};

cc_Mesh *ccm_Load(const char *filename) {
    FILE *f = fopen(filename, "rb");
    CcmHeader hdr;
    if (!f) {
        fprintf(stderr, "cc: fopen failed\n");
        set_error(3, "ccm_Load: fopen failed");
        return NULL;
    }
    if (fread(&hdr, sizeof(hdr), 1, f) != 1 || memcmp(hdr.magic, kMagic, 8) != 0) {
        fprintf(stderr, "cc: unsupported file\n");
        set_error(4, "ccm_Load: unsupported file");
        fclose(f);
        return NULL;
    }
    cc_Mesh *m = ccm_Create(hdr.vertexCount, hdr.uvCount, hdr.halfedgeCount, hdr.edgeCount, hdr.faceCount);
    if (!m) { fclose(f); return NULL; }
    const size_t V = hdr.vertexCount, U = hdr.uvCount, H = hdr.halfedgeCount, E = hdr.edgeCount, F = hdr.faceCount;
    const bool ok = fread(m->vertexToHalfedgeIDs, sizeof(int32_t), V, f) == V &&
                    fread(m->edgeToHalfedgeIDs, sizeof(int32_t), E, f) == E &&
                    fread(m->faceToHalfedgeIDs, sizeof(int32_t), F, f) == F &&
                    fread(m->vertexPoints, sizeof(cc_VertexPoint), V, f) == V &&
                    fread(m->uvs, sizeof(cc_VertexUv), U, f) == U &&
                    fread(m->creases, sizeof(cc_Crease), E, f) == E &&
                    fread(m->halfedges, sizeof(cc_Halfedge), H, f) == H;
    fclose(f);
    if (!ok) {
        fprintf(stderr, "cc: data reading failed\n");
        set_error(5, "ccm_Load: data reading failed");
        ccm_Release(m);
        return NULL;
    }
    clear_error();
    return m;
}

bool ccm_Save(const cc_Mesh *m, const char *filename) {
    CcmHeader hdr;
    memcpy(hdr.magic, kMagic, 8);
    hdr.vertexCount = m->vertexCount; hdr.uvCount = m->uvCount; hdr.halfedgeCount = m->halfedgeCount;
    hdr.edgeCount = m->edgeCount; hdr.faceCount = m->faceCount; hdr.pad = 0;
    FILE *f = fopen(filename, "wb");
    if (!f) {
        fprintf(stderr, "cc: fopen failed\n");
        set_error(3, "ccm_Save: fopen failed");
        return false;
    }
    const size_t V = m->vertexCount, U = m->uvCount, H = m->halfedgeCount, E = m->edgeCount, F = m->faceCount;
    const bool ok = fwrite(&hdr, sizeof(hdr), 1, f) == 1 &&
                    fwrite(m->vertexToHalfedgeIDs, sizeof(int32_t), V, f) == V &&
                    fwrite(m->edgeToHalfedgeIDs, sizeof(int32_t), E, f) == E &&
                    fwrite(m->faceToHalfedgeIDs, sizeof(int32_t), F, f) == F &&
                    fwrite(m->vertexPoints, sizeof(cc_VertexPoint), V, f) == V &&
                    fwrite(m->uvs, sizeof(cc_VertexUv), U, f) == U &&
                    fwrite(m->creases, sizeof(cc_Crease), E, f) == E &&
                    fwrite(m->halfedges, sizeof(cc_Halfedge), H, f) == H;
    if (fclose(f) != 0 || !ok) {
        fprintf(stderr, "cc: data dump failed\n");
        set_error(6, "ccm_Save: data dump failed");
        return false;
    }
    clear_error();
    return true;
}

// ---- subd counts (CatmullClark.h:811-946) ------------------------------------
int32_t ccs_CumulativeHalfedgeCountAtDepth(const cc_Mesh *c, int32_t d) { return depth_ok(d) ? (int32_t)counts_of(c).cumH(d) : 0; }
int32_t ccs_CumulativeFaceCountAtDepth(const cc_Mesh *c, int32_t d) { return ccs_CumulativeHalfedgeCountAtDepth(c, d) >> 2; }
int32_t ccs_CumulativeCreaseCountAtDepth(const cc_Mesh *c, int32_t d) { return depth_ok(d) ? (int32_t)counts_of(c).cumC(d) : 0; }
int32_t ccs_CumulativeVertexCountAtDepth(const cc_Mesh *c, int32_t d) { return depth_ok(d) ? (int32_t)counts_of(c).cumV(d) : 0; }
int32_t ccs_CumulativeEdgeCountAtDepth(const cc_Mesh *c, int32_t d) {
    if (!depth_ok(d)) return 0;
    const Counts n = counts_of(c);
    int64_t s = 0;
    for (int k = 1; k <= d; ++k) s += n.E[k];
    return (int32_t)s;
}
int32_t ccs_MaxDepth(const cc_Subd *s) { return s->maxDepth; }
int32_t ccs_CumulativeFaceCount(const cc_Subd *s) { return ccs_CumulativeFaceCountAtDepth(s->cage, s->maxDepth); }
int32_t ccs_CumulativeEdgeCount(const cc_Subd *s) { return ccs_CumulativeEdgeCountAtDepth(s->cage, s->maxDepth); }
int32_t ccs_CumulativeHalfedgeCount(const cc_Subd *s) { return ccs_CumulativeHalfedgeCountAtDepth(s->cage, s->maxDepth); }
int32_t ccs_CumulativeCreaseCount(const cc_Subd *s) { return ccs_CumulativeCreaseCountAtDepth(s->cage, s->maxDepth); }
int32_t ccs_CumulativeVertexCount(const cc_Subd *s) { return ccs_CumulativeVertexCountAtDepth(s->cage, s->maxDepth); }
// number of vertex points at the finest level (the reference declares it at
// CatmullClark.h:150 without defining it)
int32_t ccs_VertexCount(const cc_Subd *s) { return ccm_VertexCountAtDepth(s->cage, s->maxDepth); }

// ---- subd accessors (CatmullClark.h:995-1196) --------------------------------
static const cc_Halfedge_SemiRegular *subd_he(const cc_Subd *s, int32_t h, int32_t d) {
    return &s->halfedges[counts_of(s->cage).offH[d] + h];
}
static const cc_Crease *subd_cr(const cc_Subd *s, int32_t e, int32_t d) {
    return &s->creases[counts_of(s->cage).offC[d] + e];
}
float ccs_CreaseSharpness_Fast(const cc_Subd *s, int32_t e, int32_t d) { return subd_cr(s, e, d)->sharpness; }
float ccs_CreaseSharpness(const cc_Subd *s, int32_t e, int32_t d) {
    return e < ccm_CreaseCountAtDepth(s->cage, d) ? ccs_CreaseSharpness_Fast(s, e, d) : 0.0f;
}
int32_t ccs_CreaseNextID_Fast(const cc_Subd *s, int32_t e, int32_t d) { return subd_cr(s, e, d)->nextID; }
int32_t ccs_CreaseNextID(const cc_Subd *s, int32_t e, int32_t d) {
    return e < ccm_CreaseCountAtDepth(s->cage, d) ? ccs_CreaseNextID_Fast(s, e, d) : e;
}
int32_t ccs_CreasePrevID_Fast(const cc_Subd *s, int32_t e, int32_t d) { return subd_cr(s, e, d)->prevID; }
int32_t ccs_CreasePrevID(const cc_Subd *s, int32_t e, int32_t d) {
    return e < ccm_CreaseCountAtDepth(s->cage, d) ? ccs_CreasePrevID_Fast(s, e, d) : e;
}
int32_t ccs_HalfedgeVertexID(const cc_Subd *s, int32_t h, int32_t d) { return subd_he(s, h, d)->vertexID; }
int32_t ccs_HalfedgeTwinID(const cc_Subd *s, int32_t h, int32_t d) { return subd_he(s, h, d)->twinID; }
int32_t ccs_HalfedgeNextID(const cc_Subd *, int32_t h, int32_t) { return qnext(h); }
int32_t ccs_HalfedgePrevID(const cc_Subd *, int32_t h, int32_t) { return qprev(h); }
int32_t ccs_HalfedgeFaceID(const cc_Subd *, int32_t h, int32_t) { return h >> 2; }
int32_t ccs_HalfedgeEdgeID(const cc_Subd *s, int32_t h, int32_t d) { return subd_he(s, h, d)->edgeID; }
float ccs_HalfedgeSharpness(const cc_Subd *s, int32_t h, int32_t d) {
    return ccs_CreaseSharpness(s, ccs_HalfedgeEdgeID(s, h, d), d);
}
cc_VertexPoint ccs_VertexPoint(const cc_Subd *s, int32_t v, int32_t d) {
    return s->vertexPoints[counts_of(s->cage).offV[d] + v];
}
cc_VertexPoint ccs_HalfedgeVertexPoint(const cc_Subd *s, int32_t h, int32_t d) {
    return ccs_VertexPoint(s, ccs_HalfedgeVertexID(s, h, d), d);
}
#ifndef CC_DISABLE_UV   // absent from the reference's CC_DISABLE_UV build (CatmullClark.h:170-172)
cc_VertexUv ccs_HalfedgeVertexUv(const cc_Subd *s, int32_t h, int32_t d) {
    const UV uv = uv_decode(subd_he(s, h, d)->uvID);
    cc_VertexUv r;
    r.u = uv.u;
    r.v = uv.v;
    return r;
}
#endif
int32_t ccs_PrevVertexHalfedgeID(const cc_Subd *s, int32_t h, int32_t d) {
    return ccs_HalfedgeTwinID(s, ccs_HalfedgePrevID(s, h, d), d);
}
int32_t ccs_NextVertexHalfedgeID(const cc_Subd *s, int32_t h, int32_t d) {
    return ccs_HalfedgeNextID(s, ccs_HalfedgeTwinID(s, h, d), d);
}
int32_t ccs_FaceToHalfedgeID(const cc_Subd *, int32_t f, int32_t) { return f << 2; }

int32_t ccs_EdgeToHalfedgeID(const cc_Subd *s, int32_t e, int32_t d) {
    DCage c = host_cage_view(s->cage);
    LevelTab t = level_tab(counts_of(s->cage));
    return edge_rep(c, t, e, d);
}
int32_t ccs_VertexPointToHalfedgeID(const cc_Subd *s, int32_t v, int32_t d) {
    DCage c = host_cage_view(s->cage);
    LevelTab t = level_tab(counts_of(s->cage));
    return vertex_rep(c, t, v, d);
}
int32_t ccs_VertexToHalfedgeID(const cc_Subd *s, int32_t v, int32_t d) { return ccs_VertexPointToHalfedgeID(s, v, d); }

// Level `depth` of a refined subd (the cage for depth 0) as a Wavefront OBJ
// file in the text the reference's example writes (examples/subd_cpu.c:22-115,
// ExportToObj): "v" lines, "vt" lines (the cage's uvs at depth 0, one per
// halfedge above it), then "f" lines, 1-based, with per-halfedge uv indices
// above depth 0; "%f" formatting throughout.  Returns false when the file
// cannot be written.
bool ccs_ExportToObj(const cc_Subd *subd, int32_t depth, const char *filename) {
    clear_error();
    if (!subd || !subd->cage || !filename || depth < 0 || depth > subd->maxDepth) {
        set_error(11, "ccs_ExportToObj: null argument or depth out of range");
        return false;
    }
    FILE *pf = fopen(filename, "w");
    if (!pf) {
        set_error(3, "ccs_ExportToObj: fopen failed");
        fprintf(stderr, "cc: ccs_ExportToObj: fopen failed\n");
        return false;
    }
    const cc_Mesh *cage = subd->cage;
    const int32_t vertexCount = ccm_VertexCountAtDepth(cage, depth);
    const int32_t faceCount = ccm_FaceCountAtDepth(cage, depth);
    fprintf(pf, "# Vertices\n");
    if (depth == 0) {
        for (int32_t v = 0; v < vertexCount; ++v) {
            const cc_VertexPoint p = ccm_VertexPoint(cage, v);
            fprintf(pf, "v %f %f %f\n", p.x, p.y, p.z);
        }
        for (int32_t t = 0; t < ccm_UvCount(cage); ++t) {
            const cc_VertexUv uv = ccm_Uv(cage, t);
            fprintf(pf, "vt %f %f\n", uv.u, uv.v);
        }
    } else {
        for (int32_t v = 0; v < vertexCount; ++v) {
            const cc_VertexPoint p = ccs_VertexPoint(subd, v, depth);
            fprintf(pf, "v %f %f %f\n", p.x, p.y, p.z);
        }
#ifndef CC_DISABLE_UV
        const int32_t halfedgeCount = ccm_HalfedgeCountAtDepth(cage, depth);
        for (int32_t h = 0; h < halfedgeCount; ++h) {
            const cc_VertexUv uv = ccs_HalfedgeVertexUv(subd, h, depth);
            fprintf(pf, "vt %f %f\n", uv.u, uv.v);
        }
#endif
    }
    fprintf(pf, "\n");
    fprintf(pf, "# Topology\n");
    if (depth == 0) {
        for (int32_t f = 0; f < faceCount; ++f) {
            const int32_t h0 = ccm_FaceToHalfedgeID(cage, f);
            fprintf(pf, "f %i/%i", ccm_HalfedgeVertexID(cage, h0) + 1, ccm_HalfedgeUvID(cage, h0) + 1);
            int32_t n = 1;
            for (int32_t h = ccm_HalfedgeNextID(cage, h0); h != h0 && n < kWalkCap; h = ccm_HalfedgeNextID(cage, h), ++n)
                fprintf(pf, " %i/%i", ccm_HalfedgeVertexID(cage, h) + 1, ccm_HalfedgeUvID(cage, h) + 1);
            fprintf(pf, "\n");
        }
    } else {
        for (int32_t f = 0; f < faceCount; ++f) {
            const int32_t h = 4 * f;
#ifndef CC_DISABLE_UV
            fprintf(pf, "f %i/%i %i/%i %i/%i %i/%i\n", ccs_HalfedgeVertexID(subd, h, depth) + 1, h + 1,
                    ccs_HalfedgeVertexID(subd, h + 1, depth) + 1, h + 2, ccs_HalfedgeVertexID(subd, h + 2, depth) + 1,
                    h + 3, ccs_HalfedgeVertexID(subd, h + 3, depth) + 1, h + 4);
#else
            fprintf(pf, "f %i %i %i %i\n", ccs_HalfedgeVertexID(subd, h, depth) + 1,
                    ccs_HalfedgeVertexID(subd, h + 1, depth) + 1, ccs_HalfedgeVertexID(subd, h + 2, depth) + 1,
                    ccs_HalfedgeVertexID(subd, h + 3, depth) + 1);
#endif
        }
        fprintf(pf, "\n");
    }
    const bool ok = !ferror(pf);
    if (fclose(pf) != 0 || !ok) {
        set_error(6, "ccs_ExportToObj: write failed");
        return false;
    }
    return true;
}

}  // extern "C"
