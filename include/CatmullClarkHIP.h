/*
 * CatmullClarkHIP.h -- MI355X (gfx950) drop-in for the uniform Catmull-Clark
 * refinement path of CatmullClark.h (jdupuy/HalfedgeCatmullClark).
 *
 * Link against libcchip.so.  The types below are byte-identical to the
 * reference's (CatmullClark.h:17-71, :136-142) so existing callers compile
 * unchanged; every ccm_* / ccs_* function keeps the reference's name,
 * signature and meaning.  The ccs_Refine* functions run on the GPU and keep
 * the reference's host semantics: they read the cage from host memory, write
 * every level 1..maxDepth into the cc_Subd's host arrays and return when the
 * results are there.  Failures never abort: the call logs to stderr, leaves
 * the outputs unmodified where it can, and records a code for
 * ccs_LastError() on the calling thread.
 *
 * Extensions (no counterpart in the reference) are marked as such:
 *   - ccs_LastError / ccs_LastErrorString (error channel);
 *   - ccs_RefineCage* (the cage step 0 -> 1 on its own);
 *   - ccs_RefineVertexPoints_Adjoint (transpose of the point refinement);
 *   - ccd_* (device-resident refinement for callers that keep data in HBM);
 *   - ccb_* (batched refinement of many cages).
 */
#ifndef CC_INCLUDE_CC_HIP_H
#define CC_INCLUDE_CC_HIP_H

#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- data layouts (CatmullClark.h:17-71, :136-142) ---------------------- */
typedef union {
    struct { float x, y, z; };
    float array[3];
} cc_VertexPoint;

typedef union {
    struct { float u, v; };
    float array[2];
} cc_VertexUv;

typedef struct {
    int32_t nextID;
    int32_t prevID;
    float sharpness;
} cc_Crease;

typedef struct {
    int32_t twinID;
    int32_t nextID;
    int32_t prevID;
    int32_t faceID;
    int32_t edgeID;
    int32_t vertexID;
    int32_t uvID;
} cc_Halfedge;

/* 16 bytes; 12 without uvID when CC_DISABLE_UV is defined, as in the
 * reference (CatmullClark.h:47-55).  libcchip_nouv.so is the library built
 * that way: link it, and define CC_DISABLE_UV before including this header,
 * to keep 12-byte level records (no uv refinement). */
typedef struct {
    int32_t twinID;
    int32_t edgeID;
    int32_t vertexID;
#ifndef CC_DISABLE_UV
    int32_t uvID;
#endif
} cc_Halfedge_SemiRegular;

typedef struct {
    int32_t vertexCount;
    int32_t uvCount;
    int32_t halfedgeCount;
    int32_t edgeCount;
    int32_t faceCount;
    int32_t *vertexToHalfedgeIDs;
    int32_t *edgeToHalfedgeIDs;
    int32_t *faceToHalfedgeIDs;
    cc_VertexPoint *vertexPoints;
    cc_VertexUv *uvs;
    cc_Halfedge *halfedges;
    cc_Crease *creases;
} cc_Mesh;

typedef struct {
    const cc_Mesh *cage;
    cc_VertexPoint *vertexPoints;
    cc_Halfedge_SemiRegular *halfedges;
    cc_Crease *creases;
    int32_t maxDepth;
} cc_Subd;

/* ---- cage lifecycle and I/O (CatmullClark.h:74-83) ---------------------- */
cc_Mesh *ccm_Load(const char *filename);                       /* :3186 */
cc_Mesh *ccm_Create(int32_t vertexCount, int32_t uvCount, int32_t halfedgeCount,
                    int32_t edgeCount, int32_t faceCount);     /* :763  */
void ccm_Release(cc_Mesh *mesh);                               /* :798  */
bool ccm_Save(const cc_Mesh *mesh, const char *filename);      /* :3227 */

/* ---- cage counts (CatmullClark.h:86-101) --------------------------------- */
int32_t ccm_FaceCount(const cc_Mesh *mesh);
int32_t ccm_EdgeCount(const cc_Mesh *mesh);
int32_t ccm_HalfedgeCount(const cc_Mesh *mesh);
int32_t ccm_CreaseCount(const cc_Mesh *mesh);
int32_t ccm_VertexCount(const cc_Mesh *mesh);
int32_t ccm_UvCount(const cc_Mesh *mesh);
int32_t ccm_HalfedgeCountAtDepth(const cc_Mesh *cage, int32_t depth);      /* :523 */
int32_t ccm_CreaseCountAtDepth(const cc_Mesh *cage, int32_t depth);        /* :540 */
int32_t ccm_FaceCountAtDepth(const cc_Mesh *cage, int32_t depth);          /* :473 */
int32_t ccm_FaceCountAtDepth_Fast(const cc_Mesh *cage, int32_t depth);     /* :465 */
int32_t ccm_EdgeCountAtDepth(const cc_Mesh *cage, int32_t depth);          /* :504 */
int32_t ccm_EdgeCountAtDepth_Fast(const cc_Mesh *cage, int32_t depth);     /* :494 */
int32_t ccm_VertexCountAtDepth(const cc_Mesh *cage, int32_t depth);        /* :578 */
int32_t ccm_VertexCountAtDepth_Fast(const cc_Mesh *cage, int32_t depth);   /* :563 */

/* ---- cage accessors (CatmullClark.h:104-131) ----------------------------- */
int32_t ccm_HalfedgeTwinID(const cc_Mesh *mesh, int32_t halfedgeID);
int32_t ccm_HalfedgeNextID(const cc_Mesh *mesh, int32_t halfedgeID);
int32_t ccm_HalfedgePrevID(const cc_Mesh *mesh, int32_t halfedgeID);
int32_t ccm_HalfedgeFaceID(const cc_Mesh *mesh, int32_t halfedgeID);
int32_t ccm_HalfedgeEdgeID(const cc_Mesh *mesh, int32_t halfedgeID);
int32_t ccm_HalfedgeVertexID(const cc_Mesh *mesh, int32_t halfedgeID);
int32_t ccm_HalfedgeUvID(const cc_Mesh *mesh, int32_t halfedgeID);
float ccm_HalfedgeSharpness(const cc_Mesh *mesh, int32_t halfedgeID);
cc_VertexPoint ccm_HalfedgeVertexPoint(const cc_Mesh *mesh, int32_t halfedgeID);
cc_VertexUv ccm_HalfedgeVertexUv(const cc_Mesh *mesh, int32_t halfedgeID);
int32_t ccm_CreaseNextID(const cc_Mesh *mesh, int32_t edgeID);
int32_t ccm_CreasePrevID(const cc_Mesh *mesh, int32_t edgeID);
float ccm_CreaseSharpness(const cc_Mesh *mesh, int32_t edgeID);
cc_VertexPoint ccm_VertexPoint(const cc_Mesh *mesh, int32_t vertexID);
cc_VertexUv ccm_Uv(const cc_Mesh *mesh, int32_t uvID);
int32_t ccm_HalfedgeNextID_Quad(int32_t halfedgeID);
int32_t ccm_HalfedgePrevID_Quad(int32_t halfedgeID);
int32_t ccm_HalfedgeFaceID_Quad(int32_t halfedgeID);
int32_t ccm_VertexToHalfedgeID(const cc_Mesh *mesh, int32_t vertexID);
int32_t ccm_EdgeToHalfedgeID(const cc_Mesh *mesh, int32_t edgeID);
int32_t ccm_FaceToHalfedgeID(const cc_Mesh *mesh, int32_t faceID);
int32_t ccm_FaceToHalfedgeID_Quad(int32_t faceID);
int32_t ccm_NextVertexHalfedgeID(const cc_Mesh *mesh, int32_t halfedgeID);
int32_t ccm_PrevVertexHalfedgeID(const cc_Mesh *mesh, int32_t halfedgeID);

/* ---- subd lifecycle and counts (CatmullClark.h:145-160) ------------------ */
/* ccs_Create refuses (returns NULL, ccs_LastError() != 0) a depth below 1 or
 * one whose cumulative counts overflow the reference's int32 formulas.  It
 * allocates the three level arrays page-locked when the HIP runtime allows
 * it (else with malloc), so refinements copy back at the DMA rate;
 * ccs_Release frees either kind. */
cc_Subd *ccs_Create(const cc_Mesh *cage, int32_t maxDepth);                 /* :953 */
void ccs_Release(cc_Subd *subd);                                            /* :977 */
int32_t ccs_MaxDepth(const cc_Subd *subd);
int32_t ccs_VertexCount(const cc_Subd *subd);
int32_t ccs_CumulativeFaceCount(const cc_Subd *subd);
int32_t ccs_CumulativeEdgeCount(const cc_Subd *subd);
int32_t ccs_CumulativeCreaseCount(const cc_Subd *subd);
int32_t ccs_CumulativeVertexCount(const cc_Subd *subd);
int32_t ccs_CumulativeHalfedgeCount(const cc_Subd *subd);
int32_t ccs_CumulativeHalfedgeCountAtDepth(const cc_Mesh *cage, int32_t depth); /* :861 */
int32_t ccs_CumulativeVertexCountAtDepth(const cc_Mesh *cage, int32_t depth);   /* :916 */
int32_t ccs_CumulativeFaceCountAtDepth(const cc_Mesh *cage, int32_t depth);     /* :815 */
int32_t ccs_CumulativeEdgeCountAtDepth(const cc_Mesh *cage, int32_t depth);     /* :830 */
int32_t ccs_CumulativeCreaseCountAtDepth(const cc_Mesh *cage, int32_t depth);   /* :890 */

/* ---- subd accessors (CatmullClark.h:163-195) ----------------------------- */
int32_t ccs_HalfedgeTwinID(const cc_Subd *subd, int32_t halfedgeID, int32_t depth);
int32_t ccs_HalfedgeNextID(const cc_Subd *subd, int32_t halfedgeID, int32_t depth);
int32_t ccs_HalfedgePrevID(const cc_Subd *subd, int32_t halfedgeID, int32_t depth);
int32_t ccs_HalfedgeFaceID(const cc_Subd *subd, int32_t halfedgeID, int32_t depth);
int32_t ccs_HalfedgeEdgeID(const cc_Subd *subd, int32_t halfedgeID, int32_t depth);
int32_t ccs_HalfedgeVertexID(const cc_Subd *subd, int32_t halfedgeID, int32_t depth);
cc_VertexPoint ccs_HalfedgeVertexPoint(const cc_Subd *subd, int32_t halfedgeID, int32_t depth);
#ifndef CC_DISABLE_UV
cc_VertexUv ccs_HalfedgeVertexUv(const cc_Subd *subd, int32_t halfedgeID, int32_t depth);
#endif
float ccs_HalfedgeSharpness(const cc_Subd *subd, int32_t halfedgeID, int32_t depth);
int32_t ccs_CreaseNextID_Fast(const cc_Subd *subd, int32_t edgeID, int32_t depth);
int32_t ccs_CreaseNextID(const cc_Subd *subd, int32_t edgeID, int32_t depth);
int32_t ccs_CreasePrevID_Fast(const cc_Subd *subd, int32_t edgeID, int32_t depth);
int32_t ccs_CreasePrevID(const cc_Subd *subd, int32_t edgeID, int32_t depth);
float ccs_CreaseSharpness_Fast(const cc_Subd *subd, int32_t edgeID, int32_t depth);
float ccs_CreaseSharpness(const cc_Subd *subd, int32_t edgeID, int32_t depth);
cc_VertexPoint ccs_VertexPoint(const cc_Subd *subd, int32_t vertexID, int32_t depth);
int32_t ccs_NextVertexHalfedgeID(const cc_Subd *subd, int32_t halfedgeID, int32_t depth);
int32_t ccs_PrevVertexHalfedgeID(const cc_Subd *subd, int32_t halfedgeID, int32_t depth);
/* Declared as ccs_VertexToHalfedgeID (:187) but defined as
 * ccs_VertexPointToHalfedgeID (:1336) in the reference; both are exported. */
int32_t ccs_VertexToHalfedgeID(const cc_Subd *subd, int32_t vertexID, int32_t depth);
int32_t ccs_VertexPointToHalfedgeID(const cc_Subd *subd, int32_t vertexID, int32_t depth);
int32_t ccs_EdgeToHalfedgeID(const cc_Subd *mesh, int32_t edgeID, int32_t depth);   /* :1230 */
int32_t ccs_FaceToHalfedgeID(const cc_Subd *mesh, int32_t faceID, int32_t depth);   /* :1190 */

/* ---- refinement (CatmullClark.h:197-212), GPU-backed --------------------- */
void ccs_Refine_Gather(cc_Subd *subd);                          /* :3066 */
void ccs_Refine_Scatter(cc_Subd *subd);                         /* :3060 */
void ccs_RefineVertexPoints_Gather(cc_Subd *subd);              /* :2659 */
void ccs_RefineVertexPoints_Scatter(cc_Subd *subd);             /* :2631 */
void ccs_RefineHalfedges(cc_Subd *subd);                        /* :2803 */
void ccs_RefineCreases(cc_Subd *subd);                          /* :3033 */
#ifndef CC_DISABLE_UV
void ccs_RefineVertexUvs(cc_Subd *subd);                        /* :2928 */
#endif
void ccs_Refine_NoCreases_Gather(cc_Subd *subd);                /* :3078 */
void ccs_Refine_NoCreases_Scatter(cc_Subd *subd);               /* :3072 */
void ccs_RefineVertexPoints_NoCreases_Gather(cc_Subd *subd);    /* :2672 */
void ccs_RefineVertexPoints_NoCreases_Scatter(cc_Subd *subd);   /* :2645 */

/* ---- extension: error channel ------------------------------------------- */
/* 0 when the last ccs_/ccd_/ccb_ call on this thread succeeded. */
int32_t ccs_LastError(void);
const char *ccs_LastErrorString(void);

/* ---- extension: the cage step (level 0 -> 1) alone ----------------------
 * The reference keeps these static (ccs__RefineCageHalfedges :2693,
 * ccs__RefineCageCreases :2950, ccs__RefineCageVertexUvs :2825, the
 * ccs__Cage*Points_* kernels :1408-1991).  Each writes level 1 only. */
void ccs_RefineCageHalfedges(cc_Subd *subd);
void ccs_RefineCageCreases(cc_Subd *subd);
#ifndef CC_DISABLE_UV
void ccs_RefineCageVertexUvs(cc_Subd *subd);
#endif
void ccs_RefineCageVertexPoints_Scatter(cc_Subd *subd);
void ccs_RefineCageVertexPoints_Gather(cc_Subd *subd);

/* ---- extension: adjoint of the point refinement --------------------------
 * Refined points are linear in the cage points at fixed topology and
 * sharpness.  Given dL/d(points at level maxDepth) (V_D records) this writes
 * dL/d(cage points) (V_0 records), the transpose of ccs_RefineVertexPoints_*
 * with creases.  Topology and creases must have been refined already. */
void ccs_RefineVertexPoints_Adjoint(const cc_Subd *subd,
                                    const cc_VertexPoint *dL_dFinest,
                                    cc_VertexPoint *dL_dCage);
/* the same in double precision (V_D x 3 and V_0 x 3 doubles), for gradient
 * checks */
void ccs_RefineVertexPoints_Adjoint_f64(const cc_Subd *subd,
                                        const double *dL_dFinest,
                                        double *dL_dCage);

/* ---- extension: the cage builder on the GPU ------------------------------
 * The reference builds cages in its converter (examples/obj_to_ccm.c:88-450,
 * LoadObj): twins by a bitonic sort and binary search, edge IDs by a prefix
 * sum, the vertex -> halfedge choice, crease tags, boundary sharpness 16 and
 * crease neighbours.  The same steps run here as GPU kernels (the twin sort
 * with the converter's own compare-exchange network), so the result is the
 * converter's byte for byte.  Inputs are host arrays: faceOffsets holds
 * faceCount + 1 offsets (from 0) into faceVertices (0-based vertex
 * indices); faceUvs (per corner, or NULL for uvID 0) indexes uvs; crease
 * tags give two vertices (creaseTagVertices[2k], [2k+1], 0-based as the
 * converter reads "t crease 2/1/0 v0 v1 s") and a sharpness each; boundary
 * edges get sharpness 16 when sharpenBoundaries is nonzero.  The cage comes
 * back in host memory (ccm_Release), NULL on failure (ccs_LastError). */
cc_Mesh *ccm_CreateFromFaces(int32_t vertexCount, const cc_VertexPoint *vertexPoints, int32_t faceCount,
                             const int32_t *faceOffsets, const int32_t *faceVertices, int32_t uvCount,
                             const cc_VertexUv *uvs, const int32_t *faceUvs, int32_t creaseTagCount,
                             const int32_t *creaseTagVertices, const float *creaseTagSharpness,
                             int32_t sharpenBoundaries);
/* An OBJ file parsed as the converter parses it (v, vt, f in its four vertex
 * formats, "t crease 2/1/0" tags) and built by ccm_CreateFromFaces. */
cc_Mesh *ccm_LoadObj(const char *filename);

/* ---- extension: normals and OBJ export ----------------------------------
 * The step after refinement.  Normals of every halfedge of level maxDepth
 * (ccm_HalfedgeCountAtDepth(cage, maxDepth) records), as the reference's
 * shaders define them (glsl/CatmullClark_Scatter.glsl:931-960,
 * ccs_HalfedgeNormal_Fast / ccs_HalfedgeNormal): the corner normal
 * normalize(cross(P(next) - P(h), P(prev) - P(h))), or the normalized sum
 * of the corner normals around h's vertex (the corner normal at a boundary).
 * Computed on the GPU from the subd's host arrays of level maxDepth. */
void ccs_HalfedgeNormals(const cc_Subd *subd, cc_VertexPoint *normals);
void ccs_HalfedgeNormals_Fast(const cc_Subd *subd, cc_VertexPoint *normals);
/* Level `depth` (0: the cage) as an OBJ file, in the text the reference's
 * example writes (examples/subd_cpu.c:22-115, ExportToObj); false when the
 * file cannot be written. */
bool ccs_ExportToObj(const cc_Subd *subd, int32_t depth, const char *filename);

/* ---- extension: device-resident refinement ------------------------------
 * A ccd_Subd owns device copies of a cage and of levels 1..maxDepth laid out
 * exactly as cc_Subd's arrays.  Calls are asynchronous on `stream` (a
 * hipStream_t; NULL is the HIP null stream, which is also torch's default
 * stream).  ccd_Download and ccd_UploadCagePoints synchronize the device
 * first.  Return 0 on success, a nonzero code otherwise (also readable via
 * ccs_LastError). */
typedef struct ccd_Subd ccd_Subd;

enum {
    CCD_HALFEDGES = 1,    /* twin/edge/vertex IDs of levels 1..D           */
    CCD_CREASES = 2,      /* crease links and sharpness                    */
    CCD_UVS = 4,          /* uvIDs (only when the cage has uvs)            */
    CCD_POINTS = 8,       /* vertex points                                 */
    CCD_GATHER = 16,      /* gather point rules instead of scatter rules   */
    CCD_NOCREASES = 32,   /* plain Catmull-Clark rules (ignore sharpness)  */
    CCD_FP64 = 64,        /* ccd_RefinePointsAdjoint: gradients are doubles */
    CCD_ALL = 15
};

enum { CCD_PTR_HALFEDGES = 0, CCD_PTR_CREASES = 1, CCD_PTR_POINTS = 2, CCD_PTR_CAGE_POINTS = 3 };

ccd_Subd *ccd_Create(const cc_Mesh *cage, int32_t maxDepth);
void ccd_Release(ccd_Subd *subd);
int32_t ccd_MaxDepth(const ccd_Subd *subd);
/* exact count of `what` (0 H, 1 C, 2 F, 3 E, 4 V at `depth`; 5 Hcum, 6 Ccum,
 * 7 Vcum over levels 1..depth) */
int64_t ccd_Count(const ccd_Subd *subd, int32_t what, int32_t depth);
void *ccd_DevicePointer(const ccd_Subd *subd, int32_t which);
int32_t ccd_UploadCagePoints(ccd_Subd *subd, const cc_VertexPoint *hostPoints);
/* refines on the device; the launch sequence of each steps value is captured
 * as a hipGraph on first use and replayed on `stream` (CCB_OPT_GRAPH 0:
 * launched one by one; so are runs with timing on) */
int32_t ccd_Refine(ccd_Subd *subd, int32_t steps, void *stream);
/* points only, from cage points already on the device (or the uploaded ones
 * when devCagePoints is NULL) into devPoints (or the ccd_Subd's own buffer
 * when NULL); devPoints holds levels 1..D in cumulative order. */
int32_t ccd_RefinePoints(ccd_Subd *subd, const float *devCagePoints, float *devPoints,
                         int32_t steps, void *stream);
/* adjoint on the device: devGradFinest holds V_D x 3 floats, devGradCage
 * receives V_0 x 3 floats (doubles for both with CCD_FP64).  The transpose
 * of the scatter rules, creased unless CCD_NOCREASES; topology and creases
 * must be refined already.  devGradFinest is not modified. */
int32_t ccd_RefinePointsAdjoint(ccd_Subd *subd, const float *devGradFinest, float *devGradCage,
                                int32_t steps, void *stream);
int32_t ccd_Download(const ccd_Subd *subd, cc_Subd *dst, int32_t steps);
int32_t ccd_Synchronize(ccd_Subd *subd);
/* halfedge normals of level maxDepth into devNormals (H_D x 3 floats) from
 * the device topology and points (refined already): vertex normals when
 * smooth != 0, else corner normals (see ccs_HalfedgeNormals) */
int32_t ccd_HalfedgeNormals(ccd_Subd *subd, float *devNormals, int32_t smooth, void *stream);
/* kernel-selection options (all have working defaults; setting any drops
 * ccd_Refine's captured graphs):
 *   CCB_OPT_GRAPH      1: ccd_Refine replays captured graphs (default)
 *   CCD_OPT_FUSED      1: per level one fused kernel for children, face and
 *                      edge points plus one for vertex points (default);
 *                      2: one patch-local kernel per level for children,
 *                      face, edge and the vertex points whose rings stay in
 *                      the tile, plus one for the remaining vertices;
 *                      0: one kernel per output kind, vertex scatter with
 *                      float atomics (the reference's own structure)
 *   CCD_OPT_STORE      child-record stores: 0 direct, 1 staged through LDS
 *   CCD_OPT_XCD_REMAP  1: each XCD processes one contiguous range of a level
 *   CCD_OPT_FAN        vertex rings: 2 enumerated from the level each vertex
 *                      appeared at, owners compacted per block (default);
 *                      3 one thread per vertex; 1 enumerated, one thread per
 *                      halfedge; 0 ring walks
 *   CCD_OPT_PATCH_TILE halfedges per tile of the patch-local kernel: 128,
 *                      256 (default) or 512, every level
 *   CCD_OPT_PATCH_HPT  tiles (halfedges per thread) each workgroup of the
 *                      patch-local kernel processes in turn: 1 (default)..64
 *   CCD_OPT_SCAN       vertex kernel after the fused kernel: 0 (default)
 *                      takes the vertex representatives the fused kernel
 *                      marked in a bit mask; 2, 4 or 8 scans that many
 *                      halfedges per thread for them
 *   CCD_OPT_ADJ_VERTEX the adjoint's two vertex passes run one thread per
 *                      vertex on levels with at most this many vertices
 *                      (default 1048576), else over the vertex owners in
 *                      halfedge order; both give the same bits
 *   CCD_OPT_FUSED_BLOCK threads (parent halfedges) per workgroup of the fused
 *                      kernel: 256, 512 or 1024 (default), every level
 *   CCD_OPT_FUSED_TILES tiles of that many halfedges each fused workgroup
 *                      processes in turn: 1 (default), 2 or 4
 *   CCD_OPT_VERTEX_WAVES waves per SIMD the masked vertex kernel is compiled
 *                      for: 4, 6 or 8 (default) */
enum {
    CCD_OPT_FUSED = 0, CCD_OPT_STORE = 1, CCD_OPT_XCD_REMAP = 2, CCD_OPT_FAN = 3,
    CCD_OPT_PATCH_TILE = 4, CCD_OPT_PATCH_HPT = 5, CCD_OPT_SCAN = 6, CCD_OPT_ADJ_VERTEX = 7,
    CCD_OPT_FUSED_BLOCK = 8, CCD_OPT_FUSED_TILES = 9, CCD_OPT_VERTEX_WAVES = 10
};
int32_t ccd_SetOption(ccd_Subd *subd, int32_t key, int32_t value);
/* Picks, per level, the kernel configuration of the fused paths by timing
 * candidates on this subd: first the fused kernel's workgroup size (256,
 * 512, 1024) and tiles per workgroup (1, 2, 4); then, from each level's
 * pick, the masked vertex kernel at 4 or 6 waves per SIMD and the compacted
 * vertex kernel at 2, 4 or 8 halfedges per thread; and the patch-local
 * kernel with tiles of 128, 256 or 512 halfedges.  Every candidate gives
 * the same bits.  Runs the given steps a few times per candidate on the
 * stream, synchronizing it; the choice is kept until the next ccd_SetOption
 * of these keys.  Returns 0 or an error code. */
int32_t ccd_Autotune(ccd_Subd *subd, int32_t steps, void *stream);
/* the value of CCD_OPT_FUSED (1 or 2), _PATCH_TILE, _PATCH_HPT, _SCAN,
 * _FUSED_BLOCK, _FUSED_TILES or _VERTEX_WAVES used at one level, or -1 */
int32_t ccd_TileOption(const ccd_Subd *subd, int32_t key, int32_t level);
/* per-kernel timing with HIP events on the launch stream: enable, run, then
 * read the kernel names and mean milliseconds of the last ccd_Refine. */
int32_t ccd_SetTiming(ccd_Subd *subd, int32_t enable);
int32_t ccd_TimingCount(const ccd_Subd *subd);
const char *ccd_TimingName(const ccd_Subd *subd, int32_t i);
float ccd_TimingMs(const ccd_Subd *subd, int32_t i);

/* ---- batched refinement (extension; no counterpart in the reference) ------
 * Many cages, each with its own depth, refined together: every kernel kind
 * runs once per level over all cages (64-bit offsets between cages), and the
 * whole sequence is captured once per set of steps into a hipGraph that later
 * calls replay.  Cage i's levels 1..D_i keep the cc_Subd layout inside the
 * shared device arrays (ccb_DevicePointer) at ccb_Offset(b, what, i).
 * A batched refinement always computes points (CCD_POINTS required);
 * CCD_HALFEDGES refines uvIDs too (written 0 for cages without uvs). */
typedef struct ccb_Batch ccb_Batch;
enum {
    CCB_COUNT_HALFEDGES = 0,      /* levels 1..D, cumulative */
    CCB_COUNT_CREASES = 1,
    CCB_COUNT_POINTS = 2,
    CCB_COUNT_CAGE_POINTS = 3,
    CCB_COUNT_FINEST_POINTS = 4,  /* points of level D; the layout of gradients */
    CCB_COUNT_DEPTH = 5,
    CCB_COUNT_CAGES = 6
};
enum { CCB_OPT_GRAPH = 100 };  /* 1: capture and replay hipGraphs (default) */
ccb_Batch *ccb_Create(const cc_Mesh *const *cages, const int32_t *depths, int32_t count);
void ccb_Release(ccb_Batch *batch);
/* counts of cage i, or totals over the batch for cage = -1 (depth: maximum) */
int64_t ccb_Count(const ccb_Batch *batch, int32_t what, int32_t cage);
/* element offset of cage i's slice in the array `what` names */
int64_t ccb_Offset(const ccb_Batch *batch, int32_t what, int32_t cage);
void *ccb_DevicePointer(const ccb_Batch *batch, int32_t which);   /* CCD_PTR_* */
int32_t ccb_Refine(ccb_Batch *batch, int32_t steps, void *stream);
/* points only, from new cage points (device, ccb_Count(CAGE_POINTS) x 3
 * floats in cage order) or, with NULL, the current ones */
int32_t ccb_RefinePoints(ccb_Batch *batch, const float *devCagePoints, int32_t steps, void *stream);
/* copies every cage's finest points into devOut (FINEST_POINTS x 3 floats) */
int32_t ccb_FinestPoints(ccb_Batch *batch, float *devOut, void *stream);
/* adjoint of the points: FINEST_POINTS x 3 gradients in, CAGE_POINTS x 3 out */
int32_t ccb_RefinePointsAdjoint(ccb_Batch *batch, const float *devGradFinest, float *devGradCage,
                                int32_t steps, void *stream);
int32_t ccb_Download(const ccb_Batch *batch, int32_t cage, cc_Subd *dst, int32_t steps);
int32_t ccb_Synchronize(ccb_Batch *batch);
/* CCB_OPT_GRAPH, CCD_OPT_XCD_REMAP, CCD_OPT_FUSED, CCD_OPT_PATCH_TILE,
 * CCD_OPT_PATCH_HPT, CCD_OPT_SCAN, CCD_OPT_FUSED_BLOCK, CCD_OPT_FUSED_TILES,
 * CCD_OPT_VERTEX_WAVES */
int32_t ccb_SetOption(ccb_Batch *batch, int32_t key, int32_t value);
/* per-level choice of tile size and tiles per workgroup by timing every
 * candidate on this batch (see ccd_Autotune) */
int32_t ccb_Autotune(ccb_Batch *batch, int32_t steps, void *stream);
int32_t ccb_TileOption(const ccb_Batch *batch, int32_t key, int32_t level);
/* per-launch timing with HIP events: while enabled, runs are eager (no
 * graph) and the names, levels and milliseconds read are those of the
 * launches of the last ccb_Refine, ccb_RefinePoints or
 * ccb_RefinePointsAdjoint */
int32_t ccb_SetTiming(ccb_Batch *batch, int32_t enable);
int32_t ccb_TimingCount(const ccb_Batch *batch);
const char *ccb_TimingName(const ccb_Batch *batch, int32_t i);
int32_t ccb_TimingLevel(const ccb_Batch *batch, int32_t i);
float ccb_TimingMs(const ccb_Batch *batch, int32_t i);

/* ---- refinement that allocates nothing -------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The refinement of ccd_Refine (same kernels, same bits,
 * `steps` as there), on arrays the caller already holds in device memory:
 *   deviceCage  a host cc_Mesh whose five counts are set and whose seven
 *               array pointers are device addresses, in cc_Mesh's layout
 *               (uvs may be NULL when uvCount is 0);
 *   deviceSubd  a host cc_Subd whose vertexPoints, halfedges and creases are
 *               device arrays sized for levels 1..maxDepth (the
 *               ccs_Cumulative*CountAtDepth counts); its cage is not read.
 * The kernels' scratch comes from one array of CCD_NOALLOC_SCRATCH_BYTES per
 * GPU, which the library holds until the process ends.  A cage and depth
 * needing more are refused (error 15, nothing launched); the need is
 *   4 (V0 + F0 + H0 2^(D-1) + 1) + H_{D-1} / 8 + (H_1 + ... + H_{D-2}) / 8
 * bytes (each part rounded up to 256 bytes, the last one only for D >= 4),
 * about 26 MB for a 8,192-halfedge cage at depth 8.  Every call on one GPU
 * shares that GPU's array: calls that overlap in time on one GPU race, so a
 * caller serializes them (one stream, or events).  Asynchronous on `stream`;
 * returns 0 or an error code (also in ccs_LastError). */
#define CCD_NOALLOC_SCRATCH_BYTES (64u << 20)
int32_t ccd_RefineNoAlloc(const cc_Mesh *deviceCage, const cc_Subd *deviceSubd, int32_t steps, void *stream);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* CC_INCLUDE_CC_HIP_H */
