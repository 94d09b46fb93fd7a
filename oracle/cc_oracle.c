/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle for the uniform Catmull-Clark
 * refinement path.  Nothing in the product library links, loads or calls
 * this file; only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark use it, and only as the checker.
 *
 * It restates, in plain C, the per-level refinement rules of
 * CatmullClark.h (jdupuy/HalfedgeCatmullClark) with the same floating-point
 * operation order, so that a float build compiled with -ffp-contract=off
 * reproduces the reference bit-for-bit (points of the scatter variant only
 * when run sequentially, which is what this oracle does), and a double build
 * serves as the high-precision error oracle.
 *
 * The file is compiled twice: -DORC_REAL=float -DORC_PREFIX=orc32_ and
 * -DORC_REAL=double -DORC_PREFIX=orc64_.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   - counts: closed forms against brute-force sums (CatmullClark.h:455-585,
 *     :811-936) and against the compiled reference;
 *   - topology, creases, uvIDs: bitwise against the compiled reference
 *     (oracle/_ref) and golden fixtures, plus invariants (twin involution,
 *     Euler characteristic, crease chains);
 *   - points: bitwise against the compiled reference (float, 1 thread),
 *     against the double build of the reference, and against closed forms
 *     (bicubic B-spline refinement on regular torus patches, the cubic
 *     B-spline curve rule on infinitely sharp chains, affine invariance,
 *     partition of unity).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORC_REAL
#define ORC_REAL float
#endif
#ifndef ORC_PREFIX
#define ORC_PREFIX orc32_
#endif

#define ORC_CAT2(a, b) a##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define ORC_API(name) ORC_CAT(ORC_PREFIX, name)

typedef ORC_REAL R;

/* Layouts mirror CatmullClark.h:17-55 with float replaced by R. */
typedef struct { int32_t twin, next, prev, face, edge, vert, uv; } CageHe; /* :37-45 */
typedef struct { int32_t twin, edge, vert, uv; } He;                      /* :48-55 */
typedef struct { int32_t next, prev; R sharp; } Crease;                   /* :30-34 */
typedef struct { R p[3]; } Pt;                                            /* :18-21 */
typedef struct { R t[2]; } Uv;                                            /* :24-27 */

#define ORC_MAXD 31

typedef struct {
    int32_t V, U, H, E, F;
    const CageHe *he;
    const int32_t *v2h, *e2h, *f2h;
    const Pt *pts;
    const Uv *uvs;
    const Crease *cr;
    /* per-level counts and cumulative offsets (levels 1..d), exact in int64 */
    int64_t Hc[ORC_MAXD + 1], Cc[ORC_MAXD + 1], Fc[ORC_MAXD + 1], Ec[ORC_MAXD + 1], Vc[ORC_MAXD + 1];
    int64_t Hcum[ORC_MAXD + 1], Ccum[ORC_MAXD + 1], Vcum[ORC_MAXD + 1];
} Cage;

typedef struct {
    Cage c;
    int32_t D;
    He *he;         /* levels 1..D, cumulative */
    Crease *cr;     /* levels 1..D, cumulative */
    Pt *pts;        /* levels 1..D, cumulative */
} Subd;

/* ------------------------------------------------------------------------ */
/* Counts (CatmullClark.h:455-585, :811-936), computed exactly in int64     */
/* from the recurrences F_{d+1} = H_d, E_{d+1} = 2 E_d + H_d, H_{d+1} = 4 H_d, */
/* C_{d+1} = 2 C_d, V_{d+1} = V_d + E_d + F_d.                              */
/* ------------------------------------------------------------------------ */
static void init_counts(Cage *c)
{
    c->Hc[0] = c->H; c->Cc[0] = c->E; c->Fc[0] = c->F; c->Ec[0] = c->E; c->Vc[0] = c->V;
    c->Hcum[0] = c->Ccum[0] = c->Vcum[0] = 0;
    for (int d = 1; d <= ORC_MAXD; ++d) {
        c->Hc[d] = 4 * c->Hc[d - 1];
        c->Cc[d] = 2 * c->Cc[d - 1];
        c->Fc[d] = c->Hc[d - 1];
        c->Ec[d] = 2 * c->Ec[d - 1] + c->Hc[d - 1];
        c->Vc[d] = c->Vc[d - 1] + c->Ec[d - 1] + c->Fc[d - 1];
        c->Hcum[d] = c->Hcum[d - 1] + c->Hc[d];
        c->Ccum[d] = c->Ccum[d - 1] + c->Cc[d];
        c->Vcum[d] = c->Vcum[d - 1] + c->Vc[d];
    }
}
static int64_t hcount(const Cage *c, int d) { return c->Hc[d]; }
static int64_t ccount(const Cage *c, int d) { return c->Cc[d]; }
static int64_t fcount(const Cage *c, int d) { return c->Fc[d]; }
static int64_t ecount(const Cage *c, int d) { return c->Ec[d]; }
static int64_t vcount(const Cage *c, int d) { return c->Vc[d]; }
/* cumulative counts over levels 1..d, i.e. the offset of level d+1 */
static int64_t hcum(const Cage *c, int d) { return c->Hcum[d]; }
static int64_t ccum(const Cage *c, int d) { return c->Ccum[d]; }
static int64_t vcum(const Cage *c, int d) { return c->Vcum[d]; }

int64_t ORC_API(count)(int32_t V, int32_t H, int32_t E, int32_t F, int32_t what, int32_t depth)
{
    Cage c; memset(&c, 0, sizeof(c));
    c.V = V; c.H = H; c.E = E; c.F = F;
    init_counts(&c);
    if (depth < 0 || depth > ORC_MAXD) return -1;
    switch (what) {
    case 0: return hcount(&c, depth);
    case 1: return ccount(&c, depth);
    case 2: return fcount(&c, depth);
    case 3: return ecount(&c, depth);
    case 4: return vcount(&c, depth);
    case 5: return hcum(&c, depth);
    case 6: return ccum(&c, depth);
    case 7: return vcum(&c, depth);
    default: return -1;
    }
}

/* ------------------------------------------------------------------------ */
/* Small helpers (CatmullClark.h:296-392).                                   */
/* ------------------------------------------------------------------------ */
static R rmin(R a, R b) { return a < b ? a : b; }
static R rmax(R a, R b) { return a > b ? a : b; }
static R sat(R x) { return rmax((R)0, rmin(x, (R)1)); }
static R sgn(R x) { return x < (R)0 ? (R)-1 : (x > (R)0 ? (R)1 : (R)0); }
static int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static R lerp1(R x, R y, R u) { return x + u * (y - x); }

static Pt padd(Pt a, Pt b) { Pt r; for (int i = 0; i < 3; ++i) r.p[i] = a.p[i] + b.p[i]; return r; }
static Pt pmul(Pt a, R s) { Pt r; for (int i = 0; i < 3; ++i) r.p[i] = a.p[i] * s; return r; }
static Pt plerp(Pt a, Pt b, R u) { Pt r; for (int i = 0; i < 3; ++i) r.p[i] = lerp1(a.p[i], b.p[i], u); return r; }
static Pt pzero(void) { Pt r = {{0, 0, 0}}; return r; }

/* UV codec, CatmullClark.h:374-392.  The float->uint32 conversion follows
 * what x86-64 compilers emit for the reference (a 64-bit truncating
 * conversion whose low 32 bits are kept).  The decode divides an integer by
 * the literal 65535.0f, which stays a single-precision division even in the
 * double build of the reference, so it is single precision here too. */
static Uv uv_decode(int32_t code)
{
    uint32_t t = (uint32_t)code;
    Uv r;
    r.t[0] = (R)((float)((t >> 0) & 0xFFFFu) / 65535.0f);
    r.t[1] = (R)((float)((t >> 16) & 0xFFFFu) / 65535.0f);
    return r;
}
static uint32_t r_to_u32(R x) { return (uint32_t)(int64_t)x; }
static int32_t uv_encode(Uv uv)
{
    uint32_t u = r_to_u32(uv.t[0] * (R)65535.0f);
    uint32_t v = r_to_u32(uv.t[1] * (R)65535.0f);
    return (int32_t)((u & 0xFFFFu) | ((v & 0xFFFFu) << 16));
}
static Uv uv_lerp_half(Uv a, Uv b)
{
    Uv r;
    r.t[0] = lerp1(a.t[0], b.t[0], (R)0.5f);
    r.t[1] = lerp1(a.t[1], b.t[1], (R)0.5f);
    return r;
}

/* quad arithmetic, CatmullClark.h:667-689 (two's complement, also applied
 * to the negative twin IDs that boundaries produce) */
static int32_t qnext(int32_t h) { return (int32_t)(((uint32_t)h & ~3u) | (((uint32_t)h + 1u) & 3u)); }
static int32_t qprev(int32_t h) { return (int32_t)(((uint32_t)h & ~3u) | (((uint32_t)h - 1u) & 3u)); }
static int32_t qface(int32_t h) { return h >> 2; }
static int32_t i4(int32_t x) { return (int32_t)((uint32_t)x * 4u); }

/* level-d accessors (d >= 1), CatmullClark.h:995-1161 */
static const He *lhe(const Subd *s, int d) { return s->he + hcum(&s->c, d - 1); }
static const Crease *lcr(const Subd *s, int d) { return s->cr + ccum(&s->c, d - 1); }
static const Pt *lpt(const Subd *s, int d) { return s->pts + vcum(&s->c, d - 1); }
static R lsharp(const Subd *s, int d, int32_t e)
{
    return e < ccount(&s->c, d) ? lcr(s, d)[e].sharp : (R)0;   /* :1011-1021 */
}
static int32_t ltwin(const Subd *s, int d, int32_t h) { return lhe(s, d)[h].twin; }
static int32_t ledge(const Subd *s, int d, int32_t h) { return lhe(s, d)[h].edge; }
static int32_t lvert(const Subd *s, int d, int32_t h) { return lhe(s, d)[h].vert; }
static R lhsharp(const Subd *s, int d, int32_t h) { return lsharp(s, d, ledge(s, d, h)); }
static Pt lhpt(const Subd *s, int d, int32_t h) { return lpt(s, d)[lvert(s, d, h)]; }

/* cage accessors, CatmullClark.h:592-757 */
static R chsharp(const Cage *c, int32_t h) { return c->cr[c->he[h].edge].sharp; }
static Pt chpt(const Cage *c, int32_t h) { return c->pts[c->he[h].vert]; }

/* ------------------------------------------------------------------------ */
/* O(depth) maps, CatmullClark.h:1206-1395 (non-recursive forms).           */
/* ------------------------------------------------------------------------ */
static int32_t edge_rep_first(const Cage *c, int32_t e)            /* :1206 */
{
    if (e >= 2 * c->E) {
        int32_t h = e - 2 * c->E;
        return imax(4 * h + 1, 4 * c->he[h].next + 2);
    } else if (e & 1) {
        return 4 * c->he[c->e2h[e >> 1]].next + 3;
    } else {
        return 4 * c->e2h[e >> 1] + 0;
    }
}

static int32_t edge_rep(const Cage *c, int32_t e, int d)           /* :1230 */
{
    uint32_t path = 1u;
    int32_t rep = 0;
    int k = d;
    for (; k > 1; --k) {
        int32_t Ep = (int32_t)ecount(c, k - 1);
        if (e >= 2 * Ep) {
            int32_t h = e - 2 * Ep;
            rep = imax(4 * h + 1, 4 * qnext(h) + 2);
            break;
        }
        path = (path << 1) | (uint32_t)(e & 1);
        e >>= 1;
    }
    if (k == 1) rep = edge_rep_first(c, e);
    for (; path > 1u; path >>= 1) {
        if (path & 1u) rep = 4 * qnext(rep) + 3;
        else rep = 4 * rep + 0;
    }
    return rep;
}

static int32_t vertex_rep_first(const Cage *c, int32_t v)          /* :1314 */
{
    if (v >= c->V + c->F) return 4 * c->e2h[v - c->V - c->F] + 1;
    if (v >= c->V) return 4 * c->f2h[v - c->V] + 2;
    return 4 * c->v2h[v] + 0;
}

static int32_t vertex_rep(const Cage *c, int32_t v, int d)         /* :1336 */
{
    int shift = 0;
    int32_t rep = 0;
    int k = d;
    for (; k > 1; --k) {
        int32_t Vp = (int32_t)vcount(c, k - 1);
        int32_t Fp = (int32_t)fcount(c, k - 1);
        if (v >= Vp + Fp) { rep = 4 * edge_rep(c, v - Vp - Fp, k - 1) + 1; break; }
        if (v >= Vp) { rep = 4 * (4 * (v - Vp)) + 2; break; }
        shift += 2;
    }
    if (k == 1) rep = vertex_rep_first(c, v);
    return rep << shift;
}

int32_t ORC_API(edge_to_halfedge)(const Subd *s, int32_t e, int32_t d) { return edge_rep(&s->c, e, d); }
int32_t ORC_API(vertex_to_halfedge)(const Subd *s, int32_t v, int32_t d) { return vertex_rep(&s->c, v, d); }

/* ------------------------------------------------------------------------ */
/* Halfedge refinement, CatmullClark.h:2693-2812.                           */
/* ------------------------------------------------------------------------ */
static void emit_children(He *out, int32_t h, int32_t twin, int32_t prev, int32_t next,
                          int32_t face, int32_t edge, int32_t vert, int32_t prevEdge,
                          int32_t prevTwin, int32_t twinNext,
                          int32_t V, int32_t F, int32_t E)
{
    He *c = out + 4 * (int64_t)h;
    c[0].twin = i4(twinNext) + 3;
    c[1].twin = i4(next) + 2;
    c[2].twin = i4(prev) + 1;
    c[3].twin = i4(prevTwin) + 0;
    c[0].edge = 2 * edge + (h > twin ? 0 : 1);
    c[1].edge = 2 * E + h;
    c[2].edge = 2 * E + prev;
    c[3].edge = 2 * prevEdge + (prev > prevTwin ? 1 : 0);
    c[0].vert = vert;
    c[1].vert = V + F + edge;
    c[2].vert = V + face;
    c[3].vert = V + F + prevEdge;
}

static void refine_halfedges(Subd *s)
{
    const Cage *c = &s->c;
    He *out = s->he;
    int32_t h;
#pragma omp parallel for
    for (h = 0; h < c->H; ++h) {
        const CageHe *x = &c->he[h];
        int32_t twinNext = x->twin >= 0 ? c->he[x->twin].next : -1;
        emit_children(out, h, x->twin, x->prev, x->next, x->face, x->edge, x->vert,
                      c->he[x->prev].edge, c->he[x->prev].twin, twinNext, c->V, c->F, c->E);
    }
    for (int d = 1; d < s->D; ++d) {
        const He *in = lhe(s, d);
        He *o = s->he + hcum(c, d);
        int32_t Hd = (int32_t)hcount(c, d), Vd = (int32_t)vcount(c, d);
        int32_t Fd = (int32_t)fcount(c, d), Ed = (int32_t)ecount(c, d);
#pragma omp parallel for
        for (h = 0; h < Hd; ++h) {
            int32_t p = qprev(h);
            emit_children(o, h, in[h].twin, p, qnext(h), qface(h), in[h].edge, in[h].vert,
                          in[p].edge, in[p].twin, qnext(in[h].twin), Vd, Fd, Ed);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* UV refinement, CatmullClark.h:2825-2939.                                  */
/* ------------------------------------------------------------------------ */
static void refine_uvs(Subd *s)
{
    const Cage *c = &s->c;
    int32_t h;
    if (c->U <= 0) return;
#pragma omp parallel for
    for (h = 0; h < c->H; ++h) {
        const CageHe *x = &c->he[h];
        Uv uv = c->uvs[x->uv];
        Uv nuv = c->uvs[c->he[x->next].uv];
        Uv puv = c->uvs[c->he[x->prev].uv];
        Uv fuv = uv;
        R m = 0;
        int32_t mi = 1;
        He *o = s->he + 4 * (int64_t)h;
        for (int32_t it = x->next; it != h; it = c->he[it].next) {
            Uv t = c->uvs[c->he[it].uv];
            fuv.t[0] += t.t[0];
            fuv.t[1] += t.t[1];
            ++mi;
        }
        m = (R)mi;
        fuv.t[0] /= m;
        fuv.t[1] /= m;
        o[0].uv = uv_encode(uv);
        o[1].uv = uv_encode(uv_lerp_half(uv, nuv));
        o[2].uv = uv_encode(fuv);
        o[3].uv = uv_encode(uv_lerp_half(uv, puv));
    }
    for (int d = 1; d < s->D; ++d) {
        const He *in = lhe(s, d);
        He *out = s->he + hcum(c, d);
        int32_t Hd = (int32_t)hcount(c, d);
#pragma omp parallel for
        for (h = 0; h < Hd; ++h) {
            Uv uv = uv_decode(in[h].uv);
            Uv nuv = uv_decode(in[qnext(h)].uv);
            Uv puv = uv_decode(in[qprev(h)].uv);
            Uv fuv = uv;
            He *o = out + 4 * (int64_t)h;
            for (int32_t it = qnext(h); it != h; it = qnext(it)) {
                Uv t = uv_decode(in[it].uv);
                fuv.t[0] += t.t[0];
                fuv.t[1] += t.t[1];
            }
            fuv.t[0] /= (R)4.0f;
            fuv.t[1] /= (R)4.0f;
            o[0].uv = in[h].uv;
            o[1].uv = uv_encode(uv_lerp_half(uv, nuv));
            o[2].uv = uv_encode(fuv);
            o[3].uv = uv_encode(uv_lerp_half(uv, puv));
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Crease refinement, CatmullClark.h:2950-3042.                              */
/* ------------------------------------------------------------------------ */
static void split_crease(Crease *o, const Crease *in, int32_t e)
{
    int32_t n = in[e].next, p = in[e].prev;
    int t1 = in[n].prev == e && n != e;
    int t2 = in[p].next == e && p != e;
    R s3 = (R)3.0f * in[e].sharp;
    R ns = in[n].sharp, ps = in[p].sharp;
    o[2 * (int64_t)e + 0].next = 2 * e + 1;
    o[2 * (int64_t)e + 1].next = 2 * n + (t1 ? 0 : 1);
    o[2 * (int64_t)e + 0].prev = 2 * p + (t2 ? 1 : 0);
    o[2 * (int64_t)e + 1].prev = 2 * e + 0;
    o[2 * (int64_t)e + 0].sharp = rmax((R)0, (ps + s3) / (R)4.0f - (R)1.0f);
    o[2 * (int64_t)e + 1].sharp = rmax((R)0, (s3 + ns) / (R)4.0f - (R)1.0f);
}

static void refine_creases(Subd *s)
{
    const Cage *c = &s->c;
    int32_t e;
#pragma omp parallel for
    for (e = 0; e < c->E; ++e) split_crease(s->cr, c->cr, e);
    for (int d = 1; d < s->D; ++d) {
        const Crease *in = lcr(s, d);
        Crease *out = s->cr + ccum(c, d);
        int32_t Cd = (int32_t)ccount(c, d);
#pragma omp parallel for
        for (e = 0; e < Cd; ++e) split_crease(out, in, e);
    }
}

/* ------------------------------------------------------------------------ */
/* Cage-level points (level 0 -> 1), CatmullClark.h:1408-1991.               */
/* ------------------------------------------------------------------------ */
static void cage_faces_gather(Subd *s)                               /* :1408 */
{
    const Cage *c = &s->c;
    Pt *fp = s->pts + c->V;
    int32_t f;
#pragma omp parallel for
    for (f = 0; f < c->F; ++f) {
        int32_t h = c->f2h[f];
        Pt acc = chpt(c, h);
        R n = (R)1.0f;
        for (int32_t it = c->he[h].next; it != h; it = c->he[it].next) {
            acc = padd(acc, chpt(c, it));
            n = n + (R)1.0f;
        }
        fp[f] = pmul(acc, (R)1.0f / n);
    }
}

static void cage_faces_scatter(Subd *s)                              /* :1437 */
{
    const Cage *c = &s->c;
    Pt *fp = s->pts + c->V;
    for (int32_t h = 0; h < c->H; ++h) {
        Pt v = chpt(c, h);
        R n = (R)1.0f;
        int32_t f = c->he[h].face;
        for (int32_t it = c->he[h].next; it != h; it = c->he[it].next) n = n + (R)1.0f;
        for (int i = 0; i < 3; ++i) fp[f].p[i] += v.p[i] / n;
    }
}

/* CatmullClark.h:1476 (smooth=0 uses twin<0 as the sharp test) and :1561 */
static Pt edge_gather_value(Pt v0, Pt v1, Pt f0, Pt f1, R w, int creased)
{
    Pt t1 = padd(v0, v1), t2 = padd(f0, f1);
    Pt sharp = pmul(t1, (R)0.5f);
    Pt smooth = pmul(padd(t1, t2), (R)0.25f);
    return creased ? plerp(smooth, sharp, w) : plerp(sharp, smooth, w);
}

static void cage_edges_gather(Subd *s, int creased)
{
    const Cage *c = &s->c;
    const Pt *fp = s->pts + c->V;
    Pt *ep = s->pts + c->V + c->F;
    int32_t e;
#pragma omp parallel for
    for (e = 0; e < c->E; ++e) {
        int32_t h = c->e2h[e];
        int32_t t = c->he[h].twin;
        int32_t n = c->he[h].next;
        R w = creased ? sat(c->cr[e].sharp) : (t < 0 ? (R)0.0f : (R)1.0f);
        ep[e] = edge_gather_value(chpt(c, h), chpt(c, n), fp[c->he[h].face],
                                  fp[c->he[imax(0, t)].face], w, creased);
    }
}

static void cage_edges_scatter(Subd *s, int creased)                 /* :1517, :1602 */
{
    const Cage *c = &s->c;
    const Pt *fp = s->pts + c->V;
    Pt *ep = s->pts + c->V + c->F;
    for (int32_t h = 0; h < c->H; ++h) {
        const CageHe *x = &c->he[h];
        Pt v0 = chpt(c, h), v1 = chpt(c, x->next), f = fp[x->face];
        Pt a;
        if (creased) {
            R w = sat(c->cr[x->edge].sharp);
            Pt sharp = pmul(plerp(v0, v1, (R)0.5f), x->twin < 0 ? (R)1.0f : (R)0.5f);
            Pt smooth = pmul(plerp(v0, f, (R)0.5f), (R)0.5f);
            a = plerp(smooth, sharp, w);
        } else {
            R w = x->twin >= 0 ? (R)0.5f : (R)1.0f;
            a = plerp(pmul(f, (R)0.5f), plerp(pmul(v0, w), pmul(v1, w), (R)0.5f), w);
        }
        for (int i = 0; i < 3; ++i) ep[x->edge].p[i] += a.p[i];
    }
}

static void cage_vertices_gather_smooth(Subd *s)                     /* :1661 */
{
    const Cage *c = &s->c;
    const Pt *fp = s->pts + c->V;
    const Pt *ep = s->pts + c->V + c->F;
    Pt *vp = s->pts;
    int32_t v;
#pragma omp parallel for
    for (v = 0; v < c->V; ++v) {
        int32_t h = c->v2h[v];
        Pt old = c->pts[v];
        Pt acc = padd(pmul(fp[c->he[h].face], (R)-1.0f), pmul(ep[c->he[h].edge], (R)4.0f));
        R n = (R)1.0f;
        int32_t it;
        for (it = c->he[c->he[h].prev].twin; it >= 0 && it != h; it = c->he[c->he[it].prev].twin) {
            acc = padd(acc, pmul(fp[c->he[it].face], (R)-1.0f));
            acc = padd(acc, pmul(ep[c->he[it].edge], (R)4.0f));
            n = n + (R)1.0f;
        }
        Pt sm = padd(pmul(acc, (R)1.0f / (n * n)), pmul(old, (R)1.0f - (R)3.0f / n));
        vp[v] = plerp(old, sm, it != h ? (R)0.0f : (R)1.0f);
    }
}

static void cage_vertices_scatter_smooth(Subd *s)                    /* :1713 */
{
    const Cage *c = &s->c;
    const Pt *fp = s->pts + c->V;
    const Pt *ep = s->pts + c->V + c->F;
    Pt *vp = s->pts;
    for (int32_t h = 0; h < c->H; ++h) {
        const CageHe *x = &c->he[h];
        Pt old = c->pts[x->vert];
        int32_t n = 1, fw, bw;
        for (fw = c->he[x->prev].twin; fw >= 0 && fw != h; fw = c->he[c->he[fw].prev].twin) ++n;
        for (bw = x->twin >= 0 ? c->he[x->twin].next : -1;
             fw < 0 && bw >= 0 && bw != h;
             bw = c->he[bw].twin >= 0 ? c->he[c->he[bw].twin].next : -1) ++n;
        for (int i = 0; i < 3; ++i) {
            R w = (R)1.0f / (R)n;
            R vv = old.p[i], f = fp[x->face].p[i], e = ep[x->edge].p[i];
            R sm = fw < 0 ? (R)0.0f : (R)1.0f;
            vp[x->vert].p[i] += w * (vv + w * sm * ((R)4.0f * e - f - (R)3.0f * vv));
        }
    }
}

static void cage_vertices_gather_creased(Subd *s)                    /* :1769 */
{
    const Cage *c = &s->c;
    const Pt *fp = s->pts + c->V;
    const Pt *ep = s->pts + c->V + c->F;
    Pt *vp = s->pts;
    int32_t v;
#pragma omp parallel for
    for (v = 0; v < c->V; ++v) {
        int32_t h = c->v2h[v];
        int32_t p = c->he[h].prev;
        R ts = chsharp(c, h), ps = chsharp(c, p);
        R cw = sgn(ts), pcw = sgn(ps);
        Pt E = ep[c->he[h].edge];
        Pt PE = ep[c->he[p].edge];
        Pt PF = fp[c->he[p].face];
        Pt old = c->pts[v];
        Pt sm = padd(pmul(PF, (R)-1.0f), pmul(PE, (R)4.0f));
        Pt cr = padd(pzero(), pmul(PE, pcw));
        R avgS = ps, cc = pcw, n = (R)1.0f;
        int32_t fw;
        for (fw = c->he[p].twin; fw >= 0 && fw != h; fw = c->he[fw].twin) {
            int32_t q = c->he[fw].prev;
            Pt qE = ep[c->he[q].edge], qF = fp[c->he[q].face];
            R qs = chsharp(c, q), qw = sgn(qs);
            sm = padd(sm, pmul(qF, (R)-1.0f));
            sm = padd(sm, pmul(qE, (R)4.0f));
            n = n + (R)1.0f;
            cr = padd(cr, pmul(qE, qw));
            avgS = avgS + qs;
            cc = cc + qw;
            fw = q;
        }
        if (fw < 0) {
            cr = padd(cr, pmul(E, cw));
            cc = cc + cw;
            n = n + (R)1.0f;
        }
        sm = padd(pmul(sm, (R)1.0f / (n * n)), pmul(old, (R)1.0f - (R)3.0f / n));
        cr = padd(pmul(cr, (R)0.25f), pmul(old, (R)0.5f));
        if (cc <= (R)1.0f) vp[v] = sm;
        else if (cc >= (R)3.0f || n == (R)2.0f) vp[v] = old;
        else vp[v] = plerp(old, cr, sat(avgS * (R)0.5f));
    }
}

static void cage_vertices_scatter_creased(Subd *s)                   /* :1872 */
{
    const Cage *c = &s->c;
    const Pt *fp = s->pts + c->V;
    const Pt *ep = s->pts + c->V + c->F;
    Pt *vp = s->pts;
    for (int32_t h = 0; h < c->H; ++h) {
        const CageHe *x = &c->he[h];
        int32_t p = x->prev;
        R ts = chsharp(c, h), ps = chsharp(c, p);
        R cw = sgn(ts), pcw = sgn(ps);
        Pt PE = ep[c->he[p].edge], E = ep[x->edge], F = fp[x->face];
        Pt old = c->pts[x->vert];
        R avgS = ps, cc = pcw, n = (R)1.0f;
        int32_t fw, bw;
        Pt corner, cr, sm, a;
        for (fw = c->he[p].twin; fw >= 0 && fw != h; fw = c->he[fw].twin) {
            int32_t q = c->he[fw].prev;
            R qs = chsharp(c, q);
            n = n + (R)1.0f;
            avgS = avgS + qs;
            cc = cc + sgn(qs);
            fw = q;
        }
        for (bw = x->twin; fw < 0 && bw >= 0 && bw != h; bw = c->he[bw].twin) {
            int32_t q = c->he[bw].next;
            R qs = chsharp(c, q);
            n = n + (R)1.0f;
            avgS = avgS + qs;
            cc = cc + sgn(qs);
            bw = q;
        }
        corner = pmul(old, (R)1.0f / n);
        cr = padd(pmul(old, (R)0.25f * cw), pmul(E, (R)0.25f * cw));
        sm = padd(pmul(F, (R)-1.0f), pmul(E, (R)4.0f));
        sm = padd(sm, pmul(old, n - (R)3.0f));
        sm = pmul(sm, (R)1.0f / (n * n));
        if (fw < 0) {
            Pt t;
            cc = cc + cw;
            n = n + (R)1.0f;
            t = padd(pmul(old, (R)0.25f * pcw), pmul(PE, (R)0.25f * pcw));
            cr = padd(cr, t);
        }
        if (cc <= (R)1.0f) a = sm;
        else if (cc >= (R)3.0f || n == (R)2.0f) a = corner;
        else a = plerp(corner, cr, sat(avgS * (R)0.5f));
        for (int i = 0; i < 3; ++i) vp[x->vert].p[i] += a.p[i];
    }
}

/* ------------------------------------------------------------------------ */
/* Level-d points (d >= 1 -> d+1), CatmullClark.h:2004-2615.                */
/* ------------------------------------------------------------------------ */
typedef struct {
    const Pt *old;  /* level d points */
    Pt *vp, *fp, *ep; /* level d+1: [V_d | F_d | E_d] */
    int32_t H, V, F, E;
} Lvl;

static Lvl level_ptrs(Subd *s, int d)
{
    Lvl L;
    const Cage *c = &s->c;
    L.H = (int32_t)hcount(c, d);
    L.V = (int32_t)vcount(c, d);
    L.F = (int32_t)fcount(c, d);
    L.E = (int32_t)ecount(c, d);
    L.old = lpt(s, d);
    L.vp = s->pts + vcum(c, d);
    L.fp = L.vp + L.V;
    L.ep = L.fp + L.F;
    return L;
}

static void faces_gather(Subd *s, int d)                             /* :2004 */
{
    Lvl L = level_ptrs(s, d);
    int32_t f;
#pragma omp parallel for
    for (f = 0; f < L.F; ++f) {
        int32_t h = 4 * f;
        Pt acc = lhpt(s, d, h);
        for (int32_t it = qnext(h); it != h; it = qnext(it)) acc = padd(acc, lhpt(s, d, it));
        L.fp[f] = pmul(acc, (R)0.25f);
    }
}

static void faces_scatter(Subd *s, int d)                            /* :2032 */
{
    Lvl L = level_ptrs(s, d);
    for (int32_t h = 0; h < L.H; ++h) {
        Pt v = lhpt(s, d, h);
        for (int i = 0; i < 3; ++i) L.fp[qface(h)].p[i] += v.p[i] / (R)4.0f;
    }
}

static void edges_gather(Subd *s, int d, int creased)                /* :2065, :2151 */
{
    Lvl L = level_ptrs(s, d);
    int32_t e;
#pragma omp parallel for
    for (e = 0; e < L.E; ++e) {
        int32_t h = edge_rep(&s->c, e, d);
        int32_t t = ltwin(s, d, h);
        R w = creased ? sat(lsharp(s, d, e)) : (t < 0 ? (R)0.0f : (R)1.0f);
        L.ep[e] = edge_gather_value(lhpt(s, d, h), lhpt(s, d, qnext(h)), L.fp[qface(h)],
                                    L.fp[qface(imax(0, t))], w, creased);
    }
}

static void edges_scatter(Subd *s, int d, int creased)               /* :2107, :2194 */
{
    Lvl L = level_ptrs(s, d);
    for (int32_t h = 0; h < L.H; ++h) {
        int32_t t = ltwin(s, d, h), e = ledge(s, d, h);
        Pt v0 = lhpt(s, d, h), v1 = lhpt(s, d, qnext(h)), f = L.fp[qface(h)];
        Pt a;
        if (creased) {
            R w = sat(lsharp(s, d, e));
            Pt sharp = pmul(plerp(v0, v1, (R)0.5f), t < 0 ? (R)1.0f : (R)0.5f);
            Pt smooth = pmul(plerp(v0, f, (R)0.5f), (R)0.5f);
            a = plerp(smooth, sharp, w);
        } else {
            R w = t >= 0 ? (R)0.5f : (R)1.0f;
            a = plerp(pmul(f, (R)0.5f), plerp(pmul(v0, w), pmul(v1, w), (R)0.5f), w);
        }
        for (int i = 0; i < 3; ++i) L.ep[e].p[i] += a.p[i];
    }
}

static void vertices_gather_smooth(Subd *s, int d)                   /* :2254 */
{
    Lvl L = level_ptrs(s, d);
    int32_t v;
#pragma omp parallel for
    for (v = 0; v < L.V; ++v) {
        int32_t h = vertex_rep(&s->c, v, d);
        Pt old = L.old[v];
        Pt acc = padd(pmul(L.fp[qface(h)], (R)-1.0f), pmul(L.ep[ledge(s, d, h)], (R)4.0f));
        R n = (R)1.0f;
        int32_t it;
        for (it = ltwin(s, d, qprev(h)); it >= 0 && it != h; it = ltwin(s, d, qprev(it))) {
            acc = padd(acc, pmul(L.fp[qface(it)], (R)-1.0f));
            acc = padd(acc, pmul(L.ep[ledge(s, d, it)], (R)4.0f));
            n = n + (R)1.0f;
        }
        Pt sm = padd(pmul(acc, (R)1.0f / (n * n)), pmul(old, (R)1.0f - (R)3.0f / n));
        L.vp[v] = plerp(old, sm, it != h ? (R)0.0f : (R)1.0f);
    }
}

static void vertices_scatter_smooth(Subd *s, int d)                  /* :2307 */
{
    Lvl L = level_ptrs(s, d);
    for (int32_t h = 0; h < L.H; ++h) {
        int32_t vx = lvert(s, d, h), e = ledge(s, d, h), f = qface(h);
        Pt old = L.old[vx];
        int32_t n = 1, fw, bw;
        for (fw = ltwin(s, d, qprev(h)); fw >= 0 && fw != h; fw = ltwin(s, d, qprev(fw))) ++n;
        for (bw = qnext(ltwin(s, d, h)); fw < 0 && bw >= 0 && bw != h; bw = qnext(ltwin(s, d, bw))) ++n;
        for (int i = 0; i < 3; ++i) {
            R w = (R)1.0f / (R)n;
            R vv = old.p[i], ff = L.fp[f].p[i], ee = L.ep[e].p[i];
            R sm = fw < 0 ? (R)0.0f : (R)1.0f;
            L.vp[vx].p[i] += w * (vv + w * sm * ((R)4.0f * ee - ff - (R)3.0f * vv));
        }
    }
}

static void vertices_gather_creased(Subd *s, int d)                  /* :2364 */
{
    Lvl L = level_ptrs(s, d);
    int32_t v;
#pragma omp parallel for
    for (v = 0; v < L.V; ++v) {
        int32_t h = vertex_rep(&s->c, v, d);
        int32_t p = qprev(h);
        R ts = lhsharp(s, d, h), ps = lhsharp(s, d, p);
        R cw = sgn(ts), pcw = sgn(ps);
        Pt E = L.ep[ledge(s, d, h)];
        Pt PE = L.ep[ledge(s, d, p)];
        Pt PF = L.fp[qface(p)];
        Pt old = L.old[v];
        Pt sm = padd(pmul(PF, (R)-1.0f), pmul(PE, (R)4.0f));
        Pt cr = padd(pzero(), pmul(PE, pcw));
        R avgS = ps, cc = pcw, n = (R)1.0f;
        int32_t fw, bw;
        for (fw = ltwin(s, d, p); fw >= 0 && fw != h; fw = ltwin(s, d, fw)) {
            int32_t q = qprev(fw);
            Pt qE = L.ep[ledge(s, d, q)], qF = L.fp[qface(q)];
            R qs = lhsharp(s, d, q), qw = sgn(qs);
            sm = padd(sm, pmul(qF, (R)-1.0f));
            sm = padd(sm, pmul(qE, (R)4.0f));
            n = n + (R)1.0f;
            cr = padd(cr, pmul(qE, qw));
            avgS = avgS + qs;
            cc = cc + qw;
            fw = q;
        }
        for (bw = ltwin(s, d, h); fw < 0 && bw >= 0 && bw != h; bw = ltwin(s, d, bw)) {
            int32_t q = qnext(bw);
            Pt qE = L.ep[ledge(s, d, q)], qF = L.fp[qface(q)];
            R qs = lhsharp(s, d, q), qw = sgn(qs);
            sm = padd(sm, pmul(qF, (R)-1.0f));
            sm = padd(sm, pmul(qE, (R)4.0f));
            n = n + (R)1.0f;
            cr = padd(cr, pmul(qE, qw));
            avgS = avgS + qs;
            cc = cc + qw;
            bw = q;
        }
        if (fw < 0) {
            cr = padd(cr, pmul(E, cw));
            cc = cc + cw;
            n = n + (R)1.0f;
        }
        sm = padd(pmul(sm, (R)1.0f / (n * n)), pmul(old, (R)1.0f - (R)3.0f / n));
        cr = padd(pmul(cr, (R)0.5f / cc), pmul(old, (R)0.5f));
        if (cc <= (R)1.0f) L.vp[v] = sm;
        else if (cc >= (R)3.0f || n == (R)2.0f) L.vp[v] = old;
        else L.vp[v] = plerp(old, cr, sat(avgS * (R)0.5f));
    }
}

static void vertices_scatter_creased(Subd *s, int d)                 /* :2496 */
{
    Lvl L = level_ptrs(s, d);
    for (int32_t h = 0; h < L.H; ++h) {
        int32_t vx = lvert(s, d, h), e = ledge(s, d, h), p = qprev(h);
        R ts = lhsharp(s, d, h), ps = lhsharp(s, d, p);
        R cw = sgn(ts), pcw = sgn(ps);
        Pt PE = L.ep[ledge(s, d, p)], E = L.ep[e], F = L.fp[qface(h)];
        Pt old = L.old[vx];
        R avgS = ps, cc = pcw, n = (R)1.0f;
        int32_t fw, bw;
        Pt corner, cr, sm, a;
        for (fw = ltwin(s, d, p); fw >= 0 && fw != h; fw = ltwin(s, d, fw)) {
            int32_t q = qprev(fw);
            R qs = lhsharp(s, d, q);
            n = n + (R)1.0f;
            avgS = avgS + qs;
            cc = cc + sgn(qs);
            fw = q;
        }
        for (bw = ltwin(s, d, h); fw < 0 && bw >= 0 && bw != h; bw = ltwin(s, d, bw)) {
            int32_t q = qnext(bw);
            R qs = lhsharp(s, d, q);
            n = n + (R)1.0f;
            avgS = avgS + qs;
            cc = cc + sgn(qs);
            bw = q;
        }
        corner = pmul(old, (R)1.0f / n);
        cr = padd(pmul(old, (R)0.25f * cw), pmul(E, (R)0.25f * cw));
        sm = padd(pmul(F, (R)-1.0f), pmul(E, (R)4.0f));
        sm = padd(sm, pmul(old, n - (R)3.0f));
        sm = pmul(sm, (R)1.0f / (n * n));
        if (fw < 0) {
            Pt t;
            cc = cc + cw;
            n = n + (R)1.0f;
            t = padd(pmul(old, (R)0.25f * pcw), pmul(PE, (R)0.25f * pcw));
            cr = padd(cr, t);
        }
        /* branch order differs from the cage and gather kernels (:2598) */
        if (cc >= (R)3.0f || n == (R)2.0f) a = corner;
        else if (cc <= (R)1.0f) a = sm;
        else a = plerp(corner, cr, sat(avgS * (R)0.5f));
        for (int i = 0; i < 3; ++i) L.vp[vx].p[i] += a.p[i];
    }
}

/* ------------------------------------------------------------------------ */
/* Drivers, CatmullClark.h:2631-2683 and :3051-3082.                         */
/* ------------------------------------------------------------------------ */
enum {
    ORC_HALFEDGES = 1, ORC_CREASES = 2, ORC_UVS = 4, ORC_POINTS = 8,
    ORC_GATHER = 16, ORC_NOCREASES = 32
};

static void refine_points(Subd *s, int gather, int creased)
{
    const Cage *c = &s->c;
    if (!gather) {
        memset(s->pts, 0, (size_t)vcum(c, s->D) * sizeof(Pt));   /* :2623 */
        cage_faces_scatter(s);
        cage_edges_scatter(s, creased);
        if (creased) cage_vertices_scatter_creased(s); else cage_vertices_scatter_smooth(s);
        for (int d = 1; d < s->D; ++d) {
            faces_scatter(s, d);
            edges_scatter(s, d, creased);
            if (creased) vertices_scatter_creased(s, d); else vertices_scatter_smooth(s, d);
        }
    } else {
        cage_faces_gather(s);
        cage_edges_gather(s, creased);
        if (creased) cage_vertices_gather_creased(s); else cage_vertices_gather_smooth(s);
        for (int d = 1; d < s->D; ++d) {
            faces_gather(s, d);
            edges_gather(s, d, creased);
            if (creased) vertices_gather_creased(s, d); else vertices_gather_smooth(s, d);
        }
    }
}

/*
 * Refine a cage given as flat arrays (layouts as in CatmullClark.h with float
 * replaced by the build's real type).  Output arrays hold levels 1..maxDepth
 * in the reference's cumulative order.  flags selects the steps; nthreads
 * sets the OpenMP thread count of the deterministic loops (the scatter
 * point loops always run sequentially, in ascending halfedge order).
 */
int32_t ORC_API(refine)(int32_t V, int32_t U, int32_t H, int32_t E, int32_t F,
                        const void *cageHalfedges, const int32_t *v2h, const int32_t *e2h,
                        const int32_t *f2h, const void *points, const void *uvs,
                        const void *creases, int32_t maxDepth, int32_t flags, int32_t nthreads,
                        void *outHalfedges, void *outCreases, void *outPoints)
{
    Subd s;
    if (maxDepth < 1) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    s.c.V = V; s.c.U = U; s.c.H = H; s.c.E = E; s.c.F = F;
    s.c.he = (const CageHe *)cageHalfedges;
    s.c.v2h = v2h; s.c.e2h = e2h; s.c.f2h = f2h;
    s.c.pts = (const Pt *)points;
    s.c.uvs = (const Uv *)uvs;
    s.c.cr = (const Crease *)creases;
    init_counts(&s.c);
    if (maxDepth > ORC_MAXD) return -1;
    s.D = maxDepth;
    s.he = (He *)outHalfedges;
    s.cr = (Crease *)outCreases;
    s.pts = (Pt *)outPoints;
    if (flags & ORC_HALFEDGES) refine_halfedges(&s);
    if (flags & ORC_CREASES) refine_creases(&s);
    if (flags & ORC_UVS) refine_uvs(&s);
    if (flags & ORC_POINTS) refine_points(&s, (flags & ORC_GATHER) != 0, (flags & ORC_NOCREASES) == 0);
    return 0;
}

/* The O(depth) maps, exposed for tests (CatmullClark.h:1230, :1336). */
int32_t ORC_API(map)(int32_t V, int32_t H, int32_t E, int32_t F, const void *cageHalfedges,
                     const int32_t *v2h, const int32_t *e2h, const int32_t *f2h,
                     int32_t which, int32_t id, int32_t depth)
{
    Cage c;
    memset(&c, 0, sizeof(c));
    c.V = V; c.H = H; c.E = E; c.F = F;
    c.he = (const CageHe *)cageHalfedges;
    c.v2h = v2h; c.e2h = e2h; c.f2h = f2h;
    init_counts(&c);
    return which == 0 ? edge_rep(&c, id, depth) : vertex_rep(&c, id, depth);
}

/* ------------------------------------------------------------------------ */
/* Per-halfedge normals of one level of quads                               */
/* (glsl/CatmullClark_Scatter.glsl:931-960, the only place the reference   */
/* defines them).  The corner normal of h is                                */
/* normalize(cross(P(next h) - P(h), P(prev h) - P(h))); the vertex normal  */
/* is normalize(corner(h) + sum of corner(it)) over the halfedges it met    */
/* walking it = twin(prev(it)) from h, or corner(h) when the walk reaches   */
/* a boundary.  normalize(v) = v * (1 / sqrt(dot(v, v))) with a correctly   */
/* rounded square root.                                                     */
/* ------------------------------------------------------------------------ */
static Pt psub(Pt a, Pt b) { Pt r; for (int i = 0; i < 3; ++i) r.p[i] = a.p[i] - b.p[i]; return r; }

static Pt pcross(Pt a, Pt b)
{
    Pt r;
    r.p[0] = a.p[1] * b.p[2] - b.p[1] * a.p[2];
    r.p[1] = a.p[2] * b.p[0] - b.p[2] * a.p[0];
    r.p[2] = a.p[0] * b.p[1] - b.p[0] * a.p[1];
    return r;
}

static Pt pnormalize(Pt v)
{
    const R dd = v.p[0] * v.p[0] + v.p[1] * v.p[1] + v.p[2] * v.p[2];
    const R inv = (R)1 / (R)sqrt((double)dd);
    return pmul(v, inv);
}

static Pt corner_normal(const He *he, const Pt *pts, int32_t h)
{
    const Pt v0 = pts[he[h].vert], v1 = pts[he[qprev(h)].vert], v2 = pts[he[qnext(h)].vert];
    return pnormalize(pcross(psub(v2, v0), psub(v1, v0)));
}

int32_t ORC_API(halfedge_normals)(int64_t H, const void *levelHalfedges, const void *levelPoints, int32_t smooth,
                                  void *out)
{
    const He *he = (const He *)levelHalfedges;
    const Pt *pts = (const Pt *)levelPoints;
    Pt *o = (Pt *)out;
#pragma omp parallel for
    for (int64_t k = 0; k < H; ++k) {
        const int32_t h = (int32_t)k;
        const Pt n = corner_normal(he, pts, h);
        if (!smooth) { o[k] = n; continue; }
        Pt avg = pzero();
        int32_t it = he[qprev(h)].twin;
        for (int64_t g = 0; it >= 0 && it != h && g < H; ++g) {
            avg = padd(avg, corner_normal(he, pts, it));
            it = he[qprev(it)].twin;
        }
        o[k] = it < 0 ? n : pnormalize(padd(n, avg));
    }
    return 0;
}
