/*
 * CPU oracle for the DiffTetVR tetrahedral volume-rendering hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * library, and only as the checker; the product path (difftetvr_amd) never links or calls it.
 *
 * It restates the reference INTERSECTION renderer (Data/Shaders/Intersect/, Data/Shaders/Common/,
 * Projection/ComputeTriangleDepths.glsl) for the CPU:
 *   1. per-tet projection and front-face generation      GenerateTrianglesIntersect.glsl:74-132
 *   2. per-face depth key = NDC centroid z                ComputeTriangleDepths.glsl:99-109
 *   3. stable ascending sort of the keys                  TetMeshRendererIntersection.cpp:682-696
 *   4. rasterisation of the projected faces (pixel-centre coverage, fixed point, 8 sub-pixel bits; the
 *      reference uses the Vulkan rasteriser, whose exact rounding is not observable here)
 *   5. per-fragment ray-tet compositing, front to back    IntersectRasterization.glsl:139-273
 *   6. adjoints (exact; and the reference's own, quirks kept)  AdjointIntersectRasterization.glsl:111-362
 *   7. torch.optim.Adam + colour clamp                    pytests/train.py:318,351-356
 *   8. per-cell colours (accumulateConst)                 IntersectRasterization.glsl:250-255, TetMesh.cpp:294-367
 *   9. the Projection renderer (projected tetrahedra): per-tet class and thick vertex, 3 or 4 screen triangles
 *      sorted by centroid depth, linear interpolation of colour and thickness, alpha = 1-exp(-a d)
 *                                                         Projection/GenerateTrianglesVTK.glsl:129-403,
 *                                                         Projection/ProjectedRasterization.glsl:85-105
 *
 * Parity status (how each part is pinned; see tests/test_oracle_*.py):
 *   camera uniforms      closed-form numpy float64 check
 *   projection/keys      numpy float32 restatement of the same formulas + invariants (1..3 faces per tet)
 *   sort                 numpy argsort(kind="stable") on the same keys
 *   coverage             watertight invariant: every pixel inside a convex mesh silhouette is covered by
 *                        exactly one front face per crossed surface; brute-force point-in-triangle in Fraction
 *   compositing          closed forms (constant colour: 1-exp(-sigma*chord)); scipy quad of the linear model;
 *                        chord-length invariant on the cube grid
 *   exact adjoint        FP64 central differences of the FP64 forward (h=1e-4, masked to pixels whose fragment
 *                        lists do not change)
 *   reference adjoint    on rays through one tet: FP64 central differences of the forward with sub-segment 0
 *                        frozen (colours, density) and of the two frozen-geometry forwards whose derivatives it
 *                        adds (positions); on rays through two to four tets, the chaining of state between
 *                        fragments (sub-segment 0 never inverted) by FP64 central differences of the closed form
 *                        it gives each tet's gradient (tests/test_reference_adjoint_pinning.py)
 *   Adam                 golden vectors written by torch.optim.Adam (tests/golden/adam_golden.npz)
 *   per-cell colours     closed form of a constant field (1-exp(-sigma*chord) per cell); the per-cell path
 *                        equals the per-vertex path on the unlinked mesh whose four corners carry the cell
 *                        colour; colour gradient by FP64 central differences (tests/test_cell_colors.py)
 *   Projection           one tet, class 1 and class 2: thick vertex, crossing and thickness against an
 *                        independent float64 construction (line crossings, eye-ray closest points); seen from
 *                        afar through a narrow field of view, the interpolated thickness sums to the chord of
 *                        the constant-colour mesh; adjoint by FP64 central differences (tests/test_projection.py)
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Face f uses vertices ORC_FT[f][0..2]; ORC_FT[f][3] is the opposite vertex.  Common/TetFaceTable.glsl:29-34. */
static const int ORC_FT[4][4] = {{0, 1, 2, 3}, {1, 0, 3, 2}, {0, 2, 3, 1}, {2, 1, 3, 0}};

/* Discrete choices of the Projection renderer for one tet (pt_decide). */
typedef struct {
    int ntri;              /* triangles: 0 (culled or degenerate), 3 (class 1) or 4 (class 2) */
    int seg1[2], seg2[2];  /* the two edges whose projections are intersected, as first chosen */
    int cls2;              /* the projections cross inside edge 1: class 2, vertex 4 is the crossing */
    int flip;              /* class 1 with the crossing before edge 1's first vertex: edge 1 is reversed */
} pt_dec;

typedef struct orc_frags {
    int W, H;
    int64_t* off;   /* [W*H + 1] CSR offsets of the per-pixel fragment lists */
    int32_t* ids;   /* fragment = tet*4 + face, front to back */
    int64_t ntris;  /* visible (generated) triangles */
    uint32_t* tri_key;
    int32_t* tri_id;
    pt_dec* pt_dec;   /* Projection renderer lists only: per-tet decisions (ids are then tet*4 + triangle) */
    int64_t T;
} orc_frags;

/* ------------------------------------------------------------------------------------------------------ */
/* Camera.  Perspective RH, depth [0,1], Y flipped for a top-left origin (TetMeshRendererPPLL.cpp:675-681),
 * fovy/near/far as set by the module (Module.cpp:457-463, :897-911).  Built in double and rounded once.     */
/* Uniform layout (68 values): viewProj[16] view[16] invProj[16] invView[16] camPos[3] pad.                 */

static void mul4_d(const double* A, const double* B, double* C) {
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++)
            C[c * 4 + r] = ((A[0 * 4 + r] * B[c * 4 + 0] + A[1 * 4 + r] * B[c * 4 + 1]) + A[2 * 4 + r] * B[c * 4 + 2]) +
                           A[3 * 4 + r] * B[c * 4 + 3];
}

/* Gauss-Jordan inverse with partial pivoting of a column-major 4x4. */
static int inv4_d(const double* m, double* out) {
    double a[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            a[r][c] = m[c * 4 + r];
            a[r][c + 4] = (r == c) ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; col++) {
        int piv = col;
        for (int r = col + 1; r < 4; r++)
            if (fabs(a[r][col]) > fabs(a[piv][col])) piv = r;
        if (a[piv][col] == 0.0) return -1;
        if (piv != col)
            for (int k = 0; k < 8; k++) { double t = a[col][k]; a[col][k] = a[piv][k]; a[piv][k] = t; }
        double inv = 1.0 / a[col][col];
        for (int k = 0; k < 8; k++) a[col][k] *= inv;
        for (int r = 0; r < 4; r++) {
            if (r == col) continue;
            double f = a[r][col];
            if (f == 0.0) continue;
            for (int k = 0; k < 8; k++) a[r][k] -= f * a[col][k];
        }
    }
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) out[c * 4 + r] = a[r][c + 4];
    return 0;
}

int orc_camera(const float view[16], float fovy, float znear, float zfar, int W, int H, float* uf, double* ud) {
    double V[16], P[16] = {0}, iP[16] = {0}, VP[16], iV[16];
    for (int i = 0; i < 16; i++) V[i] = (double)view[i];
    double th = tan(0.5 * (double)fovy);
    double aspect = (double)W / (double)H;
    P[0] = 1.0 / (aspect * th);
    P[5] = -1.0 / th;
    P[10] = (double)zfar / ((double)znear - (double)zfar);
    P[11] = -1.0;
    P[14] = -((double)zfar * (double)znear) / ((double)zfar - (double)znear);
    iP[0] = 1.0 / P[0];
    iP[5] = 1.0 / P[5];
    iP[11] = 1.0 / P[14];
    iP[14] = -1.0;
    iP[15] = P[10] / P[14];
    mul4_d(P, V, VP);
    if (inv4_d(V, iV) != 0) return -1;
    double u[68];
    memcpy(u + 0, VP, sizeof(VP));
    memcpy(u + 16, V, sizeof(V));
    memcpy(u + 32, iP, sizeof(iP));
    memcpy(u + 48, iV, sizeof(iV));
    u[64] = iV[12]; u[65] = iV[13]; u[66] = iV[14]; u[67] = 0.0;
    for (int i = 0; i < 68; i++) {
        if (ud) ud[i] = u[i];
        if (uf) uf[i] = (float)u[i];
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------------------ */
/* Discrete stage (float, no contraction): triangles, keys, sort, raster coverage -> per-pixel lists.       */

static inline float sgnf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline int64_t floordiv256(int64_t a) { return a >= 0 ? a / 256 : -((-a + 255) / 256); }

typedef struct { uint32_t key; int32_t id; int32_t X[3], Y[3]; int32_t x0, x1, y0, y1; } orc_tri;

/* Guard band in NDC units: tets with a vertex beyond |ndc.xy| > ORC_GUARD are culled (the reference's
 * rasteriser clips them instead; no configuration here reaches it). */
#define ORC_GUARD 16.0f

/* Project one tet and emit its front faces (GenerateTrianglesIntersect.glsl:80-131). Returns #faces. */
static int project_tet(const int32_t* cells, const float* pos, int64_t tet, const float* u, float znear, int W, int H,
                       orc_tri out[3]) {
    const float* VP = u;
    const float cam[3] = {u[64], u[65], u[66]};
    float p[4][3], ndc[4][3];
    for (int i = 0; i < 4; i++) {
        int32_t g = cells[tet * 4 + i];
        p[i][0] = pos[3 * (int64_t)g]; p[i][1] = pos[3 * (int64_t)g + 1]; p[i][2] = pos[3 * (int64_t)g + 2];
        float c[4];
        for (int r = 0; r < 4; r++)
            c[r] = ((VP[0 * 4 + r] * p[i][0] + VP[1 * 4 + r] * p[i][1]) + VP[2 * 4 + r] * p[i][2]) + VP[3 * 4 + r] * 1.0f;
        if (!(c[3] >= znear)) return 0;
        ndc[i][0] = c[0] / c[3]; ndc[i][1] = c[1] / c[3]; ndc[i][2] = c[2] / c[3];
        if (!(fabsf(ndc[i][0]) <= ORC_GUARD) || !(fabsf(ndc[i][1]) <= ORC_GUARD)) return 0;
    }
    int n = 0;
    const float hw = 0.5f * (float)W, hh = 0.5f * (float)H;
    for (int f = 0; f < 4; f++) {
        const float* a = p[ORC_FT[f][0]]; const float* b = p[ORC_FT[f][1]];
        const float* c = p[ORC_FT[f][2]]; const float* o = p[ORC_FT[f][3]];
        float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        float nrm[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        float d = -(nrm[0] * a[0] + nrm[1] * a[1] + nrm[2] * a[2]);
        float sc = sgnf((cam[0] * nrm[0] + cam[1] * nrm[1] + cam[2] * nrm[2]) + d);
        float so = sgnf((o[0] * nrm[0] + o[1] * nrm[1] + o[2] * nrm[2]) + d);
        if (!(sc != 0.0f && sc != so)) continue;
        orc_tri* t = &out[n++];
        t->id = (int32_t)(tet * 4 + f);
        float z = ((ndc[ORC_FT[f][0]][2] + ndc[ORC_FT[f][1]][2]) + ndc[ORC_FT[f][2]][2]) / 3.0f;
        t->key = fbits(z);
        for (int k = 0; k < 3; k++) {
            const float* q = ndc[ORC_FT[f][k]];
            float xf = q[0] * hw + hw;
            float yf = q[1] * hh + hh;
            t->X[k] = (int32_t)rintf(xf * 256.0f);
            t->Y[k] = (int32_t)rintf(yf * 256.0f);
        }
        int64_t area = (int64_t)(t->X[1] - t->X[0]) * (t->Y[2] - t->Y[0]) - (int64_t)(t->Y[1] - t->Y[0]) * (t->X[2] - t->X[0]);
        if (area < 0) {
            int32_t tx = t->X[1], ty = t->Y[1];
            t->X[1] = t->X[2]; t->Y[1] = t->Y[2]; t->X[2] = tx; t->Y[2] = ty;
        }
        int32_t mnx = t->X[0], mxx = t->X[0], mny = t->Y[0], mxy = t->Y[0];
        for (int k = 1; k < 3; k++) {
            if (t->X[k] < mnx) mnx = t->X[k];
            if (t->X[k] > mxx) mxx = t->X[k];
            if (t->Y[k] < mny) mny = t->Y[k];
            if (t->Y[k] > mxy) mxy = t->Y[k];
        }
        int64_t x0 = -floordiv256(-(int64_t)mnx + 128), x1 = floordiv256((int64_t)mxx - 128);
        int64_t y0 = -floordiv256(-(int64_t)mny + 128), y1 = floordiv256((int64_t)mxy - 128);
        if (x0 < 0) x0 = 0;
        if (y0 < 0) y0 = 0;
        if (x1 > W - 1) x1 = W - 1;
        if (y1 > H - 1) y1 = H - 1;
        if (area == 0) { x0 = 1; x1 = 0; }
        t->x0 = (int32_t)x0; t->x1 = (int32_t)x1; t->y0 = (int32_t)y0; t->y1 = (int32_t)y1;
    }
    return n;
}

/* Pixel-centre coverage with a tie rule that gives each shared edge to exactly one side. */
static inline int covers(const orc_tri* t, int x, int y) {
    int64_t px = 256 * (int64_t)x + 128, py = 256 * (int64_t)y + 128;
    for (int e = 0; e < 3; e++) {
        int a = e, b = (e + 1) % 3;
        int64_t dx = (int64_t)t->X[b] - t->X[a], dy = (int64_t)t->Y[b] - t->Y[a];
        int64_t E = dx * (py - t->Y[a]) - dy * (px - t->X[a]);
        if (E < 0) return 0;
        if (E == 0 && !(dy < 0 || (dy == 0 && dx > 0))) return 0;
    }
    return 1;
}

static int tri_cmp(const void* A, const void* B) {
    const orc_tri* a = (const orc_tri*)A;
    const orc_tri* b = (const orc_tri*)B;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return (a->id > b->id) - (a->id < b->id);
}

/* Builds the ordered per-pixel fragment lists of rows y_begin, y_begin + y_step, ... < y_end. */
orc_frags* orc_build_fragments(const int32_t* cells, int64_t T, const float* pos, const float* camu, float znear, int W,
                               int H, int y_begin, int y_end, int y_step) {
    orc_frags* fr = (orc_frags*)calloc(1, sizeof(orc_frags));
    fr->W = W; fr->H = H;
    orc_tri* all = (orc_tri*)malloc(sizeof(orc_tri) * 4 * (size_t)(T > 0 ? T : 1));
    int8_t* nf = (int8_t*)malloc((size_t)(T > 0 ? T : 1));
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < T; t++) nf[t] = (int8_t)project_tet(cells, pos, t, camu, znear, W, H, &all[4 * t]);
    int64_t n = 0;
    for (int64_t t = 0; t < T; t++)
        for (int k = 0; k < nf[t]; k++) all[n++] = all[4 * t + k];
    free(nf);
    qsort(all, (size_t)n, sizeof(orc_tri), tri_cmp);
    fr->ntris = n;
    fr->tri_key = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    fr->tri_id = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) { fr->tri_key[i] = all[i].key; fr->tri_id[i] = all[i].id; }
    const int64_t P = (int64_t)W * H;
    int64_t* cnt = (int64_t*)calloc((size_t)P + 1, sizeof(int64_t));
    if (y_begin < 0) y_begin = 0;
    if (y_end > H) y_end = H;
    if (y_step < 1) y_step = 1;
    const int nb = 64;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; b++) {
        int r0 = y_begin + (int)((int64_t)(y_end - y_begin) * b / nb), r1 = y_begin + (int)((int64_t)(y_end - y_begin) * (b + 1) / nb);
        for (int64_t i = 0; i < n; i++) {
            const orc_tri* t = &all[i];
            int ya = t->y0 > r0 ? t->y0 : r0, yb = t->y1 < r1 - 1 ? t->y1 : r1 - 1;
            for (int y = ya; y <= yb; y++) {
                if ((y - y_begin) % y_step) continue;
                for (int x = t->x0; x <= t->x1; x++)
                    if (covers(t, x, y)) cnt[(int64_t)y * W + x + 1]++;
            }
        }
    }
    for (int64_t i = 0; i < P; i++) cnt[i + 1] += cnt[i];
    fr->off = cnt;
    fr->ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cnt[P] > 0 ? cnt[P] : 1));
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)P);
    memcpy(fill, cnt, sizeof(int64_t) * (size_t)P);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; b++) {
        int r0 = y_begin + (int)((int64_t)(y_end - y_begin) * b / nb), r1 = y_begin + (int)((int64_t)(y_end - y_begin) * (b + 1) / nb);
        for (int64_t i = 0; i < n; i++) {
            const orc_tri* t = &all[i];
            int ya = t->y0 > r0 ? t->y0 : r0, yb = t->y1 < r1 - 1 ? t->y1 : r1 - 1;
            for (int y = ya; y <= yb; y++) {
                if ((y - y_begin) % y_step) continue;
                for (int x = t->x0; x <= t->x1; x++)
                    if (covers(t, x, y)) fr->ids[fill[(int64_t)y * W + x]++] = t->id;
            }
        }
    }
    free(fill);
    free(all);
    return fr;
}

int64_t orc_frags_total(const orc_frags* fr) { return fr->off[(int64_t)fr->W * fr->H]; }
int64_t orc_frags_num_tris(const orc_frags* fr) { return fr->ntris; }
void orc_frags_export(const orc_frags* fr, int64_t* off, int32_t* ids) {
    int64_t P = (int64_t)fr->W * fr->H;
    memcpy(off, fr->off, sizeof(int64_t) * (size_t)(P + 1));
    memcpy(ids, fr->ids, sizeof(int32_t) * (size_t)fr->off[P]);
}
void orc_frags_export_tris(const orc_frags* fr, uint32_t* keys, int32_t* ids) {
    memcpy(keys, fr->tri_key, sizeof(uint32_t) * (size_t)fr->ntris);
    memcpy(ids, fr->tri_id, sizeof(int32_t) * (size_t)fr->ntris);
}
void orc_frags_free(orc_frags* fr) {
    if (!fr) return;
    free(fr->off); free(fr->ids); free(fr->tri_key); free(fr->tri_id); free(fr->pt_dec); free(fr);
}

/* ------------------------------------------------------------------------------------------------------ */
/* Continuous stage, float and double instances.                                                           */

#define R float
#define SFX(n) n##_f
#define RSQRT_ sqrtf
#define REXP_ expf
#define RABS_ fabsf
#define IS_DOUBLE_ 0
#include "oracle_cont.h"
#include "oracle_proj.h"
#undef IS_DOUBLE_
#undef R
#undef SFX
#undef RSQRT_
#undef REXP_
#undef RABS_

#define R double
#define SFX(n) n##_d
#define RSQRT_ sqrt
#define REXP_ exp
#define RABS_ fabs
#define IS_DOUBLE_ 1
#include "oracle_cont.h"
#include "oracle_proj.h"
#undef IS_DOUBLE_
#undef R
#undef SFX
#undef RSQRT_
#undef REXP_
#undef RABS_

/* linear_depth: 0 = INTERSECTION segment length (Euclidean), 1 = PPLL's (difference of linear eye depths).
 * f32_decisions (double versions): the ray-tet decisions (hit, entry and exit faces, camera inside) are taken
 * from the FP32 evaluation (FP32 camera and ray, positions rounded to float), the continuous maths in FP64: the
 * FP64 counterpart of the FP32 function, free of the edge-grazing rays where FP32 and FP64 decide differently. */
/* cell_col (may be NULL): per-cell colours [T*4]; col is then not read and gcol is [T*4]. */
void orc_forward_f(const orc_frags* fr, const int32_t* cells, const float* pos, const float* col, const float* cell_col,
                   const float* camu, float att, const float* clear, int use_ert, float ert_alpha, int linear_depth,
                   float* img, float* C, int32_t* cnt) {
    forward_f(fr, cells, pos, col, cell_col, camu, att, clear, use_ert, ert_alpha, linear_depth, 0, NULL, img, C, cnt);
}
/* dec_pos (may be NULL, with f32_decisions): the float positions the decisions are taken at. */
void orc_forward_d(const orc_frags* fr, const int32_t* cells, const double* pos, const double* col,
                   const double* cell_col, const double* camu, double att, const double* clear, int use_ert,
                   double ert_alpha, int linear_depth, int f32_decisions, const float* dec_pos, double* img, double* C,
                   int32_t* cnt) {
    forward_d(fr, cells, pos, col, cell_col, camu, att, clear, use_ert, ert_alpha, linear_depth, f32_decisions, dec_pos,
              img, C, cnt);
}
void orc_adjoint_exact_f(const orc_frags* fr, const int32_t* cells, const float* pos, const float* col,
                         const float* cell_col, const float* camu, float att, const float* clear, int use_ert,
                         float ert_alpha, int linear_depth, const float* dimg, int use_abs, float* gpos, float* gcol) {
    adjoint_exact_f(fr, cells, pos, col, cell_col, camu, att, clear, use_ert, ert_alpha, linear_depth, 0, dimg, use_abs,
                    gpos, gcol);
}
void orc_adjoint_exact_d(const orc_frags* fr, const int32_t* cells, const double* pos, const double* col,
                         const double* cell_col, const double* camu, double att, const double* clear, int use_ert,
                         double ert_alpha, int linear_depth, int f32_decisions, const double* dimg, int use_abs,
                         double* gpos, double* gcol) {
    adjoint_exact_d(fr, cells, pos, col, cell_col, camu, att, clear, use_ert, ert_alpha, linear_depth, f32_decisions,
                    dimg, use_abs, gpos, gcol);
}
void orc_adjoint_reference_f(const orc_frags* fr, const int32_t* cells, const float* pos, const float* col,
                             const float* camu, float att, const float* img, const float* dimg, int use_abs, float* gpos,
                             float* gcol) {
    adjoint_reference_f(fr, cells, pos, col, camu, att, img, dimg, use_abs, gpos, gcol);
}

/* Single-fragment probes for closed-form tests. */
void orc_accumulate_linear_const_d(double t, const double* c0, const double* c1, double a, double* out) {
    v4_d r = acc_lin_const_d(t, mk3_d(c0[0], c0[1], c0[2]), mk3_d(c1[0], c1[1], c1[2]), a);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
int orc_ray_tet_d(const double* ro, const double* rd, const double* verts, int* f0, int* f1, double* t0, double* t1) {
    v3_d P[4];
    for (int i = 0; i < 4; i++) P[i] = mk3_d(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    return ray_tet_d(mk3_d(ro[0], ro[1], ro[2]), mk3_d(rd[0], rd[1], rd[2]), P, f0, f1, t0, t1);
}
void orc_ray_dir_f(const float* camu, int W, int H, int x, int y, float* out) {
    cam_f c;
    cam_load_f(&c, camu, W, H);
    v3_f d = ray_dir_f(&c, x, y);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

/* ------------------------------------------------------------------------------------------------------ */
/* torch.optim.Adam (single-tensor algorithm, torch/optim/adam.py) as used by pytests/train.py:318,351,
 * followed by the colour clamp c -= min(c, 0) (train.py:352-356) when clamp_nonneg is set.              */
void orc_adam_f(float* p, const float* g, float* m, float* v, int64_t n, double lr, double b1, double b2, double eps,
                int64_t step, int clamp_nonneg) {
    double bc1 = 1.0 - pow(b1, (double)step);
    double bc2 = 1.0 - pow(b2, (double)step);
    float step_size = (float)(lr / bc1);
    float bc2s = (float)sqrt(bc2);
    float w1 = (float)(1.0 - b1), fb2 = (float)b2, w2 = (float)(1.0 - b2), feps = (float)eps;
    for (int64_t i = 0; i < n; i++) {
        float gi = g[i];
        m[i] = m[i] + w1 * (gi - m[i]);
        v[i] = v[i] * fb2 + w2 * (gi * gi);
        float denom = sqrtf(v[i]) / bc2s + feps;
        p[i] = p[i] + (-step_size) * (m[i] / denom);
        if (clamp_nonneg) p[i] = p[i] - (p[i] < 0.0f ? p[i] : 0.0f);
    }
}

/* MSE loss gradient, torch.nn.MSELoss (mean) as in train.py:325,346: dL/dI = 2 (I - I*) / N. */
double orc_mse_grad_f(const float* img, const float* tgt, int64_t n, float* dimg) {
    double loss = 0.0;
    for (int64_t i = 0; i < n; i++) {
        float d = img[i] - tgt[i];
        loss += (double)d * (double)d;
        dimg[i] = 2.0f * d / (float)n;
    }
    return loss / (double)n;
}

/* Tet regulariser (TetRegularizer.glsl:108-166): gradient accumulated into gpos, and the loss it derives from. */
void orc_tet_regularizer_f(const int32_t* cells, int64_t T, const float* pos, float lambda, float beta, float* gpos) {
    tet_regularizer_f(cells, T, pos, lambda, beta, gpos);
}
void orc_tet_regularizer_d(const int32_t* cells, int64_t T, const double* pos, double lambda, double beta,
                           double* gpos) {
    tet_regularizer_d(cells, T, pos, lambda, beta, gpos);
}
double orc_tet_regularizer_loss_d(const int32_t* cells, int64_t T, const double* pos, double lambda, double beta) {
    return tet_regularizer_loss_d(cells, T, pos, lambda, beta);
}

/* ------------------------------------------------------------------------------------------------------ */
/* Projection renderer (projected tetrahedra): GenerateTrianglesVTK.glsl:129-403, ComputeTriangleDepths.glsl
 * :99-108, ProjectedRasterization.glsl.  Each tet in view becomes 3 or 4 triangles whose corners carry NDC
 * xyz, a colour and a thickness (0 except at the thick vertex); the triangles are sorted by the NDC z of their
 * centroid and blended back to front (premultiplied), i.e. composited front to back in ascending order; each
 * fragment contributes (c.rgb alpha, alpha), alpha = 1 - exp(-c.a d), with c and d interpolated linearly in
 * screen space.  Differences from the reference, stated: a tet with a vertex behind the near plane or beyond
 * the guard band is culled (the reference's NDC tests leave such a tet in and its triangles are then
 * undefined); equal keys are ordered by triangle id = tet * 4 + k (the reference by its atomic append order);
 * coverage is the tile renderer's fixed-point pixel-centre rule, for either winding (the pipeline culls none). */

/* Corner j of triangle k of a tet with decisions D: its projected-tetrahedra vertex (0-3 tet vertex, 4 the
 * crossing).  The triangle fans of GenerateTrianglesVTK.glsl:334-400. */
static int pt_corner(const pt_dec* D, int k, int j) {
    int a0 = D->seg1[0], a1 = D->seg1[1];
    if (D->flip) { int t = a0; a0 = a1; a1 = t; }
    const int b0 = D->seg2[0], b1 = D->seg2[1];
    if (D->cls2) {
        const int idx[6] = {4, a0, b0, a1, b1, a0};
        return j == 0 ? idx[0] : idx[k + j];
    }
    const int idx[5] = {a1, a0, b0, b1, a0};
    return j == 0 ? idx[0] : idx[k + j];
}

/* Float decisions for tet `tet` (GenerateTrianglesVTK.glsl:140-273), arithmetic as the GPU's. */
static void pt_decide(const int32_t* cells, const float* pos, int64_t tet, const float* u, float znear, pt_dec* D) {
    const float* VP = u;
    float ndc[4][3];
    int cull_w = 0;
    for (int i = 0; i < 4; i++) {
        int32_t g = cells[tet * 4 + i];
        float p[3] = {pos[3 * (int64_t)g], pos[3 * (int64_t)g + 1], pos[3 * (int64_t)g + 2]};
        float c[4];
        for (int r = 0; r < 4; r++)
            c[r] = ((VP[0 * 4 + r] * p[0] + VP[1 * 4 + r] * p[1]) + VP[2 * 4 + r] * p[2]) + VP[3 * 4 + r] * 1.0f;
        if (!(c[3] >= znear)) cull_w = 1;
        for (int k = 0; k < 3; k++) ndc[i][k] = c[k] / c[3];
        if (!(fabsf(ndc[i][0]) <= ORC_GUARD) || !(fabsf(ndc[i][1]) <= ORC_GUARD)) cull_w = 1;
    }
    D->ntri = 0;
    D->cls2 = D->flip = 0;
    if (cull_w) return;
    int all_gx = 1, all_lx = 1, all_gy = 1, all_ly = 1, all_gz = 1, any_lz = 0;
    for (int i = 0; i < 4; i++) {
        all_gx &= ndc[i][0] > 1.0f;
        all_lx &= ndc[i][0] < -1.0f;
        all_gy &= ndc[i][1] > 1.0f;
        all_ly &= ndc[i][1] < -1.0f;
        all_gz &= ndc[i][2] > 1.0f;
        any_lz |= ndc[i][2] < 0.0f;
    }
    if (all_gx || all_lx || all_gy || all_ly || all_gz || any_lz) return;
    float v1x = ndc[1][0] - ndc[0][0], v1y = ndc[1][1] - ndc[0][1];
    float v2x = ndc[2][0] - ndc[0][0], v2y = ndc[2][1] - ndc[0][1];
    float v3x = ndc[3][0] - ndc[0][0], v3y = ndc[3][1] - ndc[0][1];
    float fd1 = v3x * v2y - v3y * v2x;
    float fd2 = v1x * v3y - v1y * v3x;
    float fd3 = v2x * v1y - v2y * v1x;
    if ((fd1 * fd2 >= 0.0f) && ((fd1 != 0.0f) || (fd2 != 0.0f))) {
        D->seg1[0] = 0; D->seg1[1] = 3; D->seg2[0] = 1; D->seg2[1] = 2;
    } else if (fd1 * fd3 >= 0.0f) {
        D->seg1[0] = 0; D->seg1[1] = 2; D->seg2[0] = 1; D->seg2[1] = 3;
    } else {
        D->seg1[0] = 0; D->seg1[1] = 1; D->seg2[0] = 2; D->seg2[1] = 3;
    }
    const float* P1 = ndc[D->seg1[0]];
    const float* P2 = ndc[D->seg1[1]];
    const float* P3 = ndc[D->seg2[0]];
    const float* P4 = ndc[D->seg2[1]];
    float Ax = P2[0] - P1[0], Ay = P2[1] - P1[1];
    float Bx = P4[0] - P3[0], By = P4[1] - P3[1];
    float Cx = P3[0] - P1[0], Cy = P3[1] - P1[1];
    float den = Ax * By - Ay * Bx;
    if (fabsf(den) < 1e-8f) return;  /* degenerate */
    float alpha = (By * Cx - Bx * Cy) / den;
    D->cls2 = (alpha >= 0.0f) && (alpha <= 1.0f);
    D->flip = !D->cls2 && alpha <= 0.0f;
    D->ntri = D->cls2 ? 4 : 3;
}

/* Triangles of all tets in float (decisions, keys, fixed-point screen triangles): out[4 t + k] for k < nt[t]. */
static void pt_triangles(const int32_t* cells, int64_t T, const float* pos, const float* col, const float* cell_col,
                         const float* u, float znear, int W, int H, pt_dec* dec, orc_tri* out) {
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < T; t++) {
        pt_dec* D = &dec[t];
        pt_decide(cells, pos, t, u, znear, D);
        if (!D->ntri) continue;
        float P[4][3], C[4][4], ndc[5][3], cc[5][4], dep[5];
        for (int i = 0; i < 4; i++) {
            int32_t g = cells[t * 4 + i];
            for (int k = 0; k < 3; k++) P[i][k] = pos[3 * (int64_t)g + k];
            for (int k = 0; k < 4; k++) C[i][k] = cell_col ? cell_col[4 * t + k] : col[4 * (int64_t)g + k];
        }
        pt_vertices_f(D, u, P, C, cell_col != NULL, 1.0f, ndc, cc, dep);
        const float hw = 0.5f * (float)W, hh = 0.5f * (float)H;
        for (int k = 0; k < D->ntri; k++) {
            orc_tri* tr = &out[4 * t + k];
            tr->id = (int32_t)(t * 4 + k);
            int cv[3];
            for (int j = 0; j < 3; j++) cv[j] = pt_corner(D, k, j);
            float z = ((ndc[cv[0]][2] + ndc[cv[1]][2]) + ndc[cv[2]][2]) / 3.0f;
            tr->key = fbits(z);
            for (int j = 0; j < 3; j++) {
                float xf = ndc[cv[j]][0] * hw + hw;
                float yf = ndc[cv[j]][1] * hh + hh;
                tr->X[j] = (int32_t)rintf(xf * 256.0f);
                tr->Y[j] = (int32_t)rintf(yf * 256.0f);
            }
            int64_t area = (int64_t)(tr->X[1] - tr->X[0]) * (tr->Y[2] - tr->Y[0]) -
                           (int64_t)(tr->Y[1] - tr->Y[0]) * (tr->X[2] - tr->X[0]);
            if (area < 0) {
                int32_t tx = tr->X[1], ty = tr->Y[1];
                tr->X[1] = tr->X[2]; tr->Y[1] = tr->Y[2]; tr->X[2] = tx; tr->Y[2] = ty;
            }
            int32_t mnx = tr->X[0], mxx = tr->X[0], mny = tr->Y[0], mxy = tr->Y[0];
            for (int j = 1; j < 3; j++) {
                if (tr->X[j] < mnx) mnx = tr->X[j];
                if (tr->X[j] > mxx) mxx = tr->X[j];
                if (tr->Y[j] < mny) mny = tr->Y[j];
                if (tr->Y[j] > mxy) mxy = tr->Y[j];
            }
            int64_t x0 = -floordiv256(-(int64_t)mnx + 128), x1 = floordiv256((int64_t)mxx - 128);
            int64_t y0 = -floordiv256(-(int64_t)mny + 128), y1 = floordiv256((int64_t)mxy - 128);
            if (x0 < 0) x0 = 0;
            if (y0 < 0) y0 = 0;
            if (x1 > W - 1) x1 = W - 1;
            if (y1 > H - 1) y1 = H - 1;
            if (area == 0) { x0 = 1; x1 = 0; }
            tr->x0 = (int32_t)x0; tr->x1 = (int32_t)x1; tr->y0 = (int32_t)y0; tr->y1 = (int32_t)y1;
        }
    }
}

/* Ordered per-pixel lists of the sorted triangles `all[0..n)` covering each pixel (CSR into fr). */
static void raster_lists(const orc_tri* all, int64_t n, int W, int H, orc_frags* fr) {
    const int64_t P = (int64_t)W * H;
    int64_t* cnt = (int64_t*)calloc((size_t)P + 1, sizeof(int64_t));
    const int nb = 64;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; b++) {
        int r0 = (int)((int64_t)H * b / nb), r1 = (int)((int64_t)H * (b + 1) / nb);
        for (int64_t i = 0; i < n; i++) {
            const orc_tri* t = &all[i];
            int ya = t->y0 > r0 ? t->y0 : r0, yb = t->y1 < r1 - 1 ? t->y1 : r1 - 1;
            for (int y = ya; y <= yb; y++)
                for (int x = t->x0; x <= t->x1; x++)
                    if (covers(t, x, y)) cnt[(int64_t)y * W + x + 1]++;
        }
    }
    for (int64_t i = 0; i < P; i++) cnt[i + 1] += cnt[i];
    fr->off = cnt;
    fr->ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cnt[P] > 0 ? cnt[P] : 1));
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)P);
    memcpy(fill, cnt, sizeof(int64_t) * (size_t)P);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < nb; b++) {
        int r0 = (int)((int64_t)H * b / nb), r1 = (int)((int64_t)H * (b + 1) / nb);
        for (int64_t i = 0; i < n; i++) {
            const orc_tri* t = &all[i];
            int ya = t->y0 > r0 ? t->y0 : r0, yb = t->y1 < r1 - 1 ? t->y1 : r1 - 1;
            for (int y = ya; y <= yb; y++)
                for (int x = t->x0; x <= t->x1; x++)
                    if (covers(t, x, y)) fr->ids[fill[(int64_t)y * W + x]++] = t->id;
        }
    }
    free(fill);
}

/* Projected-tetrahedra fragment lists: the decisions and triangles of the mesh at float positions `pos`
 * (triangle ids tet * 4 + k), sorted by (key, id), rasterised into per-pixel lists front to back. */
orc_frags* orc_proj_fragments(const int32_t* cells, int64_t T, const float* pos, const float* col, const float* cell_col,
                              const float* camu, float znear, int W, int H) {
    orc_frags* fr = (orc_frags*)calloc(1, sizeof(orc_frags));
    fr->W = W; fr->H = H;
    pt_dec* dec = (pt_dec*)malloc(sizeof(pt_dec) * (size_t)(T > 0 ? T : 1));
    orc_tri* all = (orc_tri*)malloc(sizeof(orc_tri) * 4 * (size_t)(T > 0 ? T : 1));
    pt_triangles(cells, T, pos, col, cell_col, camu, znear, W, H, dec, all);
    int64_t n = 0;
    for (int64_t t = 0; t < T; t++)
        for (int k = 0; k < dec[t].ntri; k++) all[n++] = all[4 * t + k];
    qsort(all, (size_t)n, sizeof(orc_tri), tri_cmp);
    fr->ntris = n;
    fr->tri_key = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    fr->tri_id = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) { fr->tri_key[i] = all[i].key; fr->tri_id[i] = all[i].id; }
    raster_lists(all, n, W, H, fr);
    fr->pt_dec = dec;
    fr->T = T;
    free(all);
    return fr;
}

#define PT_FORWARD(SUF, R_, REXPF)                                                                                    \
    /* Composites the lists of fr (from orc_proj_fragments) with values in precision R_ at positions pos. */     \
    void orc_proj_forward##SUF(const orc_frags* fr, const int32_t* cells, const R_* pos, const R_* col,            \
                               const R_* cell_col, const R_* camu, R_ att, const R_* clear, R_* img, R_* outC) {   \
        const int W = fr->W, H = fr->H;                                                                           \
        const int64_t T = fr->T;                                                                                  \
        R_(*vn)[5][3] = (R_(*)[5][3])malloc(sizeof(R_) * 15 * (size_t)(T > 0 ? T : 1));                           \
        R_(*vc)[5][4] = (R_(*)[5][4])malloc(sizeof(R_) * 20 * (size_t)(T > 0 ? T : 1));                           \
        R_(*vd)[5] = (R_(*)[5])malloc(sizeof(R_) * 5 * (size_t)(T > 0 ? T : 1));                                  \
        _Pragma("omp parallel for schedule(static)") for (int64_t t = 0; t < T; t++) {                           \
            if (!fr->pt_dec[t].ntri) continue;                                                                    \
            R_ P[4][3], C[4][4];                                                                                  \
            for (int i = 0; i < 4; i++) {                                                                         \
                int32_t g = cells[t * 4 + i];                                                                     \
                for (int k = 0; k < 3; k++) P[i][k] = pos[3 * (int64_t)g + k];                                    \
                for (int k = 0; k < 4; k++) C[i][k] = cell_col ? cell_col[4 * t + k] : col[4 * (int64_t)g + k];   \
            }                                                                                                     \
            pt_vertices##SUF(&fr->pt_dec[t], camu, (const R_(*)[3])P, (const R_(*)[4])C, cell_col != NULL, att,  \
                             vn[t], vc[t], vd[t]);                                                                \
        }                                                                                                         \
        const int64_t NP = (int64_t)W * H;                                                                        \
        _Pragma("omp parallel for schedule(dynamic, 64)") for (int64_t px = 0; px < NP; px++) {                  \
            int x = (int)(px % W), y = (int)(px / W);                                                             \
            R_ pxn = (R_)2.0 * (((R_)x + (R_)0.5) / (R_)W) - (R_)1.0;                                             \
            R_ pyn = (R_)2.0 * (((R_)y + (R_)0.5) / (R_)H) - (R_)1.0;                                             \
            R_ rc[4] = {0, 0, 0, 0};                                                                              \
            for (int64_t q = fr->off[px]; q < fr->off[px + 1]; q++) {                                             \
                int32_t id = fr->ids[q];                                                                          \
                int64_t t = id >> 2;                                                                              \
                int k = id & 3;                                                                                   \
                R_ X[3], Y[3], cc[3][4], dd[3];                                                                   \
                for (int j = 0; j < 3; j++) {                                                                     \
                    int v = pt_corner(&fr->pt_dec[t], k, j);                                                      \
                    X[j] = vn[t][v][0];                                                                           \
                    Y[j] = vn[t][v][1];                                                                           \
                    for (int c = 0; c < 4; c++) cc[j][c] = vc[t][v][c];                                           \
                    dd[j] = vd[t][v];                                                                             \
                }                                                                                                 \
                v4##SUF f = pt_fragment##SUF(X, Y, (const R_(*)[4])cc, dd, pxn, pyn);                              \
                rc[0] = rc[0] + ((R_)1.0 - rc[3]) * f.x;                                                          \
                rc[1] = rc[1] + ((R_)1.0 - rc[3]) * f.y;                                                          \
                rc[2] = rc[2] + ((R_)1.0 - rc[3]) * f.z;                                                          \
                rc[3] = rc[3] + ((R_)1.0 - rc[3]) * f.w;                                                          \
            }                                                                                                     \
            R_ Tr = (R_)1.0 - rc[3];                                                                              \
            for (int c = 0; c < 4; c++) {                                                                         \
                if (outC) outC[4 * px + c] = rc[c];                                                               \
                img[4 * px + c] = rc[c] + Tr * clear[c];                                                          \
            }                                                                                                     \
        }                                                                                                         \
        free(vn);                                                                                                 \
        free(vc);                                                                                                 \
        free(vd);                                                                                                 \
    }
PT_FORWARD(_f, float, expf)
PT_FORWARD(_d, double, exp)

/* One tet's projected-tetrahedra vertices in double with float decisions (closed-form checks).  P: 4x3 world
 * positions, C: 4x4 colours; out: ntri, cls2, corners[4][3] (PT vertex of each triangle corner, -1 beyond ntri),
 * ndc[5][3], col[5][4], dep[5]. */
void orc_proj_tet_d(const double* P, const double* C, int per_cell, const double* camu_d, const float* camu_f,
                    float znear, double att, int* ntri, int* cls2, int* corners, double* ndc, double* col,
                    double* dep) {
    int32_t cells[4] = {0, 1, 2, 3};
    float pf[12];
    for (int i = 0; i < 12; i++) pf[i] = (float)P[i];
    pt_dec D;
    pt_decide(cells, pf, 0, camu_f, znear, &D);
    *ntri = D.ntri;
    *cls2 = D.cls2;
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < 3; j++) corners[3 * k + j] = k < D.ntri ? pt_corner(&D, k, j) : -1;
    if (!D.ntri) return;
    double Pa[4][3], Ca[4][4];
    for (int i = 0; i < 4; i++) {
        for (int k = 0; k < 3; k++) Pa[i][k] = P[3 * i + k];
        for (int k = 0; k < 4; k++) Ca[i][k] = C[4 * i + k];
    }
    pt_vertices_d(&D, camu_d, (const double(*)[3])Pa, (const double(*)[4])Ca, per_cell, att, (double(*)[3])ndc,
                  (double(*)[4])col, dep);
}
