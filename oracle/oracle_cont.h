/*
 * CPU oracle: continuous part (ray-tet intersection, barycentric colours, sub-segment compositing and the
 * adjoints), instantiated twice from oracle.c with R = float (parity against the HIP path) and R = double
 * (finite-difference checks).  TEST INFRASTRUCTURE ONLY: nothing in the product links this file.
 *
 * Every function restates a reference GLSL routine; the citation is on the function.  Arithmetic is written
 * operation-by-operation in the order of the reference expressions so that, compiled with
 * -ffp-contract=off, the float instance reproduces the reference's float rounding except for exp().
 *
 * Required macros: R (float|double), SFX(name), RSQRT_ (sqrt function), REXP_ (exp function), RABS_ (fabs).
 */

typedef struct { R x, y, z; } SFX(v3);
typedef struct { R x, y, z, w; } SFX(v4);

static inline SFX(v3) SFX(mk3)(R x, R y, R z) { SFX(v3) r = {x, y, z}; return r; }
static inline SFX(v4) SFX(mk4)(R x, R y, R z, R w) { SFX(v4) r = {x, y, z, w}; return r; }
static inline SFX(v3) SFX(sub3)(SFX(v3) a, SFX(v3) b) { return SFX(mk3)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline SFX(v3) SFX(add3)(SFX(v3) a, SFX(v3) b) { return SFX(mk3)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline SFX(v3) SFX(scl3)(SFX(v3) a, R s) { return SFX(mk3)(a.x * s, a.y * s, a.z * s); }
/* GLSL dot: (a.x*b.x + a.y*b.y) + a.z*b.z */
static inline R SFX(dot3)(SFX(v3) a, SFX(v3) b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline SFX(v3) SFX(cross3)(SFX(v3) a, SFX(v3) b) {
    return SFX(mk3)(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline R SFX(len3)(SFX(v3) a) { return RSQRT_(SFX(dot3)(a, a)); }
static inline R SFX(maxr)(R a, R b) { return a > b ? a : b; }

/* Reference uniform block restated from IntersectUniform.glsl:29-41 (matrices column-major m[c*4+r]). */
typedef struct {
    R viewProj[16], view[16], invProj[16], invView[16];
    R camPos[3];
    int W, H;
} SFX(cam);

static void SFX(cam_load)(SFX(cam) * c, const R* u, int W, int H) {
    memcpy(c->viewProj, u + 0, 16 * sizeof(R));
    memcpy(c->view, u + 16, 16 * sizeof(R));
    memcpy(c->invProj, u + 32, 16 * sizeof(R));
    memcpy(c->invView, u + 48, 16 * sizeof(R));
    c->camPos[0] = u[64]; c->camPos[1] = u[65]; c->camPos[2] = u[66];
    c->W = W; c->H = H;
}

/* Camera ray through the centre of pixel (x, y): IntersectRasterization.glsl:141-144. */
static SFX(v3) SFX(ray_dir)(const SFX(cam) * c, int x, int y) {
    R fx = (R)x + (R)0.5, fy = (R)y + (R)0.5;
    R nx = (R)2.0 * (fx / (R)c->W) - (R)1.0;
    R ny = (R)2.0 * (fy / (R)c->H) - (R)1.0;
    const R* m = c->invProj;
    R t[3];
    for (int r = 0; r < 3; r++) t[r] = ((m[0 * 4 + r] * nx + m[1 * 4 + r] * ny) + m[2 * 4 + r] * (R)1.0) + m[3 * 4 + r] * (R)1.0;
    R l = RSQRT_(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    R n[3] = {t[0] / l, t[1] / l, t[2] / l};
    const R* v = c->invView;
    R d[3];
    for (int r = 0; r < 3; r++) d[r] = ((v[0 * 4 + r] * n[0] + v[1 * 4 + r] * n[1]) + v[2 * 4 + r] * n[2]) + v[3 * 4 + r] * (R)0.0;
    return SFX(mk3)(d[0], d[1], d[2]);
}

/* rayTriangleIntersect: Intersect/RayIntersectionTests.glsl:30-76. */
static int SFX(ray_tri)(SFX(v3) ro, SFX(v3) rd, SFX(v3) p0, SFX(v3) p1, SFX(v3) p2, R* t, int* isInside) {
    SFX(v3) p0p1 = SFX(sub3)(p1, p0);
    SFX(v3) p0p2 = SFX(sub3)(p2, p0);
    SFX(v3) N = SFX(cross3)(p0p1, p0p2);
    R cosNR = SFX(dot3)(N, rd);
    if (RABS_(cosNR) < (R)1e-9) return 0;
    R d = -SFX(dot3)(N, p0);
    *t = -(SFX(dot3)(N, ro) + d) / cosNR;
    SFX(v3) P = SFX(add3)(ro, SFX(scl3)(rd, *t));
    SFX(v3) p0p = SFX(sub3)(P, p0);
    if (SFX(dot3)(N, SFX(cross3)(p0p1, p0p)) < (R)0) return 0;
    SFX(v3) p2p1 = SFX(sub3)(p2, p1);
    SFX(v3) p1p = SFX(sub3)(P, p1);
    if (SFX(dot3)(N, SFX(cross3)(p2p1, p1p)) < (R)0) return 0;
    SFX(v3) p2p0 = SFX(sub3)(p0, p2);
    SFX(v3) p2p = SFX(sub3)(P, p2);
    if (SFX(dot3)(N, SFX(cross3)(p2p0, p2p)) < (R)0) return 0;
    *isInside = 1;
    return *t >= (R)0;
}

/* intersectRayTet: Intersect/RayIntersectionTests.glsl:78-106 (faces from Common/TetFaceTable.glsl:29-34). */
static int SFX(ray_tet)(SFX(v3) ro, SFX(v3) rd, const SFX(v3) P[4], int* f0, int* f1, R* t0, R* t1) {
    *t0 = (R)1e9; *t1 = (R)-1e9; *f0 = 0; *f1 = 0;
    for (int f = 0; f < 4; f++) {
        R t = (R)1e9;
        int inside = 0;
        int hit = SFX(ray_tri)(ro, rd, P[ORC_FT[f][0]], P[ORC_FT[f][1]], P[ORC_FT[f][2]], &t, &inside);
        if (hit) {
            if (t < *t0) { *f0 = f; *t0 = t; }
            if (t > *t1) { *f1 = f; *t1 = t; }
        } else if (inside && t < (R)0) {
            *f0 = f; *t0 = (R)0;
        }
    }
    return *t1 >= *t0;
}

/* barycentricInterpolation (unsigned areas, 1e-5 floor): Common/BarycentricInterpolation.glsl:29-38. */
static void SFX(bary)(SFX(v3) p0, SFX(v3) p1, SFX(v3) p2, SFX(v3) P, R* u, R* v) {
    SFX(v3) d20 = SFX(sub3)(p2, p0);
    SFX(v3) d21 = SFX(sub3)(p2, p1);
    R area = SFX(maxr)(SFX(len3)(SFX(cross3)(d20, d21)), (R)1e-5);
    *u = SFX(len3)(SFX(cross3)(d21, SFX(sub3)(P, p1))) / area;
    *v = SFX(len3)(SFX(cross3)(SFX(sub3)(P, p0), d20)) / area;
}

/* accumulateLinearConst (forward variant): Common/ForwardCommon.glsl:42-49. */
static SFX(v4) SFX(acc_lin_const)(R t, SFX(v3) c0, SFX(v3) c1, R a) {
    if (a < (R)1e-6) return SFX(mk4)(0, 0, 0, 0);
    R A = REXP_(-a * t);
    R k = (t + (R)1.0 / a) * A - (R)1.0 / a;
    return SFX(mk4)(((R)1.0 - A) * c0.x + k * (c0.x - c1.x), ((R)1.0 - A) * c0.y + k * (c0.y - c1.y),
                    ((R)1.0 - A) * c0.z + k * (c0.z - c1.z), (R)1.0 - A);
}

/* GLSL mix(x, y, a) = x*(1-a) + y*a. */
static inline R SFX(mixr)(R x, R y, R a) { return x * ((R)1.0 - a) + y * a; }

/* Everything the forward computes for one (pixel, tet) fragment. */
typedef struct {
    int hit;              /* intersectRayTet result */
    int f0, f1;
    R t0, t1;
    SFX(v3) P0, P1;
    R u0, v0, u1, v1;     /* barycentrics on the entry / exit face */
    SFX(v4) col0, col1;   /* interpolated RGBA at entry / exit */
    SFX(v4) seg[4];       /* accumulateLinearConst per sub-segment */
    SFX(v4) state[5];     /* front-to-back composite inside the tet, state[0] = 0 */
    R a[4];               /* density * attenuation per sub-segment */
    SFX(v3) sc0[4], sc1[4];
    R tscale;             /* segment length per unit ray length: 1, or the linear-depth factor (PPLL) */
    int per_cell;         /* one colour per tet: accumulateConst over the whole segment */
} SFX(frag);

/* Tet `tet`'s vertices and colours.  cell_col (may be NULL): per-cell colours [T*4] (ColorStorage.PER_CELL,
 * TetMesh.cpp:310-343); every C[i] is then the tet's colour and col is not read. */
static void SFX(load_tet)(const int32_t* cells, const R* pos, const R* col, const R* cell_col, int64_t tet,
                          SFX(v3) P[4], SFX(v4) C[4], int32_t vid[4]) {
    for (int i = 0; i < 4; i++) {
        int32_t g = cells[tet * 4 + i];
        vid[i] = g;
        P[i] = SFX(mk3)(pos[3 * (int64_t)g], pos[3 * (int64_t)g + 1], pos[3 * (int64_t)g + 2]);
        if (cell_col)
            C[i] = SFX(mk4)(cell_col[4 * tet], cell_col[4 * tet + 1], cell_col[4 * tet + 2], cell_col[4 * tet + 3]);
        else
            C[i] = SFX(mk4)(col[4 * (int64_t)g], col[4 * (int64_t)g + 1], col[4 * (int64_t)g + 2], col[4 * (int64_t)g + 3]);
    }
}

/* accumulateConst: Common/ForwardCommon.glsl:37-40 (alpha = 1 - exp(-a t), colour alpha * c). */
static SFX(v4) SFX(acc_const)(R t, SFX(v4) c, R a) {
    R A = (R)1.0 - REXP_(-a * t);
    return SFX(mk4)(A * c.x, A * c.y, A * c.z, A);
}

/* Segment length per unit length along the ray.  INTERSECTION: the Euclidean distance between the entry and
 * exit points (IntersectRasterization.glsl:211), factor 1.  PPLL: the difference of the two faces' linear eye
 * depths (PPLL/Forward.glsl:53,188), i.e. the distance along the view axis: -(row 2 of the view matrix) . d. */
static R SFX(seg_scale)(const SFX(cam) * c, SFX(v3) rd, int linear_depth) {
    if (!linear_depth) return (R)1.0;
    const R* v = c->view;
    return (((-v[2]) * rd.x + (-v[6]) * rd.y) + (-v[10]) * rd.z);
}

/* One fragment of the INTERSECTION forward: IntersectRasterization.glsl:146-241 (segment length scaled by
 * tscale, seg_scale above). */
/* Ray parameter of the hit with face (p0, p1, p2): rayTriangleIntersect's t (RayIntersectionTests.glsl:30-76). */
static R SFX(face_t)(SFX(v3) ro, SFX(v3) rd, SFX(v3) p0, SFX(v3) p1, SFX(v3) p2) {
    SFX(v3) N = SFX(cross3)(SFX(sub3)(p1, p0), SFX(sub3)(p2, p0));
    R d = -SFX(dot3)(N, p0);
    return -(SFX(dot3)(N, ro) + d) / SFX(dot3)(N, rd);
}

/* dec (may be NULL): {hit, entry face, exit face, entry clamped to 0} decided elsewhere (the FP32 decisions of
 * an FP64 evaluation); the distances are then this precision's t of those faces. */
static void SFX(frag_eval)(SFX(v3) ro, SFX(v3) rd, const SFX(v3) P[4], const SFX(v4) C[4], R att, R tscale,
                           int per_cell, const int* dec, SFX(frag) * F) {
    memset(F, 0, sizeof(*F));
    F->tscale = tscale;
    F->per_cell = per_cell;
    if (dec) {
        F->hit = dec[0];
        if (!F->hit) return;
        F->f0 = dec[1];
        F->f1 = dec[2];
        F->t0 = dec[3] ? (R)0 : SFX(face_t)(ro, rd, P[ORC_FT[F->f0][0]], P[ORC_FT[F->f0][1]], P[ORC_FT[F->f0][2]]);
        F->t1 = SFX(face_t)(ro, rd, P[ORC_FT[F->f1][0]], P[ORC_FT[F->f1][1]], P[ORC_FT[F->f1][2]]);
    } else {
        F->hit = SFX(ray_tet)(ro, rd, P, &F->f0, &F->f1, &F->t0, &F->t1);
    }
    if (!F->hit) return;
    F->P0 = SFX(add3)(ro, SFX(scl3)(rd, F->t0));
    F->P1 = SFX(add3)(ro, SFX(scl3)(rd, F->t1));
    const int* a = ORC_FT[F->f0];
    R u, v, w;
    SFX(bary)(P[a[0]], P[a[1]], P[a[2]], F->P0, &u, &v);
    w = (R)1.0 - u - v;
    F->u0 = u; F->v0 = v;
    F->col0 = SFX(mk4)((C[a[0]].x * u + C[a[1]].x * v) + C[a[2]].x * w, (C[a[0]].y * u + C[a[1]].y * v) + C[a[2]].y * w,
                       (C[a[0]].z * u + C[a[1]].z * v) + C[a[2]].z * w, (C[a[0]].w * u + C[a[1]].w * v) + C[a[2]].w * w);
    const int* b = ORC_FT[F->f1];
    SFX(bary)(P[b[0]], P[b[1]], P[b[2]], F->P1, &u, &v);
    w = (R)1.0 - u - v;
    F->u1 = u; F->v1 = v;
    F->col1 = SFX(mk4)((C[b[0]].x * u + C[b[1]].x * v) + C[b[2]].x * w, (C[b[0]].y * u + C[b[1]].y * v) + C[b[2]].y * w,
                       (C[b[0]].z * u + C[b[1]].z * v) + C[b[2]].z * w, (C[b[0]].w * u + C[b[1]].w * v) + C[b[2]].w * w);
    R t = F->t1 - F->t0;
    if (per_cell) {
        /* IntersectRasterization.glsl:250-255: the tet's colour, accumulateConst over the segment length */
        F->col0 = F->col1 = C[0];
        F->a[0] = C[0].w * att;
        F->state[4] = SFX(acc_const)(t * tscale, C[0], F->a[0]);
        return;
    }
    R tSeg = (t / (R)4.0) * tscale;
    SFX(v4) rc = SFX(mk4)(0, 0, 0, 0);
    F->state[0] = rc;
    for (int s = 0; s < 4; s++) {
        R fb = (R)s * (R)0.25, fm = ((R)s + (R)0.5) * (R)0.25, fe = ((R)s + (R)1.0) * (R)0.25;
        SFX(v3) c0 = SFX(mk3)(SFX(mixr)(F->col0.x, F->col1.x, fb), SFX(mixr)(F->col0.y, F->col1.y, fb), SFX(mixr)(F->col0.z, F->col1.z, fb));
        SFX(v3) c1 = SFX(mk3)(SFX(mixr)(F->col0.x, F->col1.x, fe), SFX(mixr)(F->col0.y, F->col1.y, fe), SFX(mixr)(F->col0.z, F->col1.z, fe));
        R al = SFX(mixr)(F->col0.w, F->col1.w, fm);
        F->a[s] = al * att;
        F->sc0[s] = c0; F->sc1[s] = c1;
        SFX(v4) cur = SFX(acc_lin_const)(tSeg, c0, c1, al * att);
        F->seg[s] = cur;
        rc.x = rc.x + ((R)1.0 - rc.w) * cur.x;
        rc.y = rc.y + ((R)1.0 - rc.w) * cur.y;
        rc.z = rc.z + ((R)1.0 - rc.w) * cur.z;
        rc.w = rc.w + ((R)1.0 - rc.w) * cur.w;
        F->state[s + 1] = rc;
    }
}

#if IS_DOUBLE_
/* The FP32 evaluation's ray-tet decisions for tet Pt seen along the FP32 ray (camf, rdf); with dec_pos (float
 * positions [V*3], may be NULL) the tet's vertices are taken from there instead (decisions frozen at other
 * positions, as for differences of the function along one branch of its discrete choices). */
static const int* decide_f32(const cam_f* camf, v3_f rdf, const v3_d Pt[4], const float* dec_pos,
                             const int32_t vid[4], int dec[4]) {
    v3_f Pf[4];
    for (int i = 0; i < 4; i++) {
        if (dec_pos)
            Pf[i] = mk3_f(dec_pos[3 * (int64_t)vid[i]], dec_pos[3 * (int64_t)vid[i] + 1], dec_pos[3 * (int64_t)vid[i] + 2]);
        else
            Pf[i] = mk3_f((float)Pt[i].x, (float)Pt[i].y, (float)Pt[i].z);
    }
    v3_f rof = mk3_f(camf->camPos[0], camf->camPos[1], camf->camPos[2]);
    float t0, t1;
    dec[0] = ray_tet_f(rof, rdf, Pf, &dec[1], &dec[2], &t0, &t1);
    dec[3] = t0 == 0.0f;
    return dec;
}
#endif

/* Forward over the per-pixel fragment lists.  Front-to-back 'over' equals the reference's back-to-front
 * premultiplied blend (TetMeshRendererIntersection.cpp:258-276); the clear colour is blended last.
 * Early ray termination follows PPLL/Forward.glsl:168-173 (checked before each fragment). */
static void SFX(forward)(const orc_frags* fr, const int32_t* cells, const R* pos, const R* col, const R* cell_col,
                         const R* camu, R att,
                         const R clear[4], int use_ert, R ert_alpha, int linear_depth, int f32_decisions,
                         const float* dec_pos, R* out_img, R* out_C, int32_t* out_cnt) {
    SFX(cam) cam;
    SFX(cam_load)(&cam, camu, fr->W, fr->H);
#if IS_DOUBLE_
    float uff[68];
    for (int i = 0; i < 68; i++) uff[i] = (float)camu[i];
    cam_f camf;
    cam_load_f(&camf, uff, fr->W, fr->H);
#else
    (void)f32_decisions;
#endif
    SFX(v3) ro = SFX(mk3)(cam.camPos[0], cam.camPos[1], cam.camPos[2]);
    const int64_t P = (int64_t)fr->W * fr->H;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t px = 0; px < P; px++) {
        int x = (int)(px % fr->W), y = (int)(px / fr->W);
        SFX(v3) rd = SFX(ray_dir)(&cam, x, y);
        const R tscale = SFX(seg_scale)(&cam, rd, linear_depth);
#if IS_DOUBLE_
        const v3_f rdf = ray_dir_f(&camf, x, y);
#endif
        SFX(v4) rc = SFX(mk4)(0, 0, 0, 0);
        int32_t cnt = 0;
        for (int64_t k = fr->off[px]; k < fr->off[px + 1]; k++) {
            if (use_ert && rc.w > ert_alpha) break;
            int32_t id = fr->ids[k];
            SFX(v3) Pt[4]; SFX(v4) Ct[4]; int32_t vid[4];
            SFX(load_tet)(cells, pos, col, cell_col, id >> 2, Pt, Ct, vid);
            SFX(frag) F;
            int dec[4];
            const int* decp = NULL;
#if IS_DOUBLE_
            if (f32_decisions) decp = decide_f32(&camf, rdf, Pt, dec_pos, vid, dec);
#else
            (void)dec;
            (void)dec_pos;
#endif
            SFX(frag_eval)(ro, rd, Pt, Ct, att, tscale, cell_col != NULL, decp, &F);
            cnt++;
            if (!F.hit) continue;
            SFX(v4) f = F.state[4];
            rc.x = rc.x + ((R)1.0 - rc.w) * f.x;
            rc.y = rc.y + ((R)1.0 - rc.w) * f.y;
            rc.z = rc.z + ((R)1.0 - rc.w) * f.z;
            rc.w = rc.w + ((R)1.0 - rc.w) * f.w;
        }
        if (out_C) { out_C[4 * px] = rc.x; out_C[4 * px + 1] = rc.y; out_C[4 * px + 2] = rc.z; out_C[4 * px + 3] = rc.w; }
        if (out_cnt) out_cnt[px] = cnt;
        R T = (R)1.0 - rc.w;
        out_img[4 * px] = rc.x + T * clear[0];
        out_img[4 * px + 1] = rc.y + T * clear[1];
        out_img[4 * px + 2] = rc.z + T * clear[2];
        out_img[4 * px + 3] = rc.w + T * clear[3];
    }
}

/* ---------------------------------------------------------------------------------------------------------
 * Exact adjoint.  Derivative of the forward above (with the barycentric floor inactive), computed from stored
 * forward states (no compositing inversion), so it is an independent check of the inversion-based HIP adjoint.
 *
 * For one fragment with entry face f0 (plane normal N0, hit P0, face barycentrics mu), exit face f1, ray
 * direction d, segment length t = t1 - t0, colour gradients dC0, dC1 and dL/dt, the position gradient of a
 * vertex j of f0 is mu_j(P0) * Q0 with
 *     Q0 = -dLdt * N0/(N0.d) - [ G0^T dC0 - N0 * (dC0 . G0 d)/(N0.d) ],
 * G0 the in-plane colour gradient of face f0 (4x3); symmetrically Q1 = +dLdt*N1/(N1.d) - [...] for f1.
 * Moving a face vertex by delta moves the entry distance by mu_j (N.delta)/(N.d) and the colour at the hit
 * point by -mu_j G (delta - d (N.delta)/(N.d)); the two terms above are those chain rules.
 * ------------------------------------------------------------------------------------------------------- */

/* In-plane gradient of the linear interpolant of colours (ca, cb, cc) over triangle (a, b, c): g[ch][xyz]. */
static void SFX(face_grad)(SFX(v3) a, SFX(v3) b, SFX(v3) c, SFX(v4) ca, SFX(v4) cb, SFX(v4) cc, R g[4][3]) {
    SFX(v3) e1 = SFX(sub3)(b, a), e2 = SFX(sub3)(c, a);
    SFX(v3) n = SFX(cross3)(e1, e2);
    R nn = SFX(dot3)(n, n);
    SFX(v3) gb = SFX(scl3)(SFX(cross3)(e2, n), (R)1.0 / nn);
    SFX(v3) gc = SFX(scl3)(SFX(cross3)(n, e1), (R)1.0 / nn);
    R db[4] = {cb.x - ca.x, cb.y - ca.y, cb.z - ca.z, cb.w - ca.w};
    R dc[4] = {cc.x - ca.x, cc.y - ca.y, cc.z - ca.z, cc.w - ca.w};
    for (int ch = 0; ch < 4; ch++) {
        g[ch][0] = db[ch] * gb.x + dc[ch] * gc.x;
        g[ch][1] = db[ch] * gb.y + dc[ch] * gc.y;
        g[ch][2] = db[ch] * gb.z + dc[ch] * gc.z;
    }
}

static SFX(v3) SFX(face_normal)(SFX(v3) p0, SFX(v3) p1, SFX(v3) p2) {
    return SFX(cross3)(SFX(sub3)(p1, p0), SFX(sub3)(p2, p0));
}

/* Q = sgn * dLdt * N/(N.d) - [ G^T dC - N (dC . G d)/(N.d) ] */
static SFX(v3) SFX(face_q)(SFX(v3) N, SFX(v3) d, R g[4][3], const R dC[4], R sgn_dLdt) {
    R nd = SFX(dot3)(N, d);
    R gtdc[3] = {0, 0, 0};
    R dcgd = 0;
    for (int ch = 0; ch < 4; ch++) {
        gtdc[0] += g[ch][0] * dC[ch];
        gtdc[1] += g[ch][1] * dC[ch];
        gtdc[2] += g[ch][2] * dC[ch];
        dcgd += dC[ch] * (g[ch][0] * d.x + g[ch][1] * d.y + g[ch][2] * d.z);
    }
    R s = (sgn_dLdt + dcgd) / nd;
    return SFX(mk3)(N.x * s - gtdc[0], N.y * s - gtdc[1], N.z * s - gtdc[2]);
}

/* Backward through one fragment given dL/dF (F = composite of the four sub-segments).  Adds per-vertex
 * gradient contributions into gp[4][3], gc[4][4] (tet-local), one contribution per face corner. */
static void SFX(frag_backward_exact)(const SFX(frag) * F, SFX(v3) rd, const SFX(v3) Pt[4], const SFX(v4) Ct[4], R att,
                                    const R dF[4], int use_abs, R gp[4][3], R gc[4][4]) {
    if (F->per_cell) {
        /* accumulateConst: F = (alpha c, alpha), alpha = 1 - A, A = exp(-a t t_s): dL/dc = alpha dF.rgb,
         * dL/dalpha = dF.a + dF.rgb . c, dalpha/dt = a A t_s, dalpha/d(density) = att t t_s A.  The tet's colour
         * gradient goes to gc[0]; the positions move only the hit points, with no colour gradient in the face. */
        R tt = (F->t1 - F->t0) * F->tscale;
        R a = F->a[0];
        R A = REXP_(-a * tt);
        SFX(v4) c = F->col0;
        R sa = dF[3] + (dF[0] * c.x + dF[1] * c.y + dF[2] * c.z);
        R dcell[4] = {((R)1.0 - A) * dF[0], ((R)1.0 - A) * dF[1], ((R)1.0 - A) * dF[2], att * tt * A * sa};
        for (int ch = 0; ch < 4; ch++) gc[0][ch] += use_abs ? RABS_(dcell[ch]) : dcell[ch];
        R dLdt = a * A * sa * F->tscale;
        R zero_g[4][3] = {{0}};
        R zero_dc[4] = {0, 0, 0, 0};
        for (int side = 0; side < 2; side++) {
            const int* fv = ORC_FT[side == 0 ? F->f0 : F->f1];
            R u = side == 0 ? F->u0 : F->u1;
            R v = side == 0 ? F->v0 : F->v1;
            R mu[3] = {u, v, (R)1.0 - u - v};
            SFX(v3) N = SFX(face_normal)(Pt[fv[0]], Pt[fv[1]], Pt[fv[2]]);
            SFX(v3) Q = SFX(face_q)(N, rd, zero_g, zero_dc, side == 0 ? -dLdt : dLdt);
            for (int k = 0; k < 3; k++) {
                R cgp[3] = {mu[k] * Q.x, mu[k] * Q.y, mu[k] * Q.z};
                for (int ch = 0; ch < 3; ch++) gp[fv[k]][ch] += use_abs ? RABS_(cgp[ch]) : cgp[ch];
            }
        }
        return;
    }
    R dR[4] = {dF[0], dF[1], dF[2], dF[3]};
    R dcol0[4] = {0, 0, 0, 0}, dcol1[4] = {0, 0, 0, 0};
    R dLdt = 0;
    R t = F->t1 - F->t0;
    R tSeg = (t / (R)4.0) * F->tscale;
    for (int s = 3; s >= 0; s--) {
        SFX(v4) sg = F->seg[s];
        SFX(v4) st = F->state[s];
        R omA = (R)1.0 - st.w;
        R dseg[4] = {omA * dR[0], omA * dR[1], omA * dR[2], omA * dR[3]};
        R dprev_a = dR[3] * ((R)1.0 - sg.w) - (dR[0] * sg.x + dR[1] * sg.y + dR[2] * sg.z);
        dR[3] = dprev_a;
        R a = F->a[s];
        SFX(v3) c0 = F->sc0[s], c1 = F->sc1[s];
        R dc0[3] = {0, 0, 0}, dc1[3] = {0, 0, 0}, da = 0, dts = 0;
        if (!(a < (R)1e-6)) {
            R A = REXP_(-a * tSeg);
            R ia = (R)1.0 / a;
            R K = (tSeg + ia) * A - ia;
            R dKda = ia * ia * ((R)1.0 - A) - tSeg * tSeg * A - tSeg * ia * A;
            R cc0[3] = {c0.x, c0.y, c0.z}, cc1[3] = {c1.x, c1.y, c1.z};
            for (int ch = 0; ch < 3; ch++) {
                R diff = cc0[ch] - cc1[ch];
                dts += dseg[ch] * (a * A * cc0[ch] - a * tSeg * A * diff);
                da += dseg[ch] * (tSeg * A * cc0[ch] + dKda * diff);
                dc0[ch] = (((R)1.0 - A) + K) * dseg[ch];
                dc1[ch] = -K * dseg[ch];
            }
            dts += dseg[3] * a * A;
            da += dseg[3] * tSeg * A;
        }
        R fb = (R)s * (R)0.25, fm = ((R)s + (R)0.5) * (R)0.25, fe = ((R)s + (R)1.0) * (R)0.25;
        for (int ch = 0; ch < 3; ch++) {
            dcol0[ch] += ((R)1.0 - fb) * dc0[ch] + ((R)1.0 - fe) * dc1[ch];
            dcol1[ch] += fb * dc0[ch] + fe * dc1[ch];
        }
        dcol0[3] += ((R)1.0 - fm) * att * da;
        dcol1[3] += fm * att * da;
        dLdt += dts / (R)4.0;
    }
    dLdt *= F->tscale;  /* d tSeg / d t1 = tscale / 4 */
    /* colour and position gradients per face corner */
    for (int side = 0; side < 2; side++) {
        int f = side == 0 ? F->f0 : F->f1;
        const int* fv = ORC_FT[f];
        R u = side == 0 ? F->u0 : F->u1;
        R v = side == 0 ? F->v0 : F->v1;
        R mu[3] = {u, v, (R)1.0 - u - v};
        const R* dC = side == 0 ? dcol0 : dcol1;
        R g[4][3];
        SFX(face_grad)(Pt[fv[0]], Pt[fv[1]], Pt[fv[2]], Ct[fv[0]], Ct[fv[1]], Ct[fv[2]], g);
        SFX(v3) N = SFX(face_normal)(Pt[fv[0]], Pt[fv[1]], Pt[fv[2]]);
        SFX(v3) Q = SFX(face_q)(N, rd, g, dC, side == 0 ? -dLdt : dLdt);
        for (int k = 0; k < 3; k++) {
            R cgc[4] = {mu[k] * dC[0], mu[k] * dC[1], mu[k] * dC[2], mu[k] * dC[3]};
            R cgp[3] = {mu[k] * Q.x, mu[k] * Q.y, mu[k] * Q.z};
            for (int ch = 0; ch < 4; ch++) gc[fv[k]][ch] += use_abs ? RABS_(cgc[ch]) : cgc[ch];
            for (int ch = 0; ch < 3; ch++) gp[fv[k]][ch] += use_abs ? RABS_(cgp[ch]) : cgp[ch];
        }
    }
}

/* tet >= 0: per-cell colours, gc[0] is the tet's colour gradient (gcol [T*4]). */
static void SFX(scatter_tet_grads)(const int32_t vid[4], R gp[4][3], R gc[4][4], int64_t tet, R* gpos, R* gcol) {
    if (tet >= 0) {
        for (int ch = 0; ch < 4; ch++) {
            R val = gc[0][ch];
#pragma omp atomic
            gcol[4 * tet + ch] += val;
        }
    }
    for (int i = 0; i < 4; i++) {
        int64_t g = vid[i];
        for (int ch = 0; ch < 3; ch++) {
            R val = gp[i][ch];
#pragma omp atomic
            gpos[3 * g + ch] += val;
        }
        if (tet >= 0) continue;
        for (int ch = 0; ch < 4; ch++) {
            R val = gc[i][ch];
#pragma omp atomic
            gcol[4 * g + ch] += val;
        }
    }
}

/* Exact adjoint over all pixels.  gpos/gcol are accumulated into (like the reference's .grad buffers,
 * TetMesh.cpp:243-249).  dimg is dL/d(blended image). */
static void SFX(adjoint_exact)(const orc_frags* fr, const int32_t* cells, const R* pos, const R* col, const R* cell_col,
                               const R* camu, R att,
                               const R clear[4], int use_ert, R ert_alpha, int linear_depth, int f32_decisions,
                               const R* dimg, int use_abs, R* gpos, R* gcol) {
    SFX(cam) cam;
    SFX(cam_load)(&cam, camu, fr->W, fr->H);
#if IS_DOUBLE_
    float uff[68];
    for (int i = 0; i < 68; i++) uff[i] = (float)camu[i];
    cam_f camf;
    cam_load_f(&camf, uff, fr->W, fr->H);
#else
    (void)f32_decisions;
#endif
    SFX(v3) ro = SFX(mk3)(cam.camPos[0], cam.camPos[1], cam.camPos[2]);
    const int64_t P = (int64_t)fr->W * fr->H;
#pragma omp parallel
    {
        int64_t cap = 0;
        SFX(v4)* states = NULL;
#pragma omp for schedule(dynamic, 64)
        for (int64_t px = 0; px < P; px++) {
            int64_t n = fr->off[px + 1] - fr->off[px];
            if (n == 0) continue;
            if (n + 1 > cap) { cap = 2 * (n + 1); states = (SFX(v4)*)realloc(states, cap * sizeof(SFX(v4))); }
            int x = (int)(px % fr->W), y = (int)(px / fr->W);
            SFX(v3) rd = SFX(ray_dir)(&cam, x, y);
            const R tscale = SFX(seg_scale)(&cam, rd, linear_depth);
            int dec[4];
            const int* decp = NULL;
#if IS_DOUBLE_
            const v3_f rdf = ray_dir_f(&camf, x, y);
#else
            (void)dec;
#endif
            /* forward, storing the composite state before every fragment */
            SFX(v4) rc = SFX(mk4)(0, 0, 0, 0);
            int64_t m = 0;
            for (int64_t k = 0; k < n; k++) {
                if (use_ert && rc.w > ert_alpha) break;
                states[k] = rc;
                SFX(v3) Pt[4]; SFX(v4) Ct[4]; int32_t vid[4];
                SFX(load_tet)(cells, pos, col, cell_col, fr->ids[fr->off[px] + k] >> 2, Pt, Ct, vid);
                SFX(frag) F;
#if IS_DOUBLE_
                if (f32_decisions) decp = decide_f32(&camf, rdf, Pt, NULL, vid, dec);
#endif
                SFX(frag_eval)(ro, rd, Pt, Ct, att, tscale, cell_col != NULL, decp, &F);
                m = k + 1;
                if (!F.hit) continue;
                SFX(v4) f = F.state[4];
                rc.x = rc.x + ((R)1.0 - rc.w) * f.x;
                rc.y = rc.y + ((R)1.0 - rc.w) * f.y;
                rc.z = rc.z + ((R)1.0 - rc.w) * f.z;
                rc.w = rc.w + ((R)1.0 - rc.w) * f.w;
            }
            /* blend onto the clear colour: img = C + (1 - C.a) * clear */
            const R* di = dimg + 4 * px;
            R dC[4] = {di[0], di[1], di[2], di[3] * ((R)1.0 - clear[3]) - (di[0] * clear[0] + di[1] * clear[1] + di[2] * clear[2])};
            for (int64_t k = m - 1; k >= 0; k--) {
                SFX(v3) Pt[4]; SFX(v4) Ct[4]; int32_t vid[4];
                SFX(load_tet)(cells, pos, col, cell_col, fr->ids[fr->off[px] + k] >> 2, Pt, Ct, vid);
                SFX(frag) F;
#if IS_DOUBLE_
                if (f32_decisions) decp = decide_f32(&camf, rdf, Pt, NULL, vid, dec);
#endif
                SFX(frag_eval)(ro, rd, Pt, Ct, att, tscale, cell_col != NULL, decp, &F);
                if (!F.hit) continue;
                SFX(v4) ci = states[k];
                SFX(v4) f = F.state[4];
                R omA = (R)1.0 - ci.w;
                R dF[4] = {omA * dC[0], omA * dC[1], omA * dC[2], omA * dC[3]};
                dC[3] = dC[3] * ((R)1.0 - f.w) - (dC[0] * f.x + dC[1] * f.y + dC[2] * f.z);
                R gp[4][3] = {{0}}, gc[4][4] = {{0}};
                SFX(frag_backward_exact)(&F, rd, Pt, Ct, att, dF, use_abs, gp, gc);
                SFX(scatter_tet_grads)(vid, gp, gc, cell_col ? (int64_t)(fr->ids[fr->off[px] + k] >> 2) : -1, gpos,
                                       gcol);
            }
        }
        free(states);
    }
}

/* ---------------------------------------------------------------------------------------------------------
 * Reference adjoint, restated as written (quirks kept): AdjointIntersectRasterization.glsl:111-362 with
 * Common/BackwardCommon.glsl:107-251.  Sub-segment loop s = 3..1 (s = 0 skipped, :204), segment-length
 * adjoint of |P0 - P1| through barycentric coordinates (:139-181), barycentric adjoint with the hit point held
 * fixed and clamped denominators (:183-251).  Pixels are visited back to front, as the interlocked raster pass
 * does.  Used to report the deviation of the reference's gradient, not as the parity target of "exact".
 * ------------------------------------------------------------------------------------------------------- */

/* accumulateLinearConstAdjoint: BackwardCommon.glsl:116-133. */
static void SFX(ref_alc_adjoint)(R t, R tf, SFX(v3) c0, SFX(v3) c1, R a, R A, const R dC[4], R* dOut_dt, R dc0[3],
                                 R dc1[3], R* da) {
    R cc0[3] = {c0.x, c0.y, c0.z}, cc1[3] = {c1.x, c1.y, c1.z};
    R dd = 0;
    for (int ch = 0; ch < 3; ch++) dd += dC[ch] * ((a * cc0[ch] + (a * cc1[ch] - a * cc0[ch]) * t) * A);
    *dOut_dt += tf * (dd + dC[3] * a * A);
    if (a < (R)1e-6) {
        R s = 0;
        for (int ch = 0; ch < 3; ch++) { dc0[ch] = 0; dc1[ch] = 0; s += dC[ch] * (t * cc0[ch] + (R)0.5 * t * t * (cc1[ch] - cc0[ch])); }
        *da = s + dC[3] * t;
    } else {
        R inva = (R)1.0 / a;
        R s = 0;
        for (int ch = 0; ch < 3; ch++) {
            dc0[ch] = (((R)1.0 - A) + ((t + inva) * A - inva)) * dC[ch];
            dc1[ch] = -((t + inva) * A - inva) * dC[ch];
            s += dC[ch] * (t * A * cc0[ch] + (inva * inva - (t * t + inva * inva + t * inva) * A) * (cc0[ch] - cc1[ch]));
        }
        *da = s + dC[3] * t * A;
    }
}

/* d|cross(a,b)|^2 helpers for the barycentric adjoint with the hit point fixed. */
static R SFX(sq)(R x) { return x * x; }

/* Gradients of u = sqrt(f1)/sqrt(f0) and v = sqrt(f2)/sqrt(f0) w.r.t. p0, p1, p2 (P fixed), with the
 * reference's clamps denom0 = max(f0^1.5, 1e-6), denom1 = max(sqrt(f1), 1e-6), denom2 = max(sqrt(f2), 1e-6)
 * (BackwardCommon.glsl:207-230).  f0 = |(p0-p2) x (p1-p2)|^2, f1 = |(P-p1) x (p1-p2)|^2, f2 = |(P-p0) x (p0-p2)|^2. */
static void SFX(ref_bary_grads)(SFX(v3) P, SFX(v3) p0, SFX(v3) p1, SFX(v3) p2, SFX(v3) du[3], SFX(v3) dv[3]) {
    SFX(v3) A0 = SFX(sub3)(p0, p2), B0 = SFX(sub3)(p1, p2);
    SFX(v3) n0 = SFX(cross3)(A0, B0);
    SFX(v3) A1 = SFX(sub3)(P, p1), B1 = SFX(sub3)(p1, p2);
    SFX(v3) n1 = SFX(cross3)(A1, B1);
    SFX(v3) A2 = SFX(sub3)(P, p0), B2 = SFX(sub3)(p0, p2);
    SFX(v3) n2 = SFX(cross3)(A2, B2);
    R f0 = SFX(dot3)(n0, n0), f1 = SFX(dot3)(n1, n1), f2 = SFX(dot3)(n2, n2);
    R den0 = SFX(maxr)((R)pow((double)f0, 1.5), (R)1e-6);
    R den1 = SFX(maxr)(RSQRT_(f1), (R)1e-6);
    R den2 = SFX(maxr)(RSQRT_(f2), (R)1e-6);
    /* gradient of |A x B|^2 w.r.t. A is 2 (B x (A x B)), w.r.t. B is 2 ((A x B) x A) */
    SFX(v3) gA0 = SFX(cross3)(B0, n0), gB0 = SFX(cross3)(n0, A0);   /* halves of df0/dA0, df0/dB0 */
    SFX(v3) gA1 = SFX(cross3)(B1, n1), gB1 = SFX(cross3)(n1, A1);
    SFX(v3) gA2 = SFX(cross3)(B2, n2), gB2 = SFX(cross3)(n2, A2);
    /* half-derivatives of f0: A0 = p0 - p2, B0 = p1 - p2 */
    SFX(v3) hf0[3] = {gA0, gB0, SFX(scl3)(SFX(add3)(gA0, gB0), (R)-1)};
    /* f1: A1 = P - p1, B1 = p1 - p2 -> p0: none, p1: -gA1 + gB1, p2: -gB1 */
    SFX(v3) z = SFX(mk3)(0, 0, 0);
    SFX(v3) hf1[3] = {z, SFX(sub3)(gB1, gA1), SFX(scl3)(gB1, (R)-1)};
    /* f2: A2 = P - p0, B2 = p0 - p2 -> p0: -gA2 + gB2, p1: none, p2: -gB2 */
    SFX(v3) hf2[3] = {SFX(sub3)(gB2, gA2), z, SFX(scl3)(gB2, (R)-1)};
    /* u = sqrt(f1)/sqrt(f0): du = (hf1 * f0 - hf0 * f1) / (sqrt(f1) f0^1.5); the p0 row simplifies to
     * -hf0 * sqrt(f1) / f0^1.5 (the reference writes it with denom1 in the numerator). */
    du[0] = SFX(scl3)(hf0[0], -den1 / den0);
    dv[1] = SFX(scl3)(hf0[1], -den2 / den0);
    for (int k = 1; k < 3; k++) du[k] = SFX(scl3)(SFX(sub3)(SFX(scl3)(hf1[k], f0), SFX(scl3)(hf0[k], f1)), (R)1.0 / (den1 * den0));
    for (int k = 0; k < 3; k += 2) dv[k] = SFX(scl3)(SFX(sub3)(SFX(scl3)(hf2[k], f0), SFX(scl3)(hf0[k], f2)), (R)1.0 / (den2 * den0));
}

static void SFX(adjoint_reference)(const orc_frags* fr, const int32_t* cells, const R* pos, const R* col, const R* camu,
                                   R att, const R* img, const R* dimg, int use_abs, R* gpos, R* gcol) {
    SFX(cam) cam;
    SFX(cam_load)(&cam, camu, fr->W, fr->H);
    SFX(v3) ro = SFX(mk3)(cam.camPos[0], cam.camPos[1], cam.camPos[2]);
    const int64_t P = (int64_t)fr->W * fr->H;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t px = 0; px < P; px++) {
        int64_t n = fr->off[px + 1] - fr->off[px];
        if (n == 0) continue;
        int x = (int)(px % fr->W), y = (int)(px / fr->W);
        SFX(v3) rd = SFX(ray_dir)(&cam, x, y);
        R cro[4] = {img[4 * px], img[4 * px + 1], img[4 * px + 2], img[4 * px + 3]};
        R dro[4] = {dimg[4 * px], dimg[4 * px + 1], dimg[4 * px + 2], dimg[4 * px + 3]};
        for (int64_t k = n - 1; k >= 0; k--) {
            SFX(v3) Pt[4]; SFX(v4) Ct[4]; int32_t vid[4];
            SFX(load_tet)(cells, pos, col, NULL, fr->ids[fr->off[px] + k] >> 2, Pt, Ct, vid);
            SFX(frag) F;
            SFX(frag_eval)(ro, rd, Pt, Ct, att, (R)1.0, 0, NULL, &F);
            if (!F.hit) continue;
            const int* a0 = ORC_FT[F.f0];
            const int* a1 = ORC_FT[F.f1];
            R t = F.t1 - F.t0, tSeg = t / (R)4.0, dOut_dt = 0;
            R dcf0[4] = {0, 0, 0, 0}, dcf1[4] = {0, 0, 0, 0};
            for (int s = 3; s > 0; s--) {
                R fb = (R)s * (R)0.25, fm = ((R)s + (R)0.5) * (R)0.25, fe = ((R)s + (R)1.0) * (R)0.25;
                SFX(v3) c0 = F.sc0[s], c1 = F.sc1[s];
                R al = SFX(mixr)(F.col0.w, F.col1.w, fm);
                R A = REXP_(-(al * att) * tSeg);
                SFX(v4) acc = SFX(acc_lin_const)(tSeg, c0, c1, al * att);
                R ain = (acc.w - cro[3]) / (acc.w - (R)1.0);
                R cin[3] = {cro[0] - ((R)1.0 - ain) * acc.x, cro[1] - ((R)1.0 - ain) * acc.y, cro[2] - ((R)1.0 - ain) * acc.z};
                R dacc[4] = {((R)1.0 - ain) * dro[0], ((R)1.0 - ain) * dro[1], ((R)1.0 - ain) * dro[2], ((R)1.0 - ain) * dro[3]};
                R anew = dro[3] * ((R)1.0 - acc.w) - (dro[0] * acc.x + dro[1] * acc.y + dro[2] * acc.z);
                dro[3] = anew;
                cro[0] = cin[0]; cro[1] = cin[1]; cro[2] = cin[2]; cro[3] = ain;
                R dc0[3], dc1[3], da;
                SFX(ref_alc_adjoint)(tSeg, (R)0.25, c0, c1, al * att, A, dacc, &dOut_dt, dc0, dc1, &da);
                da *= att;
                for (int ch = 0; ch < 3; ch++) {
                    dcf0[ch] += ((R)1.0 - fb) * dc0[ch] + ((R)1.0 - fe) * dc1[ch];
                    dcf1[ch] += fb * dc0[ch] + fe * dc1[ch];
                }
                dcf0[3] += ((R)1.0 - fm) * da;
                dcf1[3] += fm * da;
            }
            /* segmentLengthAdjoint */
            SFX(v3) q[2][3] = {{Pt[a0[0]], Pt[a0[1]], Pt[a0[2]]}, {Pt[a1[0]], Pt[a1[1]], Pt[a1[2]]}};
            R uu[2] = {F.u0, F.u1}, vv[2] = {F.v0, F.v1};
            R ww[2] = {(R)1.0 - F.u0 - F.v0, (R)1.0 - F.u1 - F.v1};
            SFX(v3) D = SFX(sub3)(SFX(add3)(SFX(add3)(SFX(scl3)(q[0][0], uu[0]), SFX(scl3)(q[0][1], vv[0])), SFX(scl3)(q[0][2], ww[0])),
                                  SFX(add3)(SFX(add3)(SFX(scl3)(q[1][0], uu[1]), SFX(scl3)(q[1][1], vv[1])), SFX(scl3)(q[1][2], ww[1])));
            R invDen = (R)1.0 / RSQRT_(SFX(maxr)(SFX(dot3)(D, D), (R)1e-6));
            R dudv[2][2];
            SFX(v3) gp6[2][3];
            for (int sd = 0; sd < 2; sd++) {
                R sgn = sd == 0 ? (R)1 : (R)-1;
                dudv[sd][0] = dOut_dt * (sgn * SFX(dot3)(SFX(sub3)(q[sd][0], q[sd][2]), D) * invDen);
                dudv[sd][1] = dOut_dt * (sgn * SFX(dot3)(SFX(sub3)(q[sd][1], q[sd][2]), D) * invDen);
                R wts[3] = {uu[sd], vv[sd], ww[sd]};
                for (int k = 0; k < 3; k++) gp6[sd][k] = SFX(scl3)(D, dOut_dt * sgn * wts[k] * invDen);
            }
            /* baryAdjoint for both faces */
            R gc6[2][3][4];
            for (int sd = 0; sd < 2; sd++) {
                const int* fv = sd == 0 ? a0 : a1;
                const R* dcf = sd == 0 ? dcf0 : dcf1;
                SFX(v3) Pp = sd == 0 ? F.P0 : F.P1;
                SFX(v4) c0 = Ct[fv[0]], c1 = Ct[fv[1]], c2 = Ct[fv[2]];
                R du = dudv[sd][0] + (dcf[0] * (c0.x - c2.x) + dcf[1] * (c0.y - c2.y) + dcf[2] * (c0.z - c2.z) + dcf[3] * (c0.w - c2.w));
                R dv = dudv[sd][1] + (dcf[0] * (c1.x - c2.x) + dcf[1] * (c1.y - c2.y) + dcf[2] * (c1.z - c2.z) + dcf[3] * (c1.w - c2.w));
                R wts[3] = {uu[sd], vv[sd], ww[sd]};
                for (int k = 0; k < 3; k++)
                    for (int ch = 0; ch < 4; ch++) gc6[sd][k][ch] = dcf[ch] * wts[k];
                SFX(v3) gu[3], gv[3];
                SFX(ref_bary_grads)(Pp, q[sd][0], q[sd][1], q[sd][2], gu, gv);
                for (int k = 0; k < 3; k++) gp6[sd][k] = SFX(add3)(gp6[sd][k], SFX(add3)(SFX(scl3)(gu[k], du), SFX(scl3)(gv[k], dv)));
            }
            R gp[4][3] = {{0}}, gc[4][4] = {{0}};
            for (int sd = 0; sd < 2; sd++) {
                const int* fv = sd == 0 ? a0 : a1;
                for (int k = 0; k < 3; k++) {
                    R pp[3] = {gp6[sd][k].x, gp6[sd][k].y, gp6[sd][k].z};
                    for (int ch = 0; ch < 3; ch++) gp[fv[k]][ch] += use_abs ? RABS_(pp[ch]) : pp[ch];
                    for (int ch = 0; ch < 4; ch++) gc[fv[k]][ch] += use_abs ? RABS_(gc6[sd][k][ch]) : gc6[sd][k][ch];
                }
            }
            SFX(scatter_tet_grads)(vid, gp, gc, -1, gpos, gcol);
        }
    }
}

/* Tet-quality regulariser gradient, TetRegularizer.glsl:108-166 restated: V, l_rms, quality
 * q = 6 sqrt(2) V / l_rms^3, dL/dq = -lambda / (1 + exp(beta q)) (mirrored softplus), dq/dp from
 * T0 = 36 sqrt(3) dot(p0-p3, cross(p0-p1, p0-p2)), T1 = sum of squared edge lengths, T2 = max(T1^{5/2}, 1e-5). */
static void SFX(tet_regularizer)(const int32_t* cells, int64_t T, const R* pos, R lambda, R beta, R* gpos) {
    for (int64_t t = 0; t < T; t++) {
        SFX(v3) p[4];
        int32_t vid[4];
        for (int i = 0; i < 4; i++) {
            vid[i] = cells[4 * t + i];
            p[i] = SFX(mk3)(pos[3 * (int64_t)vid[i]], pos[3 * (int64_t)vid[i] + 1], pos[3 * (int64_t)vid[i] + 2]);
        }
        R V = -SFX(dot3)(SFX(cross3)(SFX(sub3)(p[1], p[0]), SFX(sub3)(p[2], p[0])), SFX(sub3)(p[3], p[0])) / (R)6.0;
        R T1 = 0;
        for (int i = 0; i < 4; i++)
            for (int j = i + 1; j < 4; j++) { SFX(v3) e = SFX(sub3)(p[i], p[j]); T1 += SFX(dot3)(e, e); }
        R lrms = RSQRT_(T1 / (R)6.0);
        R quality = (R)6.0 * RSQRT_((R)2.0) * V / (lrms * lrms * lrms);
        R dLdq = -lambda / ((R)1.0 + REXP_(beta * quality));
        R T0 = (R)62.353829072479584 * SFX(dot3)(SFX(sub3)(p[0], p[3]), SFX(cross3)(SFX(sub3)(p[0], p[1]), SFX(sub3)(p[0], p[2])));
        R T2 = SFX(maxr)((R)pow((double)T1, 2.5), (R)1e-5);
        /* second term: 12 sqrt(3) * (-dD/dp_i) * T1 with D = dot(cross(p1-p0, p2-p0), p3-p0) */
        SFX(v3) a = SFX(sub3)(p[1], p[0]), b = SFX(sub3)(p[2], p[0]), c = SFX(sub3)(p[3], p[0]);
        SFX(v3) dD[4];
        dD[1] = SFX(cross3)(b, c); dD[2] = SFX(cross3)(c, a); dD[3] = SFX(cross3)(a, b);
        dD[0] = SFX(scl3)(SFX(add3)(SFX(add3)(dD[1], dD[2]), dD[3]), (R)-1.0);
        SFX(v3) sum = SFX(add3)(SFX(add3)(p[0], p[1]), SFX(add3)(p[2], p[3]));
        for (int i = 0; i < 4; i++) {
            SFX(v3) lin = SFX(sub3)(sum, SFX(scl3)(p[i], (R)4.0));  /* (-3 p_i + sum of the others) */
            SFX(v3) dq = SFX(scl3)(SFX(add3)(SFX(scl3)(lin, T0), SFX(scl3)(dD[i], (R)-20.784609690826528 * T1)), (R)1.0 / T2);
            gpos[3 * (int64_t)vid[i]] += dLdq * dq.x;
            gpos[3 * (int64_t)vid[i] + 1] += dLdq * dq.y;
            gpos[3 * (int64_t)vid[i] + 2] += dLdq * dq.z;
        }
    }
}

/* Loss of the regulariser: sum over tets of (lambda / beta) log(1 + exp(-beta q)). */
static R SFX(tet_regularizer_loss)(const int32_t* cells, int64_t T, const R* pos, R lambda, R beta) {
    R loss = 0;
    for (int64_t t = 0; t < T; t++) {
        SFX(v3) p[4];
        for (int i = 0; i < 4; i++) {
            int32_t v = cells[4 * t + i];
            p[i] = SFX(mk3)(pos[3 * (int64_t)v], pos[3 * (int64_t)v + 1], pos[3 * (int64_t)v + 2]);
        }
        R V = -SFX(dot3)(SFX(cross3)(SFX(sub3)(p[1], p[0]), SFX(sub3)(p[2], p[0])), SFX(sub3)(p[3], p[0])) / (R)6.0;
        R T1 = 0;
        for (int i = 0; i < 4; i++)
            for (int j = i + 1; j < 4; j++) { SFX(v3) e = SFX(sub3)(p[i], p[j]); T1 += SFX(dot3)(e, e); }
        R lrms = RSQRT_(T1 / (R)6.0);
        R q = (R)6.0 * RSQRT_((R)2.0) * V / (lrms * lrms * lrms);
        loss += lambda / beta * (R)log1p(exp(-(double)(beta * q)));
    }
    return loss;
}
