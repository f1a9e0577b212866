/*
 * CPU oracle: Projection renderer (projected tetrahedra), continuous part, instantiated from oracle.c with
 * R = float and R = double like oracle_cont.h.  TEST INFRASTRUCTURE ONLY: nothing in the product links it.
 *
 * The discrete choices of a tet (culling, which two edges' projections cross, the class-1 flip) are made in
 * float by pt_decide (oracle.c) and handed in as a pt_dec; this file computes the values that follow from them
 * in precision R, operation by operation in the order of the reference expressions:
 *   triangle vertices and thickness     Data/Shaders/Projection/GenerateTrianglesVTK.glsl:129-403
 *   per-fragment colour and opacity     Data/Shaders/Projection/ProjectedRasterization.glsl:85-105
 *     (vertices written with w = 1, so the rasteriser's attribute interpolation is linear in screen space)
 */

/* GetCorrectedDepth without the linear depth correction, which the renderer never enables
 * (GenerateTrianglesVTK.glsl:83-126, TetMeshRendererProjection.cpp:800-808): the eye-space distance between the
 * unprojections of (x, y, z1) and (x, y, z2). */
static R SFX(pt_corrected_depth)(const R* iP, R x, R y, R z1, R z2) {
    R e1[4], e2[4];
    for (int r = 0; r < 4; r++) {
        e1[r] = ((iP[0 * 4 + r] * x + iP[1 * 4 + r] * y) + iP[2 * 4 + r] * z1) + iP[3 * 4 + r] * (R)1.0;
        e2[r] = ((iP[0 * 4 + r] * x + iP[1 * 4 + r] * y) + iP[2 * 4 + r] * z2) + iP[3 * 4 + r] * (R)1.0;
    }
    R d0 = e1[0] / e1[3] - e2[0] / e2[3];
    R d1 = e1[1] / e1[3] - e2[1] / e2[3];
    R d2 = e1[2] / e1[3] - e2[2] / e2[3];
    return RSQRT_((d0 * d0 + d1 * d1) + d2 * d2);
}

/* The five projected-tetrahedra vertices of a tet that D keeps (0-3: the tet's vertices, 4: the crossing of
 * the two edges' projections in class 2): NDC xyz, colour and thickness (density scale: attenuation times
 * eye-space length).  P: world positions, C: colours (per cell: the tet's colour four times, left unmodified
 * at the thick vertex, as the per-cell shader variant writes the cell colour to every corner). */
static void SFX(pt_vertices)(const pt_dec* D, const R* u, const R P[4][3], const R C[4][4], int per_cell, R att,
                             R ndc[5][3], R col[5][4], R dep[5]) {
    const R* VP = u;
    for (int i = 0; i < 4; i++) {
        R c[4];
        for (int r = 0; r < 4; r++)
            c[r] = ((VP[0 * 4 + r] * P[i][0] + VP[1 * 4 + r] * P[i][1]) + VP[2 * 4 + r] * P[i][2]) + VP[3 * 4 + r] * (R)1.0;
        for (int k = 0; k < 3; k++) ndc[i][k] = c[k] / c[3];
        for (int k = 0; k < 4; k++) col[i][k] = C[i][k];
        dep[i] = 0;
    }
    for (int k = 0; k < 3; k++) ndc[4][k] = 0;
    for (int k = 0; k < 4; k++) col[4][k] = 0;
    dep[4] = 0;
    int a0 = D->seg1[0], a1 = D->seg1[1];
    const int b0 = D->seg2[0], b1 = D->seg2[1];
    const R* P1 = ndc[a0];
    const R* P2 = ndc[a1];
    const R* P3 = ndc[b0];
    const R* P4 = ndc[b1];
    R A[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
    R B[3] = {P4[0] - P3[0], P4[1] - P3[1], P4[2] - P3[2]};
    R Cv[3] = {P3[0] - P1[0], P3[1] - P1[1], P3[2] - P1[2]};
    R den = A[0] * B[1] - A[1] * B[0];
    R alpha = (B[1] * Cv[0] - B[0] * Cv[1]) / den;
    R beta = (A[1] * Cv[0] - A[0] * Cv[1]) / den;
    const R* iP = u + 32;
    if (D->cls2) {
        for (int k = 0; k < 3; k++) ndc[4][k] = P1[k] + alpha * A[k];
        R depth = SFX(pt_corrected_depth)(iP, ndc[4][0], ndc[4][1], ndc[4][2], P3[2] + beta * B[2]);
        for (int k = 0; k < 4; k++)
            col[4][k] = per_cell ? C[0][k]
                                 : (R)0.5 * (((C[a0][k] + alpha * (C[a1][k] - C[a0][k])) + C[b0][k]) +
                                             beta * (C[b1][k] - C[b0][k]));
        dep[4] = depth * att;
    } else {
        if (D->flip) {  /* segment 1 reversed so that the thick point is its second vertex */
            int t = a0; a0 = a1; a1 = t;
            alpha = (R)1.0 - alpha;
            P1 = ndc[a0];
            P2 = ndc[a1];
        }
        R edgez = P3[2] + beta * B[2];
        R pointz = P1[2];
        R facez = (edgez + (alpha - (R)1.0) * pointz) / alpha;
        R depth = SFX(pt_corrected_depth)(iP, P2[0], P2[1], P2[2], facez);
        if (!per_cell) {
            for (int k = 0; k < 4; k++) {
                R edgec = C[b0][k] + beta * (C[b1][k] - C[b0][k]);
                R pointc = C[a0][k];
                R facec = (edgec + (alpha - (R)1.0) * pointc) / alpha;
                col[a1][k] = (R)0.5 * (facec + C[a1][k]);
            }
        }
        dep[a1] = depth * att;
    }
}

/* One fragment of a projected triangle: barycentrics of the pixel centre (px, py) (NDC) in the corners' NDC
 * xy, linearly interpolated colour and thickness, alpha = 1 - exp(-a d), premultiplied (ProjectedRasterization
 * .glsl:100-105).  Returns (rgb * alpha, alpha). */
static SFX(v4) SFX(pt_fragment)(const R X[3], const R Y[3], const R col[3][4], const R dep[3], R px, R py) {
    R D = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0]);
    R l0 = ((X[1] - px) * (Y[2] - py) - (X[2] - px) * (Y[1] - py)) / D;
    R l1 = ((X[2] - px) * (Y[0] - py) - (X[0] - px) * (Y[2] - py)) / D;
    R l2 = ((X[0] - px) * (Y[1] - py) - (X[1] - px) * (Y[0] - py)) / D;
    R c[4];
    for (int k = 0; k < 4; k++) c[k] = (l0 * col[0][k] + l1 * col[1][k]) + l2 * col[2][k];
    R d = (l0 * dep[0] + l1 * dep[1]) + l2 * dep[2];
    R a = (R)1.0 - REXP_(-c[3] * d);
    return SFX(mk4)(c[0] * a, c[1] * a, c[2] * a, a);
}
