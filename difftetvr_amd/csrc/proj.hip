// Projection renderer (projected tetrahedra) on gfx950.
//
// Each tet in view becomes 3 or 4 screen triangles, Shirley and Tuchman's classes, with the thickness at the
// thick vertex taken from the crossing of two edges' projections (GenerateTrianglesVTK.glsl:129-403).  The
// triangles go through the tile renderer's binning, radix sort and per-tile compositing: one (tile, triangle)
// record per 8x8 tile the triangle may cover, keyed by the NDC z of the triangle's centroid
// (ComputeTriangleDepths.glsl:99-108), composited front to back.  The reference blends the same list back to
// front with premultiplied alpha, which sums the same terms.  Each fragment contributes (c.rgb alpha, alpha),
// alpha = 1 - exp(-c.a d), where colour c and thickness d are interpolated linearly in screen space
// (ProjectedRasterization.glsl:100-105; the vertices carry w = 1).
//
// The adjoint walks each pixel back to front as the tet renderer's does (transmittance checkpoints and
// division).  It sums each record's gradients with respect to its triangle's corner NDC xy, colours and
// thicknesses (reduce-scatters within rows of 16 lanes, summed per batch in LDS, then one atomic per value into
// the triangle's slot).  A per-tet pass carries them back through the triangle construction to the vertex
// positions and colours.
#include "raster.h"
#include "render_common.h"

namespace dtv {

// ---------------------------------------------------------------------------------------------------------
// Decisions of one tet.  pt_info word: bits 0-2 triangle count (0, 3, 4), bit 3 class 2, bit 4 flip, then the
// two edges seg1[0], seg1[1], seg2[0], seg2[1] (2 bits each, bits 5-12).
struct PtDec {
    int ntri;
    int s10, s11, s20, s21;
    bool cls2, flip;
};

__device__ __forceinline__ uint32_t pt_pack(const PtDec& D) {
    return (uint32_t)D.ntri | (D.cls2 ? 8u : 0u) | (D.flip ? 16u : 0u) | ((uint32_t)D.s10 << 5) |
           ((uint32_t)D.s11 << 7) | ((uint32_t)D.s20 << 9) | ((uint32_t)D.s21 << 11);
}
__device__ __forceinline__ PtDec pt_unpack(uint32_t w) {
    PtDec D;
    D.ntri = (int)(w & 7u);
    D.cls2 = (w >> 3) & 1u;
    D.flip = (w >> 4) & 1u;
    D.s10 = (int)((w >> 5) & 3u);
    D.s11 = (int)((w >> 7) & 3u);
    D.s20 = (int)((w >> 9) & 3u);
    D.s21 = (int)((w >> 11) & 3u);
    return D;
}

// Corner j of triangle k: projected-tetrahedra vertex 0-3 (the tet's) or 4 (the crossing); the triangle fans
// of GenerateTrianglesVTK.glsl:334-400.
__device__ __forceinline__ int pt_corner(const PtDec& D, int k, int j) {
    const int a0 = D.flip ? D.s11 : D.s10, a1 = D.flip ? D.s10 : D.s11;
    if (D.cls2) {  // fan {4, a0, b0, a1, b1, a0}
        const int i = j == 0 ? 0 : k + j;
        return i == 0 ? 4 : (i == 1 || i == 5) ? a0 : i == 2 ? D.s20 : i == 3 ? a1 : D.s21;
    }
    const int i = j == 0 ? 0 : k + j;  // fan {a1, a0, b0, b1, a0}
    return i == 0 ? a1 : (i == 1 || i == 4) ? a0 : i == 2 ? D.s20 : D.s21;
}

// Row i (0-3) of a small register array by selects: a run-time index would place the array in scratch memory.
// The rows pass through an empty asm statement that keeps them in registers, as the compiler otherwise turns
// the select chain back into an indexed load from a stack copy.
template <int R, int N>
__device__ __forceinline__ void pick_row(const float (&a)[R][N], int i, float (&out)[N]) {
#pragma unroll
    for (int k = 0; k < N; k++) {
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            v[q] = a[q][k];
            asm volatile("" : "+v"(v[q]));
        }
        float r = v[0];
#pragma unroll
        for (int q = 1; q < 4; q++) r = i == q ? v[q] : r;
        out[k] = r;
    }
}

// Float decisions (GenerateTrianglesVTK.glsl:140-273).  Beyond the reference: a tet with a vertex behind the
// near plane or beyond the guard band is culled (its triangles would be undefined there).
__device__ __forceinline__ PtDec pt_decide(const ViewUniforms& U, const float P[4][3]) {
    PtDec D{};
    D.ntri = 0;
    float ndc[4][3];
    bool cull = false;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float c[4];
#pragma unroll
        for (int r = 0; r < 4; r++)
            c[r] = ((U.viewProj[0 * 4 + r] * P[i][0] + U.viewProj[1 * 4 + r] * P[i][1]) +
                    U.viewProj[2 * 4 + r] * P[i][2]) + U.viewProj[3 * 4 + r] * 1.0f;
        if (!(c[3] >= U.znear)) cull = true;
#pragma unroll
        for (int k = 0; k < 3; k++) ndc[i][k] = c[k] / c[3];
        if (!(fabsf(ndc[i][0]) <= kGuard) || !(fabsf(ndc[i][1]) <= kGuard)) cull = true;
    }
    if (cull) return D;
    bool gx = true, lx = true, gy = true, ly = true, gz = true, lz = false;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        gx = gx && ndc[i][0] > 1.0f;
        lx = lx && ndc[i][0] < -1.0f;
        gy = gy && ndc[i][1] > 1.0f;
        ly = ly && ndc[i][1] < -1.0f;
        gz = gz && ndc[i][2] > 1.0f;
        lz = lz || ndc[i][2] < 0.0f;
    }
    if (gx || lx || gy || ly || gz || lz) return D;
    const float v1x = ndc[1][0] - ndc[0][0], v1y = ndc[1][1] - ndc[0][1];
    const float v2x = ndc[2][0] - ndc[0][0], v2y = ndc[2][1] - ndc[0][1];
    const float v3x = ndc[3][0] - ndc[0][0], v3y = ndc[3][1] - ndc[0][1];
    const float fd1 = v3x * v2y - v3y * v2x;
    const float fd2 = v1x * v3y - v1y * v3x;
    const float fd3 = v2x * v1y - v2y * v1x;
    if ((fd1 * fd2 >= 0.0f) && ((fd1 != 0.0f) || (fd2 != 0.0f))) {
        D.s10 = 0; D.s11 = 3; D.s20 = 1; D.s21 = 2;
    } else if (fd1 * fd3 >= 0.0f) {
        D.s10 = 0; D.s11 = 2; D.s20 = 1; D.s21 = 3;
    } else {
        D.s10 = 0; D.s11 = 1; D.s20 = 2; D.s21 = 3;
    }
    float n1[3], n2[3], n3[3];  // segment 1 is (0, s11), segment 2 (s20, s21)
    pick_row(ndc, D.s11, n1);
    pick_row(ndc, D.s20, n2);
    pick_row(ndc, D.s21, n3);
    const float Ax = n1[0] - ndc[0][0], Ay = n1[1] - ndc[0][1];
    const float Bx = n3[0] - n2[0], By = n3[1] - n2[1];
    const float Cx = n2[0] - ndc[0][0], Cy = n2[1] - ndc[0][1];
    const float den = Ax * By - Ay * Bx;
    if (fabsf(den) < 1e-8f) return D;  // degenerate
    const float alpha = (By * Cx - Bx * Cy) / den;
    D.cls2 = (alpha >= 0.0f) && (alpha <= 1.0f);
    D.flip = !D.cls2 && alpha <= 0.0f;
    D.ntri = D.cls2 ? 4 : 3;
    return D;
}

// ---------------------------------------------------------------------------------------------------------
// Values of the five vertices (NDC xyz, colour, thickness) given the decisions, in the decisions' arithmetic.
struct PtVerts {
    float ndc[5][3];
    float col[5][4];
    float dep[5];
};

// GetCorrectedDepth without the linear depth correction, which the renderer never enables
// (GenerateTrianglesVTK.glsl:83-126, TetMeshRendererProjection.cpp:800-808).
__device__ __forceinline__ float pt_corrected_depth(const float* iP, float x, float y, float z1, float z2) {
    float e1[4], e2[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        e1[r] = ((iP[0 * 4 + r] * x + iP[1 * 4 + r] * y) + iP[2 * 4 + r] * z1) + iP[3 * 4 + r] * 1.0f;
        e2[r] = ((iP[0 * 4 + r] * x + iP[1 * 4 + r] * y) + iP[2 * 4 + r] * z2) + iP[3 * 4 + r] * 1.0f;
    }
    const float d0 = e1[0] / e1[3] - e2[0] / e2[3];
    const float d1 = e1[1] / e1[3] - e2[1] / e2[3];
    const float d2 = e1[2] / e1[3] - e2[2] / e2[3];
    return sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
}

__device__ __forceinline__ void pt_vertices(const PtDec& D, const ViewUniforms& U, const float P[4][3],
                                            const float C[4][4], bool per_cell, PtVerts& V) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float c[4];
#pragma unroll
        for (int r = 0; r < 4; r++)
            c[r] = ((U.viewProj[0 * 4 + r] * P[i][0] + U.viewProj[1 * 4 + r] * P[i][1]) +
                    U.viewProj[2 * 4 + r] * P[i][2]) + U.viewProj[3 * 4 + r] * 1.0f;
#pragma unroll
        for (int k = 0; k < 3; k++) V.ndc[i][k] = c[k] / c[3];
#pragma unroll
        for (int k = 0; k < 4; k++) V.col[i][k] = C[i][k];
        V.dep[i] = 0.f;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) V.ndc[4][k] = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) V.col[4][k] = 0.f;
    V.dep[4] = 0.f;
    // the two edges' end points (segment 1: a0 -> a1, segment 2: b0 -> b1)
    float na0[3], na1[3], nb0[3], nb1[3], ca0[4], ca1[4], cb0[4], cb1[4];
    {
        float nd[4][3];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int k = 0; k < 3; k++) nd[i][k] = V.ndc[i][k];
        pick_row(nd, D.s10, na0);
        pick_row(nd, D.s11, na1);
        pick_row(nd, D.s20, nb0);
        pick_row(nd, D.s21, nb1);
        float cc[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int k = 0; k < 4; k++) cc[i][k] = C[i][k];
        pick_row(cc, D.s10, ca0);
        pick_row(cc, D.s11, ca1);
        pick_row(cc, D.s20, cb0);
        pick_row(cc, D.s21, cb1);
    }
    float A[3], B[3], Cv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        A[k] = na1[k] - na0[k];
        B[k] = nb1[k] - nb0[k];
        Cv[k] = nb0[k] - na0[k];
    }
    const float den = A[0] * B[1] - A[1] * B[0];
    float alpha = (B[1] * Cv[0] - B[0] * Cv[1]) / den;
    const float beta = (A[1] * Cv[0] - A[0] * Cv[1]) / den;
    if (D.cls2) {
#pragma unroll
        for (int k = 0; k < 3; k++) V.ndc[4][k] = na0[k] + alpha * A[k];
        const float depth = pt_corrected_depth(U.invProj, V.ndc[4][0], V.ndc[4][1], V.ndc[4][2], nb0[2] + beta * B[2]);
#pragma unroll
        for (int k = 0; k < 4; k++)
            V.col[4][k] = per_cell ? C[0][k]
                                   : 0.5f * (((ca0[k] + alpha * (ca1[k] - ca0[k])) + cb0[k]) +
                                             beta * (cb1[k] - cb0[k]));
        V.dep[4] = depth * U.attenuation;
    } else {
        // the thick vertex T is segment 1's second end (its first when the segment is reversed), P the other
        const bool fl = D.flip;
        const int t = fl ? D.s10 : D.s11;
        if (fl) alpha = 1.0f - alpha;
        float nT[3], nP[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            nT[k] = fl ? na0[k] : na1[k];
            nP[k] = fl ? na1[k] : na0[k];
        }
        const float edgez = nb0[2] + beta * B[2];
        const float pointz = nP[2];
        const float facez = (edgez + (alpha - 1.0f) * pointz) / alpha;
        const float depth = pt_corrected_depth(U.invProj, nT[0], nT[1], nT[2], facez);
        const float dep = depth * U.attenuation;
#pragma unroll
        for (int v = 0; v < 4; v++) V.dep[v] = v == t ? dep : V.dep[v];
        if (!per_cell) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float edgec = cb0[k] + beta * (cb1[k] - cb0[k]);
                const float pointc = fl ? ca1[k] : ca0[k];
                const float facec = (edgec + (alpha - 1.0f) * pointc) / alpha;
                const float col = 0.5f * (facec + (fl ? ca0[k] : ca1[k]));
#pragma unroll
                for (int v = 0; v < 4; v++) V.col[v][k] = v == t ? col : V.col[v][k];
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
// Per-tet state between the passes: pt_vert[5 * t + v] = {NDC x, y, z, thickness}, pt_col[5 * t + v] = colour.

__device__ __forceinline__ void pt_load_tet(const BatchItem& B, int64_t tet, float P[4][3], float C[4][4],
                                            int32_t vid[4]) {
    const int4 idx = *reinterpret_cast<const int4*>(B.cells + 4 * tet);
    vid[0] = idx.x;
    vid[1] = idx.y;
    vid[2] = idx.z;
    vid[3] = idx.w;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float3_ p = load_pos(B.pos, vid[i]);
        P[i][0] = p.x;
        P[i][1] = p.y;
        P[i][2] = p.z;
        const float4 c = B.cell_col ? *reinterpret_cast<const float4*>(B.cell_col + 4 * tet)
                                    : *reinterpret_cast<const float4*>(B.col + 4 * (int64_t)vid[i]);
        C[i][0] = c.x;
        C[i][1] = c.y;
        C[i][2] = c.z;
        C[i][3] = c.w;
    }
}

// Corners of triangle k from the tet's five vertices held in registers (selects: a run-time index would place
// the array in scratch memory).
__device__ __forceinline__ void pt_tri_corners(const float4 (&pv)[5], const PtDec& D, int k, float4 (&q)[3]) {
    float x[5], y[5], z[5];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        x[i] = pv[i].x;
        y[i] = pv[i].y;
        z[i] = pv[i].z;
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int v = pt_corner(D, k, j);
        float rx = x[0], ry = y[0], rz = z[0];
#pragma unroll
        for (int i = 1; i < 5; i++) {
            rx = v == i ? x[i] : rx;
            ry = v == i ? y[i] : ry;
            rz = v == i ? z[i] : rz;
        }
        q[j] = make_float4(rx, ry, rz, 0.f);  // pt_screen reads x, y, z
    }
}

// Fixed-point screen triangle (oriented to positive area) and centroid key of the triangle with corners qc
// (NDC x, y, z).
__device__ __forceinline__ void pt_screen(const ViewUniforms& U, const float4 (&qc)[3], ScreenTri& S, uint32_t& key) {
    const float hw = 0.5f * (float)U.width, hh = 0.5f * (float)U.height;
    float z[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const float4 q = qc[j];
        const float xf = q.x * hw + hw;
        const float yf = q.y * hh + hh;
        S.X[j] = (int32_t)rintf(xf * 256.0f);
        S.Y[j] = (int32_t)rintf(yf * 256.0f);
        z[j] = q.z;
    }
    key = __float_as_uint(((z[0] + z[1]) + z[2]) / 3.0f);
    int64_t area = (int64_t)(S.X[1] - S.X[0]) * (S.Y[2] - S.Y[0]) - (int64_t)(S.Y[1] - S.Y[0]) * (S.X[2] - S.X[0]);
    if (area < 0) {
        const int32_t tx = S.X[1], ty = S.Y[1];
        S.X[1] = S.X[2];
        S.Y[1] = S.Y[2];
        S.X[2] = tx;
        S.Y[2] = ty;
        area = -area;
    }
    S.area2 = area;
}

// One thread per tet slot: decisions, the five vertices, and the record count of its triangles.
__global__ __launch_bounds__(256) void k_pt_count(const BatchItem* __restrict__ items, int nitems,
                                                   uint32_t* __restrict__ pt_info, float4* __restrict__ pt_vert,
                                                   float4* __restrict__ pt_col, uint32_t* __restrict__ counts,
                                                   uint4* __restrict__ block_stats) {
    const BatchItem& B = items[item_of_block(items, nitems, blockIdx.x)];
    const ViewUniforms& U = B.U;
    const int64_t tet = (int64_t)(blockIdx.x - B.block0) * 256 + threadIdx.x;
    uint32_t c = 0, ntri = 0, kand = 0xFFFFFFFFu, kor = 0u;
    if (tet < B.T) {
        float P[4][3], C[4][4];
        int32_t vid[4];
        pt_load_tet(B, tet, P, C, vid);
        const PtDec D = pt_decide(U, P);
        const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;  // the tet's slot in the batch
        pt_info[g] = pt_pack(D);
        if (D.ntri) {
            PtVerts V;
            pt_vertices(D, U, P, C, B.cell_col != nullptr, V);
            float4 pv[5];
#pragma unroll
            for (int v = 0; v < 5; v++) {
                pv[v] = make_float4(V.ndc[v][0], V.ndc[v][1], V.ndc[v][2], V.dep[v]);
                pt_vert[5 * g + v] = pv[v];
                pt_col[5 * g + v] = make_float4(V.col[v][0], V.col[v][1], V.col[v][2], V.col[v][3]);
            }
            for (int k = 0; k < D.ntri; k++) {
                ScreenTri S;
                uint32_t key;
                float4 q[3];
                pt_tri_corners(pv, D, k, q);
                pt_screen(U, q, S, key);
                const uint32_t cf = face_record_count(screen_bins(U, S));
                c += cf;
                ntri++;
                if (cf) {
                    kand &= key;
                    kor |= key;
                }
            }
        }
    }
    counts[(int64_t)blockIdx.x * 256 + threadIdx.x] = c;
    __shared__ uint4 s_part[4];
    const uint32_t wtri = wave_sum_u32(ntri), wand = wave_and_u32(kand), wor = wave_or_u32(kor);
    if (lane_id() == 0) s_part[threadIdx.x / 64] = make_uint4(wtri, wand, wor, 0u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint4 r = s_part[0];
        for (int w = 1; w < (int)(blockDim.x / 64); w++) {
            r.x += s_part[w].x;
            r.y &= s_part[w].y;
            r.z |= s_part[w].z;
        }
        block_stats[blockIdx.x] = r;
    }
}

// Same grid: the tet's records at offsets[slot], triangles in order, tiles row-major; vals = tet * 4 + k.
__global__ __launch_bounds__(256) void k_pt_emit(const BatchItem* __restrict__ items, int nitems,
                                                  const uint32_t* __restrict__ pt_info,
                                                  const float4* __restrict__ pt_vert,
                                                  const uint32_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ counts, uint64_t* __restrict__ keys,
                                                  uint32_t* __restrict__ vals, uint64_t cap,
                                                  uint32_t* __restrict__ overflow) {
    const BatchItem& B = items[item_of_block(items, nitems, blockIdx.x)];
    const ViewUniforms& U = B.U;
    const int64_t tet = (int64_t)(blockIdx.x - B.block0) * 256 + threadIdx.x;
    if (tet >= B.T) return;
    const int64_t slot = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t left = counts[slot];
    if (left == 0) return;
    uint32_t o = offsets[slot];
    const PtDec D = pt_unpack(pt_info[slot]);
    float4 pv[5];
#pragma unroll
    for (int v = 0; v < 5; v++) pv[v] = pt_vert[5 * slot + v];
    for (int k = 0; k < D.ntri && left > 0; k++) {
        ScreenTri S;
        uint32_t key;
        float4 q[3];
        pt_tri_corners(pv, D, k, q);
        pt_screen(U, q, S, key);
        const FaceBins b = screen_bins(U, S);
        const bool test = test_tiles(b);
        const uint32_t id = (uint32_t)(tet * 4 + k);
        for (int ty = b.ty0; ty <= b.ty1 && left > 0; ty++)
            for (int tx = b.tx0; tx <= b.tx1 && left > 0; tx++) {
                if (test && !tile_may_cover(b.S, tx, ty, b.shift)) continue;
                if (o < cap) {
                    keys[o] = ((uint64_t)(B.tile0 + (uint32_t)(ty * U.tiles_x + tx)) << 32) | (uint64_t)key;
                    vals[o] = id;
                } else {
                    *overflow = 1u;
                }
                o++;
                left--;
            }
    }
}

// Sorted record i: out[2i + 1] = {triangle id, coverage mask low, high, corner slots} (out[2i] is not used);
// corner slots: the tet's vertex slot (0-4) of triangle corner j in bits 3 j .. 3 j + 2.
__global__ __launch_bounds__(256) void k_pt_gather(const BatchItem* __restrict__ items, int nitems,
                                                    const uint32_t* __restrict__ pt_info,
                                                    const float4* __restrict__ pt_vert,
                                                    const SortRoute* __restrict__ route, int64_t n_cap,
                                                    const uint64_t* __restrict__ n_dev, uint4* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= run_count(n_cap, n_dev)) return;
    const uint32_t id = route->vals_out[i];
    const uint32_t gtile = (uint32_t)(route->keys_out[i] >> 32);
    const BatchItem& B = items[item_of_tile(items, nitems, gtile)];
    const ViewUniforms& U = B.U;
    const int tile = (int)(gtile - B.tile0);
    const int64_t g = (int64_t)B.block0 * 256 + (id >> 2);
    const PtDec D = pt_unpack(pt_info[g]);
    ScreenTri S;
    uint32_t key;
    float4 q[3];
#pragma unroll
    for (int j = 0; j < 3; j++) q[j] = pt_vert[5 * g + pt_corner(D, (int)(id & 3), j)];
    pt_screen(U, q, S, key);
    const uint64_t m = tile_coverage(S, tile % U.tiles_x, tile / U.tiles_x, U.width, U.height);
    const int k = (int)(id & 3);
    const uint32_t slots = (uint32_t)pt_corner(D, k, 0) | ((uint32_t)pt_corner(D, k, 1) << 3) |
                           ((uint32_t)pt_corner(D, k, 2) << 6);
    out[2 * i + 1] = make_uint4(id, (uint32_t)m, (uint32_t)(m >> 32), slots);
}

// ---------------------------------------------------------------------------------------------------------
// Tile compositing.  A batch of 32 records is staged in LDS: each triangle's corner NDC xy, colours,
// thicknesses, 1 / (twice its NDC area) and coverage.
struct PtRecLds {
    float X[3][kBatch], Y[3][kBatch];
    float c[12][kBatch];  // corner j colour channel ch at [4 j + ch]
    float d[3][kBatch];
    float invD[kBatch];
    uint32_t mask[2][kBatch];
    uint32_t id[kBatch];
};

// A record's triangle corners as fetched from the per-tet state: NDC xy and thickness (q), colour (c).
struct PtRecData {
    float4 q[3], c[3];
};

// slot0: the item's first tet slot in the batch (block0 * 256).
__device__ __forceinline__ void pt_fetch(const float4* __restrict__ pt_vert, const float4* __restrict__ pt_col,
                                         int64_t slot0, uint4 rec, PtRecData& R) {
    const int64_t tet = slot0 + (rec.x >> 2);
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const int v = (int)((rec.w >> (3 * j)) & 7u);
        R.q[j] = pt_vert[5 * tet + v];
        R.c[j] = pt_col[5 * tet + v];
    }
}

__device__ __forceinline__ void pt_store(int64_t slot0, uint4 rec, const PtRecData& R, PtRecLds& L, int i) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
        L.X[j][i] = R.q[j].x;
        L.Y[j][i] = R.q[j].y;
        L.d[j][i] = R.q[j].w;
        L.c[4 * j + 0][i] = R.c[j].x;
        L.c[4 * j + 1][i] = R.c[j].y;
        L.c[4 * j + 2][i] = R.c[j].z;
        L.c[4 * j + 3][i] = R.c[j].w;
    }
    L.invD[i] = 1.0f / ((R.q[1].x - R.q[0].x) * (R.q[2].y - R.q[0].y) - (R.q[2].x - R.q[0].x) * (R.q[1].y - R.q[0].y));
    L.mask[0][i] = rec.y;
    L.mask[1][i] = rec.z;
    L.id[i] = (uint32_t)(4 * (slot0 + (rec.x >> 2))) + (rec.x & 3u);  // the triangle's slot in the batch
}

// One fragment: barycentrics of the pixel centre (px, py) (NDC), interpolated colour c and thickness d,
// A = exp(-c.a d), F = ((1 - A) c.rgb, A) (premultiplied RGB, transmittance).
struct PtFrag {
    float l[3];
    float c[4];
    float d;
    float A;
};
__device__ __forceinline__ void pt_eval(const PtRecLds& L, int j, float px, float py, PtFrag& f) {
    const float X0 = L.X[0][j], X1 = L.X[1][j], X2 = L.X[2][j];
    const float Y0 = L.Y[0][j], Y1 = L.Y[1][j], Y2 = L.Y[2][j];
    const float iD = L.invD[j];
    f.l[0] = ((X1 - px) * (Y2 - py) - (X2 - px) * (Y1 - py)) * iD;
    f.l[1] = ((X2 - px) * (Y0 - py) - (X0 - px) * (Y2 - py)) * iD;
    f.l[2] = ((X0 - px) * (Y1 - py) - (X1 - px) * (Y0 - py)) * iD;
    {
#pragma clang fp contract(fast)
#pragma unroll
        for (int ch = 0; ch < 4; ch++)
            f.c[ch] = (f.l[0] * L.c[ch][j] + f.l[1] * L.c[4 + ch][j]) + f.l[2] * L.c[8 + ch][j];
        f.d = (f.l[0] * L.d[0][j] + f.l[1] * L.d[1][j]) + f.l[2] * L.d[2][j];
    }
    f.A = __expf(-f.c[3] * f.d);
}

__device__ __forceinline__ bool pt_covers(const PtRecLds& L, int j, int lane) {
    return (L.mask[lane >> 5][j] >> (lane & 31)) & 1u;
}

constexpr float kPtTinyT = 0x1p-100f;  // transmittance below which a fragment's input is not recovered

// The forward's fragment list and round values.
constexpr int kPtRound = 128;  // fragments evaluated per round, two per lane
struct PtFragLds {
    uint16_t entry[kBatch * 64];  // (pixel lane << 5) | record index, ordered by (rank, pixel lane)
    float4 val[kPtRound];         // F = ((1 - A) c.rgb, A)
};

// Forward: per batch of 32 records the covered (pixel, record) pairs form a dense list ordered by (rank,
// pixel lane), rank r holding the r-th fragment of every pixel that has more than r; rounds of kPtRound
// fragments are evaluated across the lanes, then each pixel lane composites its fragments of the round front
// to back (about one per pixel per round).
__global__ __launch_bounds__(64) void k_pt_forward(const BatchItem* __restrict__ items, int nitems,
                                                    const uint32_t* __restrict__ order,
                                                    const float4* __restrict__ pt_vert,
                                                    const float4* __restrict__ pt_col, const uint4* __restrict__ recs,
                                                    const uint2* __restrict__ ranges, PixelState ps_all) {
    __shared__ PtRecLds L;
    __shared__ PtFragLds F;
    const uint32_t gtile = global_tile_of_block(order);
    const BatchItem& B = items[item_of_tile(items, nitems, gtile)];
    const ViewUniforms& U = B.U;
    float4* __restrict__ out_image = reinterpret_cast<float4*>(B.image);
    const PixelState ps = item_pixel_state(ps_all, B);
    const int tile = (int)(gtile - B.tile0);
    const int lane = threadIdx.x;
    const uint64_t lt = lanemask_lt();
    const int x0 = (tile % U.tiles_x) * kTileEdge, y0 = (tile / U.tiles_x) * kTileEdge;
    const int x = x0 + (lane & 7), y = y0 + (lane >> 3);
    const bool valid = x < U.width && y < U.height;
    const uint2 rg = ranges[gtile];
    const int P = U.width * U.height;
    const int pix = valid ? pixel_index(U, x, y) : 0;
    const int64_t slot0 = (int64_t)B.block0 * 256;
    float4 C = make_float4(0.f, 0.f, 0.f, 1.0f);  // premultiplied RGB, transmittance
    uint32_t n = 0, nck = 0, nlast = 0, last = rg.x;
    // Two-stage prefetch: while a batch is shaded, the next batch's corner data and the record words of the
    // batch after it are in flight.
    uint4 cur = make_uint4(0u, 0u, 0u, 0u), nxt = cur;
    PtRecData cd;
    if (lane < kBatch && rg.x + lane < rg.y) {
        cur = recs[2 * (int64_t)(rg.x + lane) + 1];
        pt_fetch(pt_vert, pt_col, slot0, cur, cd);
    }
    if (lane < kBatch && rg.x + kBatch + lane < rg.y) nxt = recs[2 * (int64_t)(rg.x + kBatch + lane) + 1];
    for (uint32_t base = rg.x; base < rg.y; base += kBatch) {
        const int cnt = (int)min((uint32_t)kBatch, rg.y - base);
        if (lane < cnt) pt_store(slot0, cur, cd, L, lane);
        if (lane < kBatch && base + kBatch + lane < rg.y) {
            cur = nxt;
            pt_fetch(pt_vert, pt_col, slot0, cur, cd);
        }
        if (lane < kBatch && base + 2 * kBatch + lane < rg.y) nxt = recs[2 * (int64_t)(base + 2 * kBatch + lane) + 1];
        __syncthreads();
        uint32_t mask = 0u;
        if (valid)
            for (int j = 0; j < cnt; j++) mask |= (uint32_t)pt_covers(L, j, lane) << j;
        // the coverage masks are consumed: their slots hold the list's rank masks
        uint64_t* rank_mask = reinterpret_cast<uint64_t*>(&L.mask[0][0]);
        const uint32_t total = build_fragment_list(F.entry, rank_mask, mask, false, lane);
        const uint32_t nmine = (uint32_t)__popc(mask);
        RankCursor cur_e;
        __syncthreads();
        for (uint32_t r0 = 0; r0 < total; r0 += kPtRound) {
            for (int i = lane; i < kPtRound; i += 64) {
                const uint32_t e = r0 + (uint32_t)i;
                if (e >= total) break;
                const uint32_t ent = F.entry[e];
                const int pl = (int)(ent >> 5), j = (int)(ent & 31u);
                const float px = 2.0f * (((float)(x0 + (pl & 7)) + 0.5f) / (float)U.width) - 1.0f;
                const float py = 2.0f * (((float)(y0 + (pl >> 3)) + 0.5f) / (float)U.height) - 1.0f;
                PtFrag f;
                pt_eval(L, j, px, py, f);
                const float a = 1.0f - f.A;
                F.val[i] = make_float4(a * f.c[0], a * f.c[1], a * f.c[2], f.A);
            }
            __syncthreads();
            uint32_t e;
            while (cur_e.next(rank_mask, nmine, lt, r0 + kPtRound, e)) {
                const float4 f = F.val[e - r0];
                const float Tout = C.w * f.w;
                // checkpoint rules of the tet renderer: before an underflow, and periodically
                if (C.w >= kPtTinyT && Tout < kPtTinyT && nlast != n && (int)nck < U.max_ckpt) {
                    ps.ckpt_t[(int64_t)nck * P + pix] = C.w;
                    ps.ckpt_n[(int64_t)nck * P + pix] = n;
                    nck++;
                    nlast = n;
                }
                {
#pragma clang fp contract(fast)
                    C.x = C.x + C.w * f.x;
                    C.y = C.y + C.w * f.y;
                    C.z = C.z + C.w * f.z;
                }
                C.w = Tout;
                n++;
                last = base + (F.entry[e] & 31u) + 1;
                if ((int)(n - nlast) >= U.ckpt_interval && C.w >= kPtTinyT && (int)nck < U.max_ckpt - 1) {
                    ps.ckpt_t[(int64_t)nck * P + pix] = C.w;
                    ps.ckpt_n[(int64_t)nck * P + pix] = n;
                    nck++;
                    nlast = n;
                }
            }
            __syncthreads();
        }
    }
    if (valid) {
        const float T = C.w;
        out_image[pix] = make_float4(C.x + T * U.clear[0], C.y + T * U.clear[1], C.z + T * U.clear[2],
                                     (1.0f - T) + T * U.clear[3]);
        ps.ray_color[pix] = C;
        ps.frag_count[pix] = n;
        ps.last_rec[pix] = last;
        ps.num_ckpt[pix] = nck;
    }
    const uint32_t tile_total = wave_inclusive_scan(n);
    if (lane == 63) ps.tile_frags[gtile] = tile_total;
}

// Adjoint, per pixel back to front.  The wave's lanes cover the tile as four 4x4 sub-tiles, one per row of 16
// lanes (sub-tile g = lane / 16 at (g & 1, g >> 1)).  Each row walks the records covering its sub-tile, so one
// step of the loop advances four records, one per row; a record covering several sub-tiles is met by each of
// their rows.  Per step every row sums its record's 21 gradient terms (corner j: -lambda_j G.x, -lambda_j G.y,
// lambda_j dL/dc, lambda_j dL/dd) by a reduce-scatter within the row into its own LDS copy of the batch's sums
// (plain adds: the rows never share a copy).  After the batch the four copies are added and each nonzero sum
// goes into tri_grad[21 * id] with one atomic, 21 consecutive words per record.  The loop itself issues no
// global atomics, so its transmittance checkpoint loads never wait on them.
__global__ __launch_bounds__(64) void k_pt_adjoint(const BatchItem* __restrict__ items,
                                                    const AdjointTarget* __restrict__ targets, int nitems,
                                                    const uint32_t* __restrict__ order,
                                                    const float4* __restrict__ pt_vert,
                                                    const float4* __restrict__ pt_col, const uint4* __restrict__ recs,
                                                    const uint2* __restrict__ ranges, PixelState ps_all,
                                                    float* __restrict__ tri_grad, int use_abs) {
    __shared__ PtRecLds L;
    __shared__ float s_acc[4][kBatch * 21];  // per row: the batch's sums, record j at [21 j]
    const uint32_t gtile = global_tile_of_block(order);
    const int item = item_of_tile(items, nitems, gtile);
    const BatchItem& B = items[item];
    const ViewUniforms& U = B.U;
    const float4* __restrict__ dimage = reinterpret_cast<const float4*>(targets[item].dimage);
    if (dimage == nullptr) return;
    const PixelState ps = item_pixel_state(ps_all, B);
    const int tile = (int)(gtile - B.tile0);
    const int lane = threadIdx.x;
    const int grp = lane >> 4;
    const int lx = 4 * (grp & 1) + (lane & 3), ly = 4 * (grp >> 1) + ((lane >> 2) & 3);
    const int cbit = 8 * ly + lx;  // the pixel's bit in a record's coverage mask
    const int x = (tile % U.tiles_x) * kTileEdge + lx, y = (tile / U.tiles_x) * kTileEdge + ly;
    const bool valid = x < U.width && y < U.height;
    const float px = 2.0f * (((float)x + 0.5f) / (float)U.width) - 1.0f;
    const float py = 2.0f * (((float)y + 0.5f) / (float)U.height) - 1.0f;
    const uint2 rg = ranges[gtile];
    const int P = U.width * U.height;
    const int pix = valid ? pixel_index(U, x, y) : 0;
    float Cw = 1.0f;
    float dC[4] = {0.f, 0.f, 0.f, 0.f};
    uint32_t n = 0, last = 0;
    int jck = -1;
    uint32_t ck_n = 0xFFFFFFFFu, ck_n2 = 0xFFFFFFFFu;
    float ck_t = 1.0f;
    if (valid) {
        Cw = ps.ray_color[pix].w;
        n = ps.frag_count[pix];
        last = ps.last_rec[pix];
        jck = (int)ps.num_ckpt[pix] - 1;
        if (jck >= 0) {
            ck_n = ps.ckpt_n[(int64_t)jck * P + pix];
            ck_t = ps.ckpt_t[(int64_t)jck * P + pix];
            if (jck >= 1) ck_n2 = ps.ckpt_n[(int64_t)(jck - 1) * P + pix];
        }
        const float4 di = dimage[pix];
        dC[0] = di.x;
        dC[1] = di.y;
        dC[2] = di.z;
        dC[3] = di.w * (1.0f - U.clear[3]) - (di.x * U.clear[0] + di.y * U.clear[1] + di.z * U.clear[2]);
    }
    if (ballot(valid && n > 0) == 0ull) return;
    // AbsGS-style absolute colour gradients: the per-cell variant only, as the reference's
    // (AdjointProjectedRasterization.glsl:214-222)
    const bool abs_col = use_abs && B.cell_col != nullptr;
    const int qs = row_scatter_index(lane);  // this lane's terms of the row's sums: qs and 16 + qs
    const uint32_t nrec = rg.y - rg.x;
    const int nbatch = (int)((nrec + kBatch - 1) / kBatch);
    for (int bi = nbatch - 1; bi >= 0; bi--) {
        const uint32_t base = rg.x + (uint32_t)bi * kBatch;
        const int cnt = (int)min((uint32_t)kBatch, rg.y - base);
        uint32_t sub = 0u;  // sub-tiles the record of this lane covers
        if (lane < cnt) {
            const uint4 rec = recs[2 * (int64_t)(base + lane) + 1];
            PtRecData R;
            pt_fetch(pt_vert, pt_col, (int64_t)B.block0 * 256, rec, R);
            pt_store((int64_t)B.block0 * 256, rec, R, L, lane);
            sub = ((rec.y & 0x0F0F0F0Fu) ? 1u : 0u) | ((rec.y & 0xF0F0F0F0u) ? 2u : 0u) |
                  ((rec.z & 0x0F0F0F0Fu) ? 4u : 0u) | ((rec.z & 0xF0F0F0F0u) ? 8u : 0u);
        }
        // the records of this lane's sub-tile
        uint32_t gm;
        {
            const uint32_t m0 = (uint32_t)ballot(sub & 1u), m1 = (uint32_t)ballot(sub & 2u);
            const uint32_t m2 = (uint32_t)ballot(sub & 4u), m3 = (uint32_t)ballot(sub & 8u);
            gm = grp == 0 ? m0 : grp == 1 ? m1 : grp == 2 ? m2 : m3;
        }
        for (int e = lane; e < kBatch * 21; e += 64)
            reinterpret_cast<float4*>(&s_acc[0][0])[e] = make_float4(0.f, 0.f, 0.f, 0.f);  // all four copies
        __syncthreads();
        while (ballot(gm != 0u) != 0ull) {
            const int j = gm != 0u ? 31 - __clz(gm) : -1;  // back to front
            if (j >= 0) gm &= ~(1u << j);
            const bool live = j >= 0 && valid && n > 0 && base + (uint32_t)j < last &&
                              ((L.mask[cbit >> 5][j] >> (cbit & 31)) & 1u);
            float g[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            float lam[3] = {0.f, 0.f, 0.f};
            if (live) {
                PtFrag f;
                pt_eval(L, j, px, py, f);
                if (jck >= 0 && ck_n == n) {
                    Cw = ck_t;
                    jck--;
                    ck_n = ck_n2;
                    if (jck >= 0) {
                        ck_t = ps.ckpt_t[(int64_t)jck * P + pix];
                        ck_n2 = jck >= 1 ? ps.ckpt_n[(int64_t)(jck - 1) * P + pix] : 0xFFFFFFFFu;
                    }
                }
                float Tin = Cw < kPtTinyT ? 0.0f : div_nr(Cw, f.A);
                Tin = n == 1 ? 1.0f : Tin;
                Tin = (jck >= 0 && ck_n == n - 1) ? ck_t : Tin;
                {
#pragma clang fp contract(fast)
                const float a = 1.0f - f.A;
                // dL/dF (alpha form in .w), then the composite's adjoint moves past this fragment
                const float dF[4] = {Tin * dC[0], Tin * dC[1], Tin * dC[2], Tin * dC[3]};
                dC[3] = dC[3] * f.A - (dC[0] * (a * f.c[0]) + dC[1] * (a * f.c[1]) + dC[2] * (a * f.c[2]));
                Cw = Tin;
                n--;
                // F = (alpha c.rgb, alpha), alpha = 1 - exp(-c.a d)
                const float dal = dF[3] + (dF[0] * f.c[0] + dF[1] * f.c[1] + dF[2] * f.c[2]);
                const float dc[4] = {a * dF[0], a * dF[1], a * dF[2], dal * f.d * f.A};
                const float dd = dal * f.c[3] * f.A;
                // dL/dlambda_k = dc . c_k + dd d_k; moving corner k by delta (NDC xy) changes lambda by
                // -lambda_k grad(lambda) . delta, so dL/d(corner k) = -lambda_k G, G = sum_k dL/dlambda_k grad lambda_k
                float gl[3];
#pragma unroll
                for (int k = 0; k < 3; k++)
                    gl[k] = ((dc[0] * L.c[4 * k][j] + dc[1] * L.c[4 * k + 1][j]) +
                             (dc[2] * L.c[4 * k + 2][j] + dc[3] * L.c[4 * k + 3][j])) + dd * L.d[k][j];
                const float X0 = L.X[0][j], X1 = L.X[1][j], X2 = L.X[2][j];
                const float Y0 = L.Y[0][j], Y1 = L.Y[1][j], Y2 = L.Y[2][j];
                const float iD = L.invD[j];
                const float Gx = (gl[0] * (Y1 - Y2) + gl[1] * (Y2 - Y0) + gl[2] * (Y0 - Y1)) * iD;
                const float Gy = (gl[0] * (X2 - X1) + gl[1] * (X0 - X2) + gl[2] * (X1 - X0)) * iD;
                g[0] = -Gx;
                g[1] = -Gy;
                g[2] = abs_col ? fabsf(dc[0]) : dc[0];
                g[3] = abs_col ? fabsf(dc[1]) : dc[1];
                g[4] = abs_col ? fabsf(dc[2]) : dc[2];
                g[5] = abs_col ? fabsf(dc[3]) : dc[3];
                g[6] = dd;
                lam[0] = f.l[0];
                lam[1] = f.l[1];
                lam[2] = f.l[2];
                }
            }
            float p[21];
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int q = 0; q < 7; q++) p[7 * k + q] = lam[k] * g[q];
            float s[2];
            row_reduce_scatter(p, s);
            if (j >= 0) {
                float* a = &s_acc[grp][21 * j];
                a[qs] += s[0];
                if (qs < 5) a[16 + qs] += s[1];
            }
        }
        __syncthreads();
        for (int e = lane; e < cnt * 21; e += 64) {
            const float v = (s_acc[0][e] + s_acc[1][e]) + (s_acc[2][e] + s_acc[3][e]);
            const int j = e / 21;
            if (v != 0.f) atomic_add_f32(&tri_grad[21 * (int64_t)L.id[j] + (e - 21 * j)], v);
        }
        __syncthreads();
    }
}

// Reverse pass of pt_corrected_depth: the depth's gradient gdep carried to x, y, z1 and z2 (zero where the two
// points coincide, as the square root's derivative is taken there).
__device__ __forceinline__ void pt_corrected_depth_adj(const float* iP, float x, float y, float z1, float z2,
                                                       float gdep, float& gx, float& gy, float& gz1, float& gz2) {
    float e1[4], e2[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        e1[r] = ((iP[0 * 4 + r] * x + iP[1 * 4 + r] * y) + iP[2 * 4 + r] * z1) + iP[3 * 4 + r];
        e2[r] = ((iP[0 * 4 + r] * x + iP[1 * 4 + r] * y) + iP[2 * 4 + r] * z2) + iP[3 * 4 + r];
    }
    const float i1 = 1.0f / e1[3], i2 = 1.0f / e2[3];
    float p1[3], p2[3], dd[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        p1[k] = e1[k] * i1;
        p2[k] = e2[k] * i2;
        dd[k] = p1[k] - p2[k];
    }
    const float len = sqrtf((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]);
    const float s = len > 0.f ? gdep / len : 0.f;
    // dL/dp1 = s dd, dL/dp2 = -s dd; p = e.xyz / e.w
    float ge1[4], ge2[4];
    float w1 = 0.f, w2 = 0.f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float u = s * dd[k];
        ge1[k] = u * i1;
        ge2[k] = -u * i2;
        w1 += u * p1[k];
        w2 -= u * p2[k];
    }
    ge1[3] = -w1 * i1;
    ge2[3] = -w2 * i2;
    gx = gy = gz1 = gz2 = 0.f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        gx += iP[0 * 4 + r] * (ge1[r] + ge2[r]);
        gy += iP[1 * 4 + r] * (ge1[r] + ge2[r]);
        gz1 += iP[2 * 4 + r] * ge1[r];
        gz2 += iP[2 * 4 + r] * ge2[r];
    }
}

// One tet: the gradients of its triangles' corners, summed per vertex, carried back through the construction
// (pt_vertices) to the vertex positions and colours by a hand-written reverse pass.  The tet's vertices are taken
// in role order: 0 = seg1[0], 1 = seg1[1], 2 = seg2[0], 3 = seg2[1] (a permutation of the four), 4 = the crossing;
// every index below is then a compile-time constant and the state stays in registers.  Out: the four vertices'
// ids and gradients {position xyz, colour rgba} (colour zero for per-cell colours, whose gradient goes to the
// tet here).  False when the tet made no triangles.
__device__ __forceinline__ bool pt_backprop_tet(const BatchItem& B, float* __restrict__ grad_col, int64_t tet,
                                                int64_t slot, uint32_t info, const float* __restrict__ tri_grad,
                                                int32_t vid[4], float val[4][7]) {
    const ViewUniforms& U = B.U;
    const PtDec D = pt_unpack(info);
    if (D.ntri == 0) return false;
    // corner gradients by role: g[r] = {NDC x, NDC y, colour rgba, thickness}
    float g[5][7];
#pragma unroll
    for (int r = 0; r < 5; r++)
#pragma unroll
        for (int q = 0; q < 7; q++) g[r][q] = 0.f;
    const float4* tg4 = reinterpret_cast<const float4*>(tri_grad + 84 * slot);  // 4 triangles x 3 corners x 7
    float tg[84];
#pragma unroll
    for (int i = 0; i < 21; i++) {
        const float4 t = tg4[i];
        tg[4 * i] = t.x;
        tg[4 * i + 1] = t.y;
        tg[4 * i + 2] = t.z;
        tg[4 * i + 3] = t.w;
    }
    // fans (pt_corner): class 2 {4, a0, b0, a1, b1, a0}; class 1 {T, P, b0, b1, P} with the thick vertex T and
    // the other end P of segment 1, accumulated here as roles 1 and 0 and exchanged below when flipped
    if (D.cls2) {
        constexpr int fan[6] = {4, 0, 2, 1, 3, 0};
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int j = 0; j < 3; j++)
#pragma unroll
                for (int q = 0; q < 7; q++) g[fan[j == 0 ? 0 : k + j]][q] += tg[21 * k + 7 * j + q];
    } else {
        constexpr int fan[5] = {1, 0, 2, 3, 0};
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int j = 0; j < 3; j++)
#pragma unroll
                for (int q = 0; q < 7; q++) g[fan[j == 0 ? 0 : k + j]][q] += tg[21 * k + 7 * j + q];
    }
    const bool per_cell = B.cell_col != nullptr;
    // vertices in role order
    const int4 idx = *reinterpret_cast<const int4*>(B.cells + 4 * tet);
    vid[0] = idx.x;  // seg1[0] is always vertex 0
    {
        const int32_t v4[4] = {idx.x, idx.y, idx.z, idx.w};
        const int s[3] = {D.s11, D.s20, D.s21};
#pragma unroll
        for (int r = 1; r < 4; r++) vid[r] = s[r - 1] == 1 ? v4[1] : s[r - 1] == 2 ? v4[2] : v4[3];
    }
    float n[4][3], iw[4], C[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const float3_ p = load_pos(B.pos, vid[r]);
        const float P[3] = {p.x, p.y, p.z};
        float c[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
            c[k] = ((U.viewProj[0 * 4 + k] * P[0] + U.viewProj[1 * 4 + k] * P[1]) + U.viewProj[2 * 4 + k] * P[2]) +
                   U.viewProj[3 * 4 + k] * 1.0f;
#pragma unroll
        for (int k = 0; k < 3; k++) n[r][k] = c[k] / c[3];
        iw[r] = 1.0f / c[3];
        const float4 cc = per_cell ? *reinterpret_cast<const float4*>(B.cell_col + 4 * tet)
                                   : *reinterpret_cast<const float4*>(B.col + 4 * (int64_t)vid[r]);
        C[r][0] = cc.x;
        C[r][1] = cc.y;
        C[r][2] = cc.z;
        C[r][3] = cc.w;
    }
    // alpha, beta: the crossing of segment 1 (roles 0 -> 1) and segment 2 (roles 2 -> 3), as pt_vertices
    float A[3], Bv[3], Cv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        A[k] = n[1][k] - n[0][k];
        Bv[k] = n[3][k] - n[2][k];
        Cv[k] = n[2][k] - n[0][k];
    }
    const float den = A[0] * Bv[1] - A[1] * Bv[0];
    const float alpha = (Bv[1] * Cv[0] - Bv[0] * Cv[1]) / den;
    const float beta = (A[1] * Cv[0] - A[0] * Cv[1]) / den;
    float gn[4][3], gC[4][4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        gn[r][0] = g[r][0];
        gn[r][1] = g[r][1];
        gn[r][2] = 0.f;
#pragma unroll
        for (int ch = 0; ch < 4; ch++) gC[r][ch] = g[r][2 + ch];
    }
    float galpha = 0.f, gbeta = 0.f;
    const float gdep = (D.cls2 ? g[4][6] : g[1][6]) * U.attenuation;
    if (D.cls2) {
        // crossing: ndc4 = n0 + alpha A, depth from (ndc4, n2.z + beta B.z), colour the mean of the two edge points
        float x4[3];
#pragma unroll
        for (int k = 0; k < 3; k++) x4[k] = n[0][k] + alpha * A[k];
        float dx, dy, dz1, dz2;
        pt_corrected_depth_adj(U.invProj, x4[0], x4[1], x4[2], n[2][2] + beta * Bv[2], gdep, dx, dy, dz1, dz2);
        const float g4[3] = {g[4][0] + dx, g[4][1] + dy, dz1};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            gn[0][k] += (1.0f - alpha) * g4[k];
            gn[1][k] += alpha * g4[k];
            galpha += A[k] * g4[k];
        }
        gn[2][2] += (1.0f - beta) * dz2;
        gn[3][2] += beta * dz2;
        gbeta += Bv[2] * dz2;
        if (!per_cell) {
#pragma unroll
            for (int ch = 0; ch < 4; ch++) {
                const float h = 0.5f * g[4][2 + ch];
                gC[0][ch] += (1.0f - alpha) * h;
                gC[1][ch] += alpha * h;
                gC[2][ch] += (1.0f - beta) * h;
                gC[3][ch] += beta * h;
                galpha += h * (C[1][ch] - C[0][ch]);
                gbeta += h * (C[3][ch] - C[2][ch]);
            }
        }
    } else {
        // thick vertex T (role 1, or role 0 when flipped) with the depth to the face point on the line from P
        // through T: facez = (edgez + (alpha' - 1) P.z) / alpha', alpha' = flip ? 1 - alpha : alpha
        const bool fl = D.flip;
        const float al = fl ? 1.0f - alpha : alpha;
        float nT[3], nP[3], gT[3] = {0.f, 0.f, 0.f}, gP[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            nT[k] = fl ? n[0][k] : n[1][k];
            nP[k] = fl ? n[1][k] : n[0][k];
        }
        const float edgez = n[2][2] + beta * Bv[2];
        const float facez = (edgez + (al - 1.0f) * nP[2]) / al;
        float dx, dy, dz1, dz2;
        pt_corrected_depth_adj(U.invProj, nT[0], nT[1], nT[2], facez, gdep, dx, dy, dz1, dz2);
        gT[0] = dx;
        gT[1] = dy;
        gT[2] = dz1;
        const float ia = 1.0f / al;
        const float gedge = dz2 * ia;
        gP[2] += dz2 * (al - 1.0f) * ia;
        float gal = dz2 * (nP[2] - facez) * ia;
        gn[2][2] += (1.0f - beta) * gedge;
        gn[3][2] += beta * gedge;
        gbeta += Bv[2] * gedge;
        float gCT[4] = {0.f, 0.f, 0.f, 0.f}, gCP[4] = {0.f, 0.f, 0.f, 0.f};
        if (!per_cell) {
            // colour of T: 0.5 (facec + C_T), facec = (edgec + (alpha' - 1) C_P) / alpha'
#pragma unroll
            for (int ch = 0; ch < 4; ch++) {
                float c0 = C[0][ch], c1 = C[1][ch];
                // keeps the two colours in registers: a select between two rows of C is otherwise formed as a
                // dynamically indexed load from a stack copy of C
                asm volatile("" : "+v"(c0), "+v"(c1));
                const float CT = fl ? c0 : c1, CP = fl ? c1 : c0;
                const float h = 0.5f * g[1][2 + ch];
                const float edgec = C[2][ch] + beta * (C[3][ch] - C[2][ch]);
                const float facec = (edgec + (al - 1.0f) * CP) / al;
                gCT[ch] = h;
                const float ge = h * ia;
                gCP[ch] = h * (al - 1.0f) * ia;
                gC[2][ch] += (1.0f - beta) * ge;
                gC[3][ch] += beta * ge;
                gbeta += ge * (C[3][ch] - C[2][ch]);
                gal += h * (CP - facec) * ia;
            }
        }
        // slots 0 and 1 hold P's and T's corner gradients: P is role 0 and T role 1, exchanged when flipped
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const float sP = gn[0][k] + gP[k], sT = gn[1][k] + gT[k];
            gn[0][k] = fl ? sT : sP;
            gn[1][k] = fl ? sP : sT;
        }
        if (!per_cell) {
#pragma unroll
            for (int ch = 0; ch < 4; ch++) {
                // T's corner colour is the constructed one: T's input colour receives gCT only
                const float sP = gC[0][ch] + gCP[ch], sT = gCT[ch];
                gC[0][ch] = fl ? sT : sP;
                gC[1][ch] = fl ? sP : sT;
            }
        }
        galpha = fl ? -gal : gal;
    }
    // alpha = (B.y Cv.x - B.x Cv.y) / den, beta = (A.y Cv.x - A.x Cv.y) / den, den = A.x B.y - A.y B.x
    {
        const float ga = galpha / den, gb = gbeta / den;
        const float gden = -(ga * alpha + gb * beta);
        const float gA0 = -gb * Cv[1] + gden * Bv[1], gA1 = gb * Cv[0] - gden * Bv[0];
        const float gB0 = -ga * Cv[1] - gden * A[1], gB1 = ga * Cv[0] + gden * A[0];
        const float gC0 = ga * Bv[1] + gb * A[1], gC1 = -ga * Bv[0] - gb * A[0];
        gn[1][0] += gA0;
        gn[1][1] += gA1;
        gn[0][0] -= gA0 + gC0;
        gn[0][1] -= gA1 + gC1;
        gn[3][0] += gB0;
        gn[3][1] += gB1;
        gn[2][0] += gC0 - gB0;
        gn[2][1] += gC1 - gB1;
    }
    // NDC = c.xyz / c.w, c = viewProj (P, 1)
#pragma unroll
    for (int r = 0; r < 4; r++) {
        float gc[4];
#pragma unroll
        for (int k = 0; k < 3; k++) gc[k] = gn[r][k] * iw[r];
        gc[3] = -((gn[r][0] * n[r][0] + gn[r][1] * n[r][1]) + gn[r][2] * n[r][2]) * iw[r];
#pragma unroll
        for (int j = 0; j < 3; j++)
            val[r][j] = (U.viewProj[j * 4 + 0] * gc[0] + U.viewProj[j * 4 + 1] * gc[1]) +
                        (U.viewProj[j * 4 + 2] * gc[2] + U.viewProj[j * 4 + 3] * gc[3]);
#pragma unroll
        for (int ch = 0; ch < 4; ch++) val[r][3 + ch] = per_cell ? 0.f : gC[r][ch];
    }
    if (per_cell) {  // every corner carries the cell's colour
#pragma unroll
        for (int ch = 0; ch < 4; ch++) {
            const float s = ((g[0][2 + ch] + g[1][2 + ch]) + (g[2][2 + ch] + g[3][2 + ch])) + g[4][2 + ch];
            if (s != 0.f) atomic_add_f32(&grad_col[4 * tet + ch], s);
        }
    }
    return true;
}

// One wave per 64 tets (a quarter of a 256-tet block).  The wave's 256 (tet, vertex) slots are summed per vertex
// before the global atomics, as the tile renderer's adjoint does: each slot finds its vertex's entry in an LDS
// hash table keyed by vertex id (compare-and-swap; the first slot of a vertex creates the entry), stores its
// values in its own LDS row and links the row into the entry's list with one exchange; the entry's creator sums
// the list, and eight lanes per distinct vertex add its seven sums to the vertex's 32-byte row of the
// interleaved accumulator, folded into the gradients after the launch.  In a mesh ordered by locality 64
// consecutive tets share most vertices (about five slots per distinct vertex on a grid).
__global__ __launch_bounds__(64) void k_pt_backprop(const BatchItem* __restrict__ items,
                                                     const AdjointTarget* __restrict__ targets, int nitems,
                                                     const uint32_t* __restrict__ pt_info,
                                                     const float* __restrict__ tri_grad) {
    constexpr int kTab = 320;  // > 256 slots, so every slot finds an entry
    __shared__ int32_t keys[kTab], head[kTab], list[256];
    __shared__ float rows[256][8];  // 7 values, next slot
    const unsigned blk = blockIdx.x >> 2;  // the 256-tet block
    const int item = item_of_block(items, nitems, blk);
    const BatchItem& B = items[item];
    const AdjointTarget& Tg = targets[item];
    if (Tg.dimage == nullptr) return;  // uniform over the block
    const int lane = threadIdx.x;
    const int64_t tet = (int64_t)(blk - B.block0) * 256 + (blockIdx.x & 3u) * 64 + lane;
    const int64_t slot = (int64_t)blk * 256 + (blockIdx.x & 3u) * 64 + lane;
    int32_t vid[4] = {0, 0, 0, 0};
    float val[4][7];
    const bool live = tet < B.T && pt_backprop_tet(B, Tg.grad_col, tet, slot, pt_info[slot], tri_grad, vid, val);
    if (ballot(live) == 0ull) return;
    for (int e = lane; e < kTab; e += 64) {
        keys[e] = -1;
        head[e] = -1;
    }
    __syncthreads();
    int ent[4];
    bool owner[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        ent[k] = -1;
        owner[k] = false;
        if (!live) continue;
        const int32_t v = vid[k];
        int e = (int)((((uint32_t)v * 2654435761u) >> 16) * (uint32_t)kTab >> 16);  // multiplicative hash
        for (int probe = 0; probe < kTab; probe++) {
            const int32_t old = atomicCAS(&keys[e], -1, v);
            if (old == -1 || old == v) {
                ent[k] = e;
                owner[k] = old == -1;
                break;
            }
            e = e + 1 == kTab ? 0 : e + 1;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (ent[k] < 0) continue;
        const int sl = 64 * k + lane;
#pragma unroll
        for (int c = 0; c < 7; c++) rows[sl][c] = val[k][c];
        rows[sl][7] = __int_as_float(atomicExch(&head[ent[k]], sl));
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (!owner[k]) continue;
        float sum[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int sl = head[ent[k]]; sl >= 0;) {
#pragma unroll
            for (int c = 0; c < 7; c++) sum[c] += rows[sl][c];
            sl = __float_as_int(rows[sl][7]);
        }
        const int own = 64 * k + lane;
#pragma unroll
        for (int c = 0; c < 7; c++) rows[own][c] = sum[c];
        rows[own][7] = __int_as_float(vid[k]);
    }
    uint32_t nown = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint64_t m = ballot(owner[k]);
        if (owner[k]) list[nown + (uint32_t)popc64(m & lanemask_lt())] = 64 * k + lane;
        nown += (uint32_t)popc64(m);
    }
    __syncthreads();
    float* __restrict__ acc = Tg.acc;
    const int c = lane & 7;
    for (uint32_t b0 = 0; b0 < nown; b0 += 8) {
        const uint32_t li = b0 + (uint32_t)(lane >> 3);
        if (li < nown && c < 7) {
            const int sl = list[li];
            const float v = rows[sl][c];
            if (v != 0.f) atomic_add_f32(&acc[8 * (int64_t)__float_as_int(rows[sl][7]) + c], v);
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
void launch_pt_count(const BatchItem* items, int nitems, int64_t nblocks, uint32_t* pt_info, float4* pt_vert,
                     float4* pt_col, uint32_t* counts, uint32_t* block_stats, hipStream_t s) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(k_pt_count, dim3((unsigned)nblocks), dim3(256), 0, s, items, nitems, pt_info, pt_vert, pt_col,
                       counts, reinterpret_cast<uint4*>(block_stats));
}

void launch_pt_emit(const BatchItem* items, int nitems, int64_t nblocks, const uint32_t* pt_info,
                    const float4* pt_vert, const uint32_t* offsets, const uint32_t* counts, uint64_t* keys,
                    uint32_t* vals, int64_t cap, uint32_t* overflow, hipStream_t s) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(k_pt_emit, dim3((unsigned)nblocks), dim3(256), 0, s, items, nitems, pt_info, pt_vert, offsets,
                       counts, keys, vals, (uint64_t)cap, overflow);
}

void launch_pt_gather(const BatchItem* items, int nitems, const uint32_t* pt_info, const float4* pt_vert,
                      const SortRoute* route, int64_t n, const uint64_t* n_dev, uint4* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pt_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, items, nitems, pt_info,
                       pt_vert, route, n, n_dev, out);
}

void launch_pt_forward(const BatchItem* items, int nitems, int64_t ntiles, const uint32_t* order,
                       const uint32_t* pt_info, const float4* pt_vert, const float4* pt_col, const uint4* recs,
                       const uint2* ranges, PixelState ps, hipStream_t s) {
    if (ntiles <= 0) return;
    (void)pt_info;  // the records carry their corners' vertex slots
    hipLaunchKernelGGL(k_pt_forward, dim3((unsigned)ntiles), dim3(64), 0, s, items, nitems, order, pt_vert, pt_col,
                       recs, ranges, ps);
}

void launch_pt_adjoint(const BatchItem* items, const AdjointTarget* targets, int nitems, int64_t ntiles,
                       int64_t nblocks, const uint32_t* order, const uint32_t* pt_info, const float4* pt_vert,
                       const float4* pt_col, const uint4* recs, const uint2* ranges, PixelState ps, float* tri_grad,
                       int64_t tri_grad_floats, int use_abs, hipStream_t s) {
    if (ntiles <= 0) return;
    fill_u32(reinterpret_cast<uint32_t*>(tri_grad), 0u, tri_grad_floats, s);
    hipLaunchKernelGGL(k_pt_adjoint, dim3((unsigned)ntiles), dim3(64), 0, s, items, targets, nitems, order, pt_vert,
                       pt_col, recs, ranges, ps, tri_grad, use_abs);
    if (nblocks > 0)
        hipLaunchKernelGGL(k_pt_backprop, dim3((unsigned)(4 * nblocks)), dim3(64), 0, s, items, targets, nitems,
                           pt_info, tri_grad);
}

}  // namespace dtv
