// Record staging and per-fragment maths shared by the tile renderer's forward, adjoint and export kernels.
#pragma once
#include "dtv_kernels.h"
#include "dtv_prims.h"
#include "raster.h"

namespace dtv {

constexpr int kTileEdge = 8;
constexpr int kBatch = 32;  // records staged per batch (one 32-bit coverage mask per pixel)

struct alignas(16) RecordLds {  // 16-byte aligned: vid[] is also read as int4
    float p[12][kBatch];      // tet vertex positions, [3*v + c]
    float c[16][kBatch];      // tet vertex colours, [4*v + ch]
    float N[12][kBatch];      // face plane normals cross(p1-p0, p2-p0), [3*f + c]
    float k[4][kBatch];       // dot(N, camPos) - dot(N, p0) per face
    uint32_t mask[2][kBatch]; // coverage of the tile's pixels 0-31 / 32-63 (pixel = ly * 8 + lx = lane)
    uint32_t meta[kBatch];    // bit 3: face normals point outward; bits 4-31: tet index (num_cells <= 2^28)
    int32_t vid[4][kBatch];
};

// One record of the tile renderer's stream (two uint4 per record): the tet's vertex indices, and
// {tet * 4 + face, coverage mask low, high, 0}.
struct RecordIn {
    uint4 vid;
    uint4 rec;
};

__device__ __forceinline__ RecordIn load_record(const uint4* __restrict__ recs, uint32_t i) {
    RecordIn r;
    r.vid = recs[2 * (int64_t)i];
    r.rec = recs[2 * (int64_t)i + 1];
    return r;
}

// Stages record r in slot i (world-space data only: the coverage was decided at binning).
// Vertex data of one record, fetched ahead of its staging so that the loads overlap other work.  (On CDNA one
// counter tracks loads and stores alike: a load issued after the gradient atomics would wait for them too.)
struct RecordData {
    float3_ p[4];
    float4 c[4];
};

// kCell: col holds one colour per tet (ColorStorage.PER_CELL), fetched into c[0]; else one per vertex (or
// none, col NULL).
template <bool kCell = false>
__device__ __forceinline__ void fetch_record_data(const float* __restrict__ pos, const float* __restrict__ col,
                                                  const RecordIn& r, RecordData& d) {
    const int32_t vid[4] = {(int32_t)r.vid.x, (int32_t)r.vid.y, (int32_t)r.vid.z, (int32_t)r.vid.w};
#pragma unroll
    for (int v = 0; v < 4; v++) {
        d.p[v] = load_pos(pos, vid[v]);
        if constexpr (kCell) {
            d.c[v] = v == 0 ? *reinterpret_cast<const float4*>(col + 4 * (int64_t)(r.rec.x >> 2))
                            : make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
            d.c[v] = col != nullptr ? *reinterpret_cast<const float4*>(col + 4 * (int64_t)vid[v])
                                    : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
}

// kStash (binning tiles larger than 8x8): stash [6][kBatch] int32 receives the record's face as the oriented
// fixed-point screen triangle of binning (project_face + face_screen arithmetic), for subtile_masks.
template <bool kStash = false>
__device__ __forceinline__ void store_record(const ViewUniforms& U, const RecordIn& r, const RecordData& d,
                                             RecordLds& L, int i, int32_t* stash = nullptr) {
    const uint4 rec = r.rec;
    const int32_t vid[4] = {(int32_t)r.vid.x, (int32_t)r.vid.y, (int32_t)r.vid.z, (int32_t)r.vid.w};
    const float3_* p = d.p;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        L.p[3 * v + 0][i] = p[v].x;
        L.p[3 * v + 1][i] = p[v].y;
        L.p[3 * v + 2][i] = p[v].z;
        L.c[4 * v + 0][i] = d.c[v].x;
        L.c[4 * v + 1][i] = d.c[v].y;
        L.c[4 * v + 2][i] = d.c[v].z;
        L.c[4 * v + 3][i] = d.c[v].w;
        L.vid[v][i] = vid[v];
    }
    L.mask[0][i] = rec.y;
    L.mask[1][i] = rec.z;
    const float3_ ro = f3(U.camPos[0], U.camPos[1], U.camPos[2]);
    uint32_t meta = (rec.x >> 2) << 4;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const float3_ a = p[kFaceTable[g][0]], b = p[kFaceTable[g][1]], c = p[kFaceTable[g][2]];
        const float3_ n = cross(sub(b, a), sub(c, a));
        const float d = -dot(n, a);
        L.N[3 * g + 0][i] = n.x;
        L.N[3 * g + 1][i] = n.y;
        L.N[3 * g + 2][i] = n.z;
        L.k[g][i] = dot(n, ro) + d;
        if (g == 0 && dot(n, sub(p[3], p[0])) < 0.0f) meta |= 8u;
    }
    L.meta[i] = meta;
    if constexpr (kStash) {
        const int f = (int)(rec.x & 3u);
        const float hw = 0.5f * (float)U.width, hh = 0.5f * (float)U.height;
        int32_t X[3], Y[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int v = face_vertex(f, k);
            const float px = sel4(v, p[0].x, p[1].x, p[2].x, p[3].x);
            const float py = sel4(v, p[0].y, p[1].y, p[2].y, p[3].y);
            const float pz = sel4(v, p[0].z, p[1].z, p[2].z, p[3].z);
            float c[4];
#pragma unroll
            for (int rr = 0; rr < 4; rr++)
                c[rr] = ((U.viewProj[0 * 4 + rr] * px + U.viewProj[1 * 4 + rr] * py) + U.viewProj[2 * 4 + rr] * pz) +
                        U.viewProj[3 * 4 + rr] * 1.0f;
            const float nx = c[0] / c[3], ny = c[1] / c[3];
            const float xf = nx * hw + hw;
            const float yf = ny * hh + hh;
            X[k] = (int32_t)rintf(xf * 256.0f);
            Y[k] = (int32_t)rintf(yf * 256.0f);
        }
        const int64_t area =
            (int64_t)(X[1] - X[0]) * (Y[2] - Y[0]) - (int64_t)(Y[1] - Y[0]) * (X[2] - X[0]);
        if (area < 0) {
            const int32_t tx = X[1], ty = Y[1];
            X[1] = X[2];
            Y[1] = Y[2];
            X[2] = tx;
            Y[2] = ty;
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            stash[k * kBatch + i] = X[k];
            stash[(3 + k) * kBatch + i] = Y[k];
        }
    }
}

// Coverage masks of the batch's records for this wave's 8x8 sub-tile (binning tiles larger than 8x8), from the
// screen triangles store_record stashed: L.mask as the binning's gather writes it for 8x8 tiles.
__device__ __forceinline__ void subtile_masks(RecordLds& L, const int32_t* stash, int cnt, int x, int y,
                                              bool valid, int lane) {
    for (int j = 0; j < cnt; j++) {
        const int32_t X[3] = {stash[j], stash[kBatch + j], stash[2 * kBatch + j]};
        const int32_t Y[3] = {stash[3 * kBatch + j], stash[4 * kBatch + j], stash[5 * kBatch + j]};
        const uint64_t m = ballot(valid && pixel_inside(X, Y, x, y));
        if (lane == 0) {
            L.mask[0][j] = (uint32_t)m;
            L.mask[1][j] = (uint32_t)(m >> 32);
        }
    }
}

// The 8x8 sub-tile a workgroup renders: global sub-tile gsub of item B, its binning tile (records and ranges
// are per binning tile) and its first pixel.
struct SubTile {
    uint32_t gtile;
    int x0, y0;
};
__device__ __forceinline__ SubTile sub_tile(const BatchItem& B, uint32_t gsub) {
    const ViewUniforms& U = B.U;
    const int sub = (int)(gsub - B.sub0);
    const int sx = sub % U.sub_x, sy = sub / U.sub_x;
    const int k = U.tile_shift - 3;
    SubTile t;
    t.gtile = B.tile0 + (uint32_t)((sy >> k) * U.tiles_x + (sx >> k));
    t.x0 = sx * kTileEdge;
    t.y0 = sy * kTileEdge;
    return t;
}

__device__ __forceinline__ void setup_record(const ViewUniforms& U, const float* __restrict__ pos,
                                             const float* __restrict__ col, const RecordIn& r, RecordLds& L, int i) {
    RecordData d;
    fetch_record_data(pos, col, r, d);
    store_record(U, r, d, L, i);
}

// 32-bit mask of the batch's records (bit j: record j, j < cnt) that cover this lane's pixel.
__device__ __forceinline__ uint32_t lane_coverage(const RecordLds& L, int cnt, int lane) {
    const uint32_t* w = L.mask[lane >> 5];
    const int b = lane & 31;
    uint32_t m = 0;
    for (int j = 0; j < cnt; j++) m |= ((w[j] >> b) & 1u) << j;
    return m;
}

__device__ __forceinline__ float3_ lds_p(const RecordLds& L, int v, int j) {
    return f3(L.p[3 * v][j], L.p[3 * v + 1][j], L.p[3 * v + 2][j]);
}
__device__ __forceinline__ float4 lds_c(const RecordLds& L, int v, int j) {
    return make_float4(L.c[4 * v][j], L.c[4 * v + 1][j], L.c[4 * v + 2][j], L.c[4 * v + 3][j]);
}

struct FragEval {
    bool hit;
    int f0, f1;
    float t0, t1;
    float u0, v0, u1, v1;
    float4 col0, col1;
    float4 seg[4];
    float4 pre[4];   // composite inside the tet before sub-segment s (RGB, transmittance)
    float4 F;        // composite of the 4 sub-segments (premultiplied RGB, transmittance of the fragment)
};

__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }

// Vector helpers for the continuous maths (colours, sub-segment integrals, compositing, gradients), where the
// products and sums may fuse into FMAs.  The discrete maths (projection, coverage, ray-tet decisions) keeps the
// separately rounded operations of dtv_device.h, which the binning and the oracle evaluate alike.
__device__ __forceinline__ float cdot(float3_ a, float3_ b) {
#pragma clang fp contract(fast)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ float3_ ccross(float3_ a, float3_ b) {
#pragma clang fp contract(fast)
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float3_ caxpy(float s, float3_ a, float3_ b) {  // s a + b
#pragma clang fp contract(fast)
    return {s * a.x + b.x, s * a.y + b.y, s * a.z + b.z};
}

// The hardware's reciprocal, square root and reciprocal square root (about 1 ulp; one instruction each, where
// x / y and sqrtf expand into correctly rounded sequences of ten or more).  Only for continuous quantities
// (colours, barycentrics, sub-segment opacity, gradients): no discrete decision may depend on them.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_div(float x, float y) { return x * __builtin_amdgcn_rcpf(y); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
// x / y from the hardware reciprocal and one Newton correction of the quotient (residual by FMA): within about
// half an ulp, for a chain of divisions whose errors compound.  Still not for discrete decisions.
__device__ __forceinline__ float div_nr(float x, float y) {
    const float r = __builtin_amdgcn_rcpf(y);
    const float q = x * r;
    return __builtin_fmaf(__builtin_fmaf(-q, y, x), r, q);
}

// The 4 sub-segments of a segment of length 4 tSeg between the entry and exit colours (RGB + density):
// seg[s] = accumulateLinearConst of sub-segment s (ForwardCommon.glsl:42-49, colour endpoints mixed at s/4 and
// (s+1)/4, density at the midpoint, IntersectRasterization.glsl:228-241), pre[s] = their composite before s,
// F = the composite of all four.  The colour at fraction f is formed as c0 + f (c1 - c0).  Composites carry
// premultiplied RGB and, in .w, the transmittance (1 - alpha) as a product of the sub-segments' exp(-a tSeg):
// a product keeps its relative precision at any depth, where 1 - alpha formed from an alpha near 1 does not.
// It also returns each sub-segment's density a[s], transmittance A[s] = exp(-a tSeg) and 1 / a (A = 1 and
// 1 / a = 0 where a < 1e-6: the sub-segment is empty).  The reference adjoint's evaluation takes seg/pre from
// it; the forward and the exact backward need only F and the scalars (sub_segments_F, sub_segments_w below,
// the same arithmetic per sub-segment), which the backward recomputes instead of keeping them from phase A.
__device__ __forceinline__ void sub_segments_ex(float4 col0, float4 col1, float tSeg, float att, float4 seg[4],
                                                float4 pre[4], float4& F, float a_s[4], float A_s[4],
                                                float ia_s[4]) {
#pragma clang fp contract(fast)
    const float dx = col1.x - col0.x, dy = col1.y - col0.y, dz = col1.z - col0.z, dw = col1.w - col0.w;
    float4 rc = make_float4(0.f, 0.f, 0.f, 1.0f);
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const float fb = (float)s * 0.25f, fm = ((float)s + 0.5f) * 0.25f;
        const float c0x = col0.x + fb * dx, c0y = col0.y + fb * dy, c0z = col0.z + fb * dz;
        const float a = (col0.w + fm * dw) * att;
        float4 cur = make_float4(0.f, 0.f, 0.f, 0.f);
        a_s[s] = a;
        A_s[s] = 1.0f;
        ia_s[s] = 0.0f;
        if (!(a < 1e-6f)) {
            const float A = __expf(-a * tSeg);
            const float ia = fast_rcp(a);
            // accumulateLinearConst: (1 - A) c0 + ((tSeg + 1/a) A - 1/a) (c0 - c1), c0 - c1 = -(c1 - c0) / 4
            const float kq = -0.25f * ((tSeg + ia) * A - ia);
            cur = make_float4((1.0f - A) * c0x + kq * dx, (1.0f - A) * c0y + kq * dy, (1.0f - A) * c0z + kq * dz,
                              1.0f - A);
            A_s[s] = A;
            ia_s[s] = ia;
        }
        pre[s] = rc;
        seg[s] = cur;
        rc.x = rc.x + rc.w * cur.x;
        rc.y = rc.y + rc.w * cur.y;
        rc.z = rc.z + rc.w * cur.z;
        rc.w = rc.w * A_s[s];
    }
    F = rc;
}

__device__ __forceinline__ void sub_segments(float4 col0, float4 col1, float tSeg, float att, float4 seg[4],
                                             float4 pre[4], float4& F) {
    float a_s[4], A_s[4], ia_s[4];
    sub_segments_ex(col0, col1, tSeg, att, seg, pre, F, a_s, A_s, ia_s);
}

// The scalar part of sub_segments_ex: per sub-segment s the density a[s], transmittance A[s], 1 / a, the weight
// kq[s] of the colour step in its colour (seg[s].rgb = (1 - A[s]) c0_s + kq[s] (col1 - col0), c0_s = col0 + s/4
// (col1 - col0)) and the transmittance T[s] before it; returns the fragment's transmittance.
__device__ __forceinline__ float sub_segments_w(float4 col0, float4 col1, float tSeg, float att, float a_s[4],
                                                float A_s[4], float ia_s[4], float kq_s[4], float T_s[4]) {
#pragma clang fp contract(fast)
    const float dw = col1.w - col0.w;
    float T = 1.0f;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const float fm = ((float)s + 0.5f) * 0.25f;
        const float a = (col0.w + fm * dw) * att;
        a_s[s] = a;
        A_s[s] = 1.0f;
        ia_s[s] = 0.0f;
        kq_s[s] = 0.0f;
        T_s[s] = T;
        if (!(a < 1e-6f)) {
            const float A = __expf(-a * tSeg);
            const float ia = fast_rcp(a);
            A_s[s] = A;
            ia_s[s] = ia;
            kq_s[s] = -0.25f * ((tSeg + ia) * A - ia);
        }
        T = T * A_s[s];
    }
    return T;
}

// sub_segments_ex's F alone: the RGB composite sum_s T[s] seg[s].rgb is col0 times sum_s T[s] (1 - A[s]) plus
// (col1 - col0) times sum_s T[s] ((1 - A[s]) s/4 + kq[s]): two scalar sums instead of a composite per channel.
__device__ __forceinline__ float4 sub_segments_F(float4 col0, float4 col1, float tSeg, float att) {
#pragma clang fp contract(fast)
    float a_s[4], A_s[4], ia_s[4], kq_s[4], T_s[4];
    const float Tend = sub_segments_w(col0, col1, tSeg, att, a_s, A_s, ia_s, kq_s, T_s);
    float S0 = 0.f, S1 = 0.f;
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const float om = 1.0f - A_s[s];
        S0 += T_s[s] * om;
        S1 += T_s[s] * (om * ((float)s * 0.25f) + kq_s[s]);
    }
    return make_float4(S0 * col0.x + S1 * (col1.x - col0.x), S0 * col0.y + S1 * (col1.y - col0.y),
                       S0 * col0.z + S1 * (col1.z - col0.z), Tend);
}

// Barycentric coordinates with unsigned areas and the 1e-5 area floor (BarycentricInterpolation.glsl:29-38).
__device__ __forceinline__ void bary_ref(float3_ p0, float3_ p1, float3_ p2, float3_ P, float& u, float& v) {
    const float3_ d20 = sub(p2, p0), d21 = sub(p2, p1);
    const float area = fmaxf(length(cross(d20, d21)), 1e-5f);
    u = length(cross(d21, sub(P, p1))) / area;
    v = length(cross(sub(P, p0), d20)) / area;
}

// Inside test of rayTriangleIntersect (RayIntersectionTests.glsl:45-68) for face g at ray parameter t (the t
// that the reference computes for that face): P = ro + t rd must lie on the inner side of all three edges.
__device__ __forceinline__ bool ref_face_contains(const RecordLds& L, int j, int g, float3_ ro, float3_ rd, float t) {
    const float3_ p0 = lds_p(L, face_vertex(g, 0), j), p1 = lds_p(L, face_vertex(g, 1), j),
                  p2 = lds_p(L, face_vertex(g, 2), j);
    const float3_ N = f3(L.N[3 * g][j], L.N[3 * g + 1][j], L.N[3 * g + 2][j]);
    const float3_ P = add(ro, scale(rd, t));
    if (dot(N, cross(sub(p1, p0), sub(P, p0))) < 0.0f) return false;
    if (dot(N, cross(sub(p2, p1), sub(P, p1))) < 0.0f) return false;
    if (dot(N, cross(sub(p0, p2), sub(P, p2))) < 0.0f) return false;
    return true;
}

// Signed barycentrics of P = ro + t rd in face g: the shares of the sub-triangles opposite p0 and p1,
// u = dot(N, (p2 - p1) x (P - p1)) / |N|^2 and v = dot(N, (p0 - p2) x (P - p2)) / |N|^2 with N = (p1 - p0) x
// (p2 - p0) as staged.  Inside the triangle they equal the unsigned area ratios of
// BarycentricInterpolation.glsl:29-38 (|(p2 - p1) x (P - p1)| / |N| and |(p0 - p2) x (P - p2)| / |N|), and
// 1 - u - v is the share opposite p2.  Returns false for faces whose |N| is below the reference's 1e-5 area
// floor (the caller then takes the reference procedure).
__device__ __forceinline__ bool face_bary(const RecordLds& L, int j, int g, float3_ ro, float3_ rd, float t, float& u,
                                          float& v) {
#pragma clang fp contract(fast)
    const float3_ p0 = lds_p(L, face_vertex(g, 0), j), p1 = lds_p(L, face_vertex(g, 1), j),
                  p2 = lds_p(L, face_vertex(g, 2), j);
    const float3_ N = f3(L.N[3 * g][j], L.N[3 * g + 1][j], L.N[3 * g + 2][j]);
    const float3_ P = caxpy(t, rd, ro);
    const float3_ c1 = ccross(sub(p2, p1), sub(P, p1));
    const float3_ c2 = ccross(sub(p0, p2), sub(P, p2));
    const float nn = cdot(N, N);
    const float inn = fast_rcp(nn);
    u = cdot(N, c1) * inn;
    v = cdot(N, c2) * inn;
    return nn >= 1e-10f;
}

constexpr float kEdgeMargin = 1e-4f;

// Entry/exit faces and distances with the semantics of intersectRayTet (RayIntersectionTests.glsl:78-106).
// Fast path: the plane-slab test (last entering / first exiting plane) proposes the two faces, and it stands when
// both hits lie at least kEdgeMargin (in barycentric coordinates) inside their faces: there the reference's
// inside test passes for the two faces and for no other.  Otherwise the reference's four-face procedure decides
// (edges, vertices and silhouettes, where float rounding can make the reference reject a face, and the camera
// inside the tet).  On the fast path bary = {u, v} at the entry and exit points (face_bary) and have_bary = true.
__device__ __forceinline__ bool ray_tet(const RecordLds& L, int j, float3_ ro, float3_ rd, int& f0, int& f1, float& t0,
                                        float& t1, float bary[4], bool& have_bary) {
    const bool outward = (L.meta[j] >> 3) & 1u;
    float tin = -__builtin_inff(), tout = __builtin_inff();
    int fi = -1, fo = -1;
    float tf[4];
    bool par[4];
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const float3_ n = f3(L.N[3 * g][j], L.N[3 * g + 1][j], L.N[3 * g + 2][j]);
        const float den = dot(n, rd);
        par[g] = fabsf(den) < 1e-9f;
        tf[g] = -L.k[g][j] / den;
        if (par[g]) continue;
        const bool entering = outward ? (den < 0.0f) : (den > 0.0f);
        if (entering) {
            if (tf[g] > tin) { tin = tf[g]; fi = g; }
        } else {
            if (tf[g] < tout) { tout = tf[g]; fo = g; }
        }
    }
    if (fi >= 0 && fo >= 0 && tout >= tin && tin >= 0.0f && face_bary(L, j, fi, ro, rd, tin, bary[0], bary[1]) &&
        face_bary(L, j, fo, ro, rd, tout, bary[2], bary[3]) &&
        fminf(fminf(fminf(bary[0], bary[1]), 1.0f - bary[0] - bary[1]),
              fminf(fminf(bary[2], bary[3]), 1.0f - bary[2] - bary[3])) >= kEdgeMargin) {
        f0 = fi;
        f1 = fo;
        t0 = tin;
        t1 = tout;
        have_bary = true;
        return true;
    }
    have_bary = false;
    t0 = 1e9f;
    t1 = -1e9f;
    f0 = 0;
    f1 = 0;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        if (par[g]) continue;
        if (ref_face_contains(L, j, g, ro, rd, tf[g])) {
            if (tf[g] >= 0.0f) {
                if (tf[g] < t0) { f0 = g; t0 = tf[g]; }
                if (tf[g] > t1) { f1 = g; t1 = tf[g]; }
            } else {
                f0 = g;
                t0 = 0.0f;
            }
        }
    }
    return t1 >= t0;
}

// Segment length per unit length along the ray: 1 for the INTERSECTION renderer's Euclidean segments
// (IntersectRasterization.glsl:211); for the PPLL renderer's differences of linear eye depths (PPLL/Forward.glsl
// :53,188), the ray's component along the view axis.
__device__ __forceinline__ float seg_scale(const ViewUniforms& U, float3_ rd) {
    return U.linear_depth ? ((U.depth_axis[0] * rd.x + U.depth_axis[1] * rd.y) + U.depth_axis[2] * rd.z) : 1.0f;
}

// kCell: the tet's own colour (staged as vertex 0's) over the whole segment, accumulateConst
// (IntersectRasterization.glsl:250-255, Common/ForwardCommon.glsl:37-40): F = ((1 - A) c, A), A = exp(-a t);
// the barycentrics are then needed only by the adjoint (kBary).
// kSeg: also E.seg / E.pre (the reference adjoint's sub-segment values); else E.F alone.
template <bool kCell = false, bool kBary = true, bool kSeg = false>
__device__ __forceinline__ void eval_fragment(const RecordLds& L, int j, float3_ ro, float3_ rd, float att,
                                              float tscale, FragEval& E) {
    int fi, fo;
    float tin, tout, bary[4];
    bool have_bary;
    E.hit = ray_tet(L, j, ro, rd, fi, fo, tin, tout, bary, have_bary);
    if (!E.hit) {
        E.F = make_float4(0.f, 0.f, 0.f, 1.0f);
        return;
    }
    E.f0 = fi;
    E.f1 = fo;
    E.t0 = tin;
    E.t1 = tout;
    if constexpr (kCell) {
        if constexpr (kBary) {
            if (!have_bary) {
                const float3_ P0 = add(ro, scale(rd, tin)), P1 = add(ro, scale(rd, tout));
                bary_ref(lds_p(L, face_vertex(fi, 0), j), lds_p(L, face_vertex(fi, 1), j),
                         lds_p(L, face_vertex(fi, 2), j), P0, bary[0], bary[1]);
                bary_ref(lds_p(L, face_vertex(fo, 0), j), lds_p(L, face_vertex(fo, 1), j),
                         lds_p(L, face_vertex(fo, 2), j), P1, bary[2], bary[3]);
            }
            E.u0 = bary[0];
            E.v0 = bary[1];
            E.u1 = bary[2];
            E.v1 = bary[3];
        }
        const float4 c = lds_c(L, 0, j);
        E.col0 = c;
        E.col1 = c;
        const float A = __expf(-(c.w * att) * ((tout - tin) * tscale));
        E.F = make_float4((1.0f - A) * c.x, (1.0f - A) * c.y, (1.0f - A) * c.z, A);
        return;
    }
    if (!have_bary) {  // edges, vertices, silhouettes, camera inside the tet: the faces' own barycentrics
        const float3_ P0 = add(ro, scale(rd, tin)), P1 = add(ro, scale(rd, tout));
        bary_ref(lds_p(L, face_vertex(fi, 0), j), lds_p(L, face_vertex(fi, 1), j), lds_p(L, face_vertex(fi, 2), j),
                 P0, bary[0], bary[1]);
        bary_ref(lds_p(L, face_vertex(fo, 0), j), lds_p(L, face_vertex(fo, 1), j), lds_p(L, face_vertex(fo, 2), j),
                 P1, bary[2], bary[3]);
    }
    {
#pragma clang fp contract(fast)
        const int a = face_vertex(fi, 0), b = face_vertex(fi, 1), c = face_vertex(fi, 2);
        const float u = bary[0], v = bary[1];
        const float w = 1.0f - u - v;
        const float4 ca = lds_c(L, a, j), cb = lds_c(L, b, j), cc = lds_c(L, c, j);
        E.u0 = u;
        E.v0 = v;
        E.col0 = make_float4((ca.x * u + cb.x * v) + cc.x * w, (ca.y * u + cb.y * v) + cc.y * w,
                             (ca.z * u + cb.z * v) + cc.z * w, (ca.w * u + cb.w * v) + cc.w * w);
    }
    {
#pragma clang fp contract(fast)
        const int a = face_vertex(fo, 0), b = face_vertex(fo, 1), c = face_vertex(fo, 2);
        const float u = bary[2], v = bary[3];
        const float w = 1.0f - u - v;
        const float4 ca = lds_c(L, a, j), cb = lds_c(L, b, j), cc = lds_c(L, c, j);
        E.u1 = u;
        E.v1 = v;
        E.col1 = make_float4((ca.x * u + cb.x * v) + cc.x * w, (ca.y * u + cb.y * v) + cc.y * w,
                             (ca.z * u + cb.z * v) + cc.z * w, (ca.w * u + cb.w * v) + cc.w * w);
    }
    if constexpr (kSeg)
        sub_segments(E.col0, E.col1, ((tout - tin) / 4.0f) * tscale, att, E.seg, E.pre, E.F);
    else
        E.F = sub_segments_F(E.col0, E.col1, ((tout - tin) / 4.0f) * tscale, att);
}

__device__ __forceinline__ int pixel_index(const ViewUniforms& U, int x, int y) { return y * U.width + x; }

// Batch tile processed by this workgroup: order[blockIdx.x] (a permutation of the batch's tiles), or
// blockIdx.x when order is NULL.
__device__ __forceinline__ uint32_t global_tile_of_block(const uint32_t* __restrict__ order) {
    return order != nullptr ? order[blockIdx.x] : blockIdx.x;
}

// Per-pixel state of one batch item: its slice of the batch's concatenated arrays.
__device__ __forceinline__ PixelState item_pixel_state(const PixelState& all, const BatchItem& B) {
    PixelState ps;
    ps.ray_color = all.ray_color + B.pixel0;
    ps.frag_count = all.frag_count + B.pixel0;
    ps.last_rec = all.last_rec + B.pixel0;
    ps.num_ckpt = all.num_ckpt + B.pixel0;
    ps.ckpt_t = all.ckpt_t + (int64_t)B.U.max_ckpt * B.pixel0;
    ps.ckpt_n = all.ckpt_n + (int64_t)B.U.max_ckpt * B.pixel0;
    ps.counters = all.counters;
    ps.tile_frags = all.tile_frags;  // indexed by batch tile
    return ps;
}

// ---------------------------------------------------------------------------------------------------------
// Exact adjoint of one fragment given dL/dF.  Writes per-local-vertex gradients gp[v][xyz], gc[v][rgba].
// Position gradient of face vertex j:  mu_j(P) * Q, Q = s*dLdt*N/(N.d) - [G^T dC - N (dC.G d)/(N.d)] with G the
// in-plane colour gradient of the face and s = -1 (entry face) / +1 (exit face).
__device__ __forceinline__ void face_side_q(const RecordLds& L, int j, int f, float3_ rd, const float dC[4],
                                            float sdLdt, float3_& Q) {
#pragma clang fp contract(fast)
    const int ia = face_vertex(f, 0), ib = face_vertex(f, 1), ic = face_vertex(f, 2);
    const float3_ a = lds_p(L, ia, j), b = lds_p(L, ib, j), c = lds_p(L, ic, j);
    const float4 ca = lds_c(L, ia, j), cb = lds_c(L, ib, j), cc = lds_c(L, ic, j);
    const float3_ e1 = sub(b, a), e2 = sub(c, a);
    const float3_ n = f3(L.N[3 * f][j], L.N[3 * f + 1][j], L.N[3 * f + 2][j]);  // = cross(e1, e2), staged
    const float inn = fast_rcp(cdot(n, n));
    // G^T dC and dC . (G d), with the in-plane gradients of the barycentrics of b and c
    // gb = (e2 x n) / |n|^2 and gc = (n x e1) / |n|^2:  gb sb + gc sc = ((sb e2 - sc e1) x n) / |n|^2
    const float db[4] = {cb.x - ca.x, cb.y - ca.y, cb.z - ca.z, cb.w - ca.w};
    const float dcv[4] = {cc.x - ca.x, cc.y - ca.y, cc.z - ca.z, cc.w - ca.w};
    float sb = 0.f, sc = 0.f;
#pragma unroll
    for (int ch = 0; ch < 4; ch++) {
        sb += db[ch] * dC[ch];
        sc += dcv[ch] * dC[ch];
    }
    const float3_ gtdc = scale(ccross(caxpy(-sc, e1, scale(e2, sb)), n), inn);
    const float dcgd = cdot(gtdc, rd);
    const float s = fast_div(sdLdt + dcgd, cdot(n, rd));
    Q = caxpy(s, n, scale(gtdc, -1.0f));
}

__device__ __forceinline__ void frag_backward_exact(const FragEval& E, const RecordLds& L, int j, float3_ rd,
                                                    float att, float tscale, float4 dF, float lam0[4],
                                                    float lam1[4], float dcol0[4], float dcol1[4], float3_& Q0,
                                                    float3_& Q1) {
#pragma clang fp contract(fast)
    float dR[4] = {dF.x, dF.y, dF.z, dF.w};
#pragma unroll
    for (int ch = 0; ch < 4; ch++) {
        dcol0[ch] = 0.f;
        dcol1[ch] = 0.f;
    }
    float dLdt = 0.f;
    const float tSeg = ((E.t1 - E.t0) / 4.0f) * tscale;
    float a_s[4], A_s[4], ia_s[4], kq_s[4], T_s[4];
    // as in eval_fragment: not kept live since phase A
    (void)sub_segments_w(E.col0, E.col1, tSeg, att, a_s, A_s, ia_s, kq_s, T_s);
    // dR.rgb is the same for every sub-segment: sum over RGB of dR c0 (sub-segment s starts at col0 + s/4
    // (col1 - col0)) = p0 + (s/4) p1, and of dR (c0 - c1) = -p1 / 4
    const float p0 = (dR[0] * E.col0.x + dR[1] * E.col0.y) + dR[2] * E.col0.z;
    const float p1 =
        (dR[0] * (E.col1.x - E.col0.x) + dR[1] * (E.col1.y - E.col0.y)) + dR[2] * (E.col1.z - E.col0.z);
    const float sdf = -0.25f * p1;
    float w0s = 0.f, w1s = 0.f;  // the RGB colour weights summed over the sub-segments
#pragma unroll
    for (int s = 3; s >= 0; s--) {
        const float omA = T_s[s];  // transmittance before s: dL/d(sub-segment s) = omA dR
        const float fb = (float)s * 0.25f, fm = ((float)s + 0.5f) * 0.25f, fe = ((float)s + 1.0f) * 0.25f;
        const float sc0 = p0 + fb * p1;
        const float dsw = omA * dR[3];
        // dR.rgb . seg[s].rgb = (1 - A) sc0 + kq p1
        dR[3] = dR[3] * A_s[s] - ((1.0f - A_s[s]) * sc0 + kq_s[s] * p1);
        const float a = a_s[s];
        float da = 0.f, dts = 0.f, w0 = 0.f, w1 = 0.f;
        if (!(a < 1e-6f)) {
            // accumulateLinearConst: out = (1 - A) c0 + K (c0 - c1), alpha 1 - A, A = exp(-a t),
            // K = (t + 1/a) A - 1/a; dA/dt = -a A, dA/da = -t A, dK/dt = -a t A, dK/da below
            const float A = A_s[s];
            const float ia = ia_s[s];
            const float K = (tSeg + ia) * A - ia;
            const float aA = a * A, tA = tSeg * A;
            const float dKda = ia * ia * (1.0f - A) - tA * (tSeg + ia);
            dts = aA * (omA * (sc0 - tSeg * sdf) + dsw);
            da = omA * (tA * sc0 + dKda * sdf) + dsw * tA;
            // d/dc0 = (1 - A + K), d/dc1 = -K; c0, c1 mix the entry/exit colours at fb, fe
            const float kc0 = 1.0f - A + K;
            w0 = omA * ((1.0f - fb) * kc0 - (1.0f - fe) * K);
            w1 = omA * (fb * kc0 - fe * K);
        }
        w0s += w0;
        w1s += w1;
        dcol0[3] += (1.0f - fm) * att * da;
        dcol1[3] += fm * att * da;
        dLdt += 0.25f * dts;
    }
#pragma unroll
    for (int ch = 0; ch < 3; ch++) {
        dcol0[ch] = w0s * dR[ch];
        dcol1[ch] = w1s * dR[ch];
    }
    dLdt *= tscale;  // d tSeg / d t1 = tscale / 4
    face_side_q(L, j, E.f0, rd, dcol0, -dLdt, Q0);
    face_side_q(L, j, E.f1, rd, dcol1, dLdt, Q1);
    // barycentric weights over the 4 local vertices (0 at the face's opposite vertex)
    const float mu0[3] = {E.u0, E.v0, 1.0f - E.u0 - E.v0};
    const float mu1[3] = {E.u1, E.v1, 1.0f - E.u1 - E.v1};
#pragma unroll
    for (int v = 0; v < 4; v++) {
        float l0 = 0.f, l1 = 0.f;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            l0 = face_vertex(E.f0, k) == v ? mu0[k] : l0;
            l1 = face_vertex(E.f1, k) == v ? mu1[k] : l1;
        }
        lam0[v] = l0;
        lam1[v] = l1;
    }
}

// Exact adjoint of one per-cell fragment (eval_fragment<true>) given dL/dF, whose .w is the derivative with
// respect to the fragment's alpha 1 - A (as throughout the adjoint): dcell = dL/d(tet colour), and the position
// weights as frag_backward_exact's (the hit points move; the face carries no colour gradient).  With
// F = (alpha c, alpha), alpha = 1 - exp(-a t): dL/dc = alpha dF.rgb, dL/dalpha = dF.w + dF.rgb . c,
// dalpha/dt = a A tscale, dalpha/d(density) = att t A.
__device__ __forceinline__ void frag_backward_cell(const FragEval& E, const RecordLds& L, int j, float3_ rd,
                                                   float att, float tscale, float4 dF, float lam0[4], float lam1[4],
                                                   float dcell[4], float3_& Q0, float3_& Q1) {
#pragma clang fp contract(fast)
    const float4 c = E.col0;
    const float tt = (E.t1 - E.t0) * tscale;
    const float a = c.w * att;
    const float A = __expf(-a * tt);
    const float dal = dF.w + (dF.x * c.x + dF.y * c.y + dF.z * c.z);
    dcell[0] = (1.0f - A) * dF.x;
    dcell[1] = (1.0f - A) * dF.y;
    dcell[2] = (1.0f - A) * dF.z;
    dcell[3] = (att * tt * A) * dal;
    const float dLdt = (a * A) * dal * tscale;
    {
        const int f = E.f0;
        const float3_ n = f3(L.N[3 * f][j], L.N[3 * f + 1][j], L.N[3 * f + 2][j]);
        Q0 = scale(n, fast_div(-dLdt, cdot(n, rd)));
    }
    {
        const int f = E.f1;
        const float3_ n = f3(L.N[3 * f][j], L.N[3 * f + 1][j], L.N[3 * f + 2][j]);
        Q1 = scale(n, fast_div(dLdt, cdot(n, rd)));
    }
    const float mu0[3] = {E.u0, E.v0, 1.0f - E.u0 - E.v0};
    const float mu1[3] = {E.u1, E.v1, 1.0f - E.u1 - E.v1};
#pragma unroll
    for (int v = 0; v < 4; v++) {
        float l0 = 0.f, l1 = 0.f;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            l0 = face_vertex(E.f0, k) == v ? mu0[k] : l0;
            l1 = face_vertex(E.f1, k) == v ? mu1[k] : l1;
        }
        lam0[v] = l0;
        lam1[v] = l1;
    }
}

// ---------------------------------------------------------------------------------------------------------
// The reference's own adjoint of one fragment (AdjointIntersectRasterization.glsl:111-362 with
// Common/BackwardCommon.glsl:116-251), quirks kept: the caller supplies dL/d(sub-segment colour) for
// sub-segments 3, 2, 1 only (sub-segment 0 is skipped), and the hit points move with fixed barycentrics.

// accumulateLinearConstAdjoint (BackwardCommon.glsl:116-133) for one sub-segment of length t.
__device__ __forceinline__ void ref_alc_adjoint(float t, float tf, float3_ c0, float3_ c1, float a, float A,
                                                float4 dC, float& dOut_dt, float3_& dc0, float3_& dc1, float& da) {
    const float dd = (dC.x * ((a * c0.x + (a * c1.x - a * c0.x) * t) * A) +
                      dC.y * ((a * c0.y + (a * c1.y - a * c0.y) * t) * A)) +
                     dC.z * ((a * c0.z + (a * c1.z - a * c0.z) * t) * A);
    dOut_dt += tf * (dd + dC.w * a * A);
    if (a < 1e-6f) {
        dc0 = f3(0.f, 0.f, 0.f);
        dc1 = f3(0.f, 0.f, 0.f);
        const float s = (dC.x * (t * c0.x + 0.5f * t * t * (c1.x - c0.x)) +
                         dC.y * (t * c0.y + 0.5f * t * t * (c1.y - c0.y))) +
                        dC.z * (t * c0.z + 0.5f * t * t * (c1.z - c0.z));
        da = s + dC.w * t;
    } else {
        const float inva = 1.0f / a;
        const float k0 = (1.0f - A) + ((t + inva) * A - inva), k1 = -((t + inva) * A - inva);
        dc0 = f3(k0 * dC.x, k0 * dC.y, k0 * dC.z);
        dc1 = f3(k1 * dC.x, k1 * dC.y, k1 * dC.z);
        const float kq = inva * inva - (t * t + inva * inva + t * inva) * A;
        const float s = (dC.x * (t * A * c0.x + kq * (c0.x - c1.x)) + dC.y * (t * A * c0.y + kq * (c0.y - c1.y))) +
                        dC.z * (t * A * c0.z + kq * (c0.z - c1.z));
        da = s + dC.w * t * A;
    }
}

// Derivatives of the unsigned-area barycentrics u = |(P-p1) x (p1-p2)| / |(p0-p2) x (p1-p2)| and
// v = |(P-p0) x (p0-p2)| / |...| with respect to the triangle's vertices at a fixed hit point P
// (baryAdjoint, BackwardCommon.glsl:183-251).
__device__ __forceinline__ void ref_bary_grads(float3_ P, float3_ p0, float3_ p1, float3_ p2, float3_ du[3],
                                               float3_ dv[3]) {
    const float3_ A0 = sub(p0, p2), B0 = sub(p1, p2), n0 = cross(A0, B0);
    const float3_ A1 = sub(P, p1), B1 = sub(p1, p2), n1 = cross(A1, B1);
    const float3_ A2 = sub(P, p0), B2 = sub(p0, p2), n2 = cross(A2, B2);
    const float f0 = dot(n0, n0), f1 = dot(n1, n1), f2 = dot(n2, n2);
    const float den0 = fmaxf(f0 * sqrtf(f0), 1e-6f);
    const float den1 = fmaxf(sqrtf(f1), 1e-6f);
    const float den2 = fmaxf(sqrtf(f2), 1e-6f);
    // halves of the gradients of |A x B|^2: w.r.t. A is B x (A x B), w.r.t. B is (A x B) x A
    const float3_ gA0 = cross(B0, n0), gB0 = cross(n0, A0);
    const float3_ gA1 = cross(B1, n1), gB1 = cross(n1, A1);
    const float3_ gA2 = cross(B2, n2), gB2 = cross(n2, A2);
    const float3_ hf0[3] = {gA0, gB0, scale(add(gA0, gB0), -1.0f)};
    const float3_ z = f3(0.f, 0.f, 0.f);
    const float3_ hf1[3] = {z, sub(gB1, gA1), scale(gB1, -1.0f)};
    const float3_ hf2[3] = {sub(gB2, gA2), z, scale(gB2, -1.0f)};
    du[0] = scale(hf0[0], -den1 / den0);
    dv[1] = scale(hf0[1], -den2 / den0);
    const float r1 = 1.0f / (den1 * den0), r2 = 1.0f / (den2 * den0);
#pragma unroll
    for (int k = 1; k < 3; k++) du[k] = scale(sub(scale(hf1[k], f0), scale(hf0[k], f1)), r1);
#pragma unroll
    for (int k = 0; k < 3; k += 2) dv[k] = scale(sub(scale(hf2[k], f0), scale(hf0[k], f2)), r2);
}

// Parameter gradients of one fragment from dacc[s - 1] = dL/d(colour of sub-segment s), s = 1..3.
// g[7 v + c]: tet-local vertex v, c = 0..2 position, 3..6 colour.
__device__ __forceinline__ void frag_backward_reference(const FragEval& E, const RecordLds& L, int j, float3_ ro,
                                                        float3_ rd, float att, const float4 dacc[3], bool use_abs,
                                                        float g[28]) {
    const float tSeg = (E.t1 - E.t0) / 4.0f;
    float dOut_dt = 0.f;
    float dcf0[4] = {0.f, 0.f, 0.f, 0.f}, dcf1[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 3; s > 0; s--) {
        const float fb = (float)s * 0.25f, fm = ((float)s + 0.5f) * 0.25f, fe = ((float)s + 1.0f) * 0.25f;
        const float3_ c0 = f3(mixf(E.col0.x, E.col1.x, fb), mixf(E.col0.y, E.col1.y, fb), mixf(E.col0.z, E.col1.z, fb));
        const float3_ c1 = f3(mixf(E.col0.x, E.col1.x, fe), mixf(E.col0.y, E.col1.y, fe), mixf(E.col0.z, E.col1.z, fe));
        const float al = mixf(E.col0.w, E.col1.w, fm);
        const float A = expf(-(al * att) * tSeg);
        float3_ dc0, dc1;
        float da;
        ref_alc_adjoint(tSeg, 0.25f, c0, c1, al * att, A, dacc[s - 1], dOut_dt, dc0, dc1, da);
        da *= att;
        dcf0[0] += (1.0f - fb) * dc0.x + (1.0f - fe) * dc1.x;
        dcf0[1] += (1.0f - fb) * dc0.y + (1.0f - fe) * dc1.y;
        dcf0[2] += (1.0f - fb) * dc0.z + (1.0f - fe) * dc1.z;
        dcf1[0] += fb * dc0.x + fe * dc1.x;
        dcf1[1] += fb * dc0.y + fe * dc1.y;
        dcf1[2] += fb * dc0.z + fe * dc1.z;
        dcf0[3] += (1.0f - fm) * da;
        dcf1[3] += fm * da;
    }
    // segmentLengthAdjoint: t = |P0 - P1| with P0, P1 on fixed barycentrics of the entry / exit faces
    const int fv[2][3] = {{face_vertex(E.f0, 0), face_vertex(E.f0, 1), face_vertex(E.f0, 2)},
                          {face_vertex(E.f1, 0), face_vertex(E.f1, 1), face_vertex(E.f1, 2)}};
    float3_ q[2][3];
#pragma unroll
    for (int sd = 0; sd < 2; sd++)
#pragma unroll
        for (int k = 0; k < 3; k++) q[sd][k] = lds_p(L, fv[sd][k], j);
    const float uu[2] = {E.u0, E.u1}, vv[2] = {E.v0, E.v1};
    const float ww[2] = {1.0f - E.u0 - E.v0, 1.0f - E.u1 - E.v1};
    const float3_ D = sub(add(add(scale(q[0][0], uu[0]), scale(q[0][1], vv[0])), scale(q[0][2], ww[0])),
                          add(add(scale(q[1][0], uu[1]), scale(q[1][1], vv[1])), scale(q[1][2], ww[1])));
    const float invDen = 1.0f / sqrtf(fmaxf(dot(D, D), 1e-6f));
#pragma unroll
    for (int q28 = 0; q28 < 28; q28++) g[q28] = 0.f;
#pragma unroll
    for (int sd = 0; sd < 2; sd++) {
        const float sgn = sd == 0 ? 1.0f : -1.0f;
        const float du_s = dOut_dt * (sgn * dot(sub(q[sd][0], q[sd][2]), D) * invDen);
        const float dv_s = dOut_dt * (sgn * dot(sub(q[sd][1], q[sd][2]), D) * invDen);
        const float wts[3] = {uu[sd], vv[sd], ww[sd]};
        const float* dcf = sd == 0 ? dcf0 : dcf1;
        const float3_ Pp = add(ro, scale(rd, sd == 0 ? E.t0 : E.t1));
        const float4 c0 = lds_c(L, fv[sd][0], j), c1 = lds_c(L, fv[sd][1], j), c2 = lds_c(L, fv[sd][2], j);
        const float du = du_s + (((dcf[0] * (c0.x - c2.x) + dcf[1] * (c0.y - c2.y)) + dcf[2] * (c0.z - c2.z)) +
                                 dcf[3] * (c0.w - c2.w));
        const float dv = dv_s + (((dcf[0] * (c1.x - c2.x) + dcf[1] * (c1.y - c2.y)) + dcf[2] * (c1.z - c2.z)) +
                                 dcf[3] * (c1.w - c2.w));
        float3_ gu[3], gv[3];
        ref_bary_grads(Pp, q[sd][0], q[sd][1], q[sd][2], gu, gv);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float3_ gp = scale(D, dOut_dt * sgn * wts[k] * invDen);
            gp = add(gp, add(scale(gu[k], du), scale(gv[k], dv)));
            float gc[4] = {dcf[0] * wts[k], dcf[1] * wts[k], dcf[2] * wts[k], dcf[3] * wts[k]};
            if (use_abs) {
                gp = f3(fabsf(gp.x), fabsf(gp.y), fabsf(gp.z));
#pragma unroll
                for (int ch = 0; ch < 4; ch++) gc[ch] = fabsf(gc[ch]);
            }
            const int v = fv[sd][k];
            // v is a runtime face-table lookup: select its 7 slots without dynamic register indexing
#pragma unroll
            for (int vv2 = 0; vv2 < 4; vv2++) {
                if (vv2 != v) continue;
                g[7 * vv2 + 0] += gp.x;
                g[7 * vv2 + 1] += gp.y;
                g[7 * vv2 + 2] += gp.z;
#pragma unroll
                for (int ch = 0; ch < 4; ch++) g[7 * vv2 + 3 + ch] += gc[ch];
            }
        }
    }
}

// Coverage masks -> dense fragment list ordered by (rank, pixel lane): rank r holds the r-th fragment (by record
// index, ascending, or descending when `reverse`) of every pixel that has more than r, in lane order.  A round of
// the list then holds about one fragment per pixel, so the per-pixel compositing (or its inversion) that walks
// each pixel's fragments in order takes few serial steps per round.  rank_mask[r] = the lanes with more than r
// fragments (at most 32 ranks).  Returns the number of fragments.
__device__ __forceinline__ uint32_t build_fragment_list(uint16_t* entry, uint64_t* rank_mask, uint32_t mask,
                                                        bool reverse, int lane) {
    const uint64_t lt = lanemask_lt();
    uint32_t m = mask;
    uint64_t act = ballot(m != 0u);
    uint32_t base = 0;
    for (int r = 0; act != 0ull; r++) {
        if (m != 0u) {
            const int j = reverse ? 31 - __clz(m) : __ffs(m) - 1;
            entry[base + (uint32_t)popc64(act & lt)] = (uint16_t)((lane << 5) | j);
            m &= ~(1u << j);
        }
        if (lane == 0) rank_mask[r] = act;
        base += (uint32_t)popc64(act);
        act = ballot(m != 0u);
    }
    return base;
}

// Coverage masks -> dense fragment list with each lane's fragments contiguous, ordered by record index
// (ascending, or descending when `reverse`): a round then holds few records, each with several pixels, as the
// adjoint's per-record reduction of a round wants.  Returns the lane's first entry; `total` = all fragments.
__device__ __forceinline__ uint32_t build_fragment_list_by_pixel(uint16_t* entry, uint32_t mask, bool reverse,
                                                                 int lane, uint32_t& total) {
    const uint32_t nm = (uint32_t)__popc(mask);
    const uint32_t inc = wave_inclusive_scan(nm);
    total = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    const uint32_t off = inc - nm;
    uint32_t k = off;
    uint32_t m = mask;
    while (m) {
        const int j = reverse ? 31 - __clz(m) : __ffs(m) - 1;
        entry[k++] = (uint16_t)((lane << 5) | j);
        m &= ~(1u << j);
    }
    return off;
}

// A pixel lane's walk over its own fragments of a ranked list, round by round: the position of its next
// fragment (rank `rank`), while that lies before `end`.
struct RankCursor {
    uint32_t rank = 0, base = 0;  // next rank; first position of that rank
    __device__ __forceinline__ bool next(const uint64_t* rank_mask, uint32_t count, uint64_t lt, uint32_t end,
                                         uint32_t& e) {
        if (rank >= count) return false;
        const uint64_t act = rank_mask[rank];
        e = base + (uint32_t)popc64(act & lt);
        if (e >= end) return false;
        base += (uint32_t)popc64(act);
        rank++;
        return true;
    }
};

}  // namespace dtv
