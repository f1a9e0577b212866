// Device-side definitions shared by the gfx950 kernels of difftetvr_amd.
//
// All of this file is compiled with -ffp-contract=off: the "discrete" maths (projection, face culling, depth
// keys, fixed-point coverage) must round identically in every kernel that evaluates it (binning and both
// render passes), so that a (tile, face) record, its coverage and its order never disagree between passes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dtv {

// Face f of a tet uses local vertices kFaceTable[f][0..2]; [3] is the opposite vertex
// (Common/TetFaceTable.glsl:29-34).
__device__ constexpr int kFaceTable[4][4] = {{0, 1, 2, 3}, {1, 0, 3, 2}, {0, 2, 3, 1}, {2, 1, 3, 0}};

// Tets with a vertex beyond |ndc.xy| > kGuard are culled (the reference's rasteriser clips them).
constexpr float kGuard = 16.0f;
constexpr int kSubPixelBits = 8;

// kFaceTable packed as 2-bit fields: vertex k of face f = (kFacePacked >> (8f + 2k)) & 3.
constexpr uint32_t kFacePacked = 0x3678B1E4u;
__device__ __forceinline__ int face_vertex(int f, int k) { return (kFacePacked >> (8 * f + 2 * k)) & 3; }

// Register-only selection of element i of a 4-element array (avoids scratch for runtime indices).
__device__ __forceinline__ float sel4(int i, float a0, float a1, float a2, float a3) {
    return i == 0 ? a0 : (i == 1 ? a1 : (i == 2 ? a2 : a3));
}
__device__ __forceinline__ int32_t sel4i(int i, int32_t a0, int32_t a1, int32_t a2, int32_t a3) {
    return i == 0 ? a0 : (i == 1 ? a1 : (i == 2 ? a2 : a3));
}

// Per-view uniforms (IntersectUniform.glsl:29-41 plus the raster/tile setup).  Matrices column-major m[c*4+r].
struct ViewUniforms {
    float viewProj[16];
    float invProj[16];
    float invView[16];
    float camPos[3];
    float znear;
    int32_t width, height;
    int32_t tiles_x, tiles_y;
    int32_t tile_shift;      // log2(tile size)
    float attenuation;
    float ert_alpha;
    int32_t use_ert;
    float clear[4];
    int32_t ckpt_interval;
    int32_t max_ckpt;
    float depth_axis[3];     // unit view axis in world space, -(row 2 of the view matrix)
    int32_t linear_depth;    // 1: segment length = difference of linear eye depths (PPLL), 0: Euclidean
    int32_t sub_x, sub_y;    // 8x8 sub-tiles across and down (the tile renderers' workgroups); tiles_x/tiles_y
                             // count the binning tiles of 8 << (tile_shift - 3) pixels
};

// One render of a batch: a mesh seen through one camera.  A batch is a device array of these, ordered so that
// block0, tile0 and pixel0 ascend; its tiles are numbered globally (tile0 + local tile) in the record keys, and
// its per-pixel state is one concatenation (pixel0 + local pixel).
struct BatchItem {
    ViewUniforms U;
    const int32_t* cells;
    const float* pos;
    const float* col;       // per-vertex colours [V * 4], or NULL when cell_col is set
    const float* cell_col;  // per-cell colours [T * 4] (ColorStorage.PER_CELL), or NULL
    float* image;        // forward output [height * width * 4]
    int64_t T;           // tets
    int64_t pixel0;      // first pixel in the batch's per-pixel state
    uint32_t block0;     // first 256-tet block in the batch's count / emit grids
    uint32_t tile0;      // first tile in the batch's tile numbering
    uint32_t sub0;       // first 8x8 sub-tile in the batch's sub-tile numbering (== tile0 for 8x8 tiles)
};

// Adjoint inputs and gradient targets of one batch item (targets of items sharing a mesh may alias).
struct AdjointTarget {
    const float* dimage; // [height * width * 4]
    float* grad_pos;     // [V * 3], accumulated
    float* grad_col;     // [V * 4] (per-cell items: [T * 4]), accumulated
    float* acc;          // zeroed [V * 8] rows (xyz, rgba, pad) the adjoint accumulates vertex gradients into,
                         // folded into grad_pos / grad_col afterwards (items of one mesh share them)
};

// Item owning global block / tile `b` (the last item whose first block / tile <= b).  Uniform per workgroup.
__device__ __forceinline__ int item_of_block(const BatchItem* __restrict__ items, int n, uint32_t b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].block0 <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ int item_of_tile(const BatchItem* __restrict__ items, int n, uint32_t t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].tile0 <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Item owning global 8x8 sub-tile `t`.
__device__ __forceinline__ int item_of_sub(const BatchItem* __restrict__ items, int n, uint32_t t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].sub0 <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

struct float3_ {
    float x, y, z;
};

__device__ __forceinline__ float3_ f3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ float3_ sub(float3_ a, float3_ b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float3_ add(float3_ a, float3_ b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ float3_ scale(float3_ a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(float3_ a, float3_ b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3_ cross(float3_ a, float3_ b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(float3_ a) { return sqrtf(dot(a, a)); }

__device__ __forceinline__ float3_ load_pos(const float* __restrict__ pos, int32_t v) {
    const float* p = pos + 3 * (int64_t)v;
    return {p[0], p[1], p[2]};
}

__device__ __forceinline__ float signf_(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// Projected tet: NDC of the 4 vertices and the visibility mask of its faces.
struct ProjectedTet {
    float3_ p[4];     // world positions
    float ndc[4][3];
    uint32_t visible; // bit f set when face f is a front face (GenerateTrianglesIntersect.glsl:104-121)
    bool culled;
};

// Projection and face culling of one tet (GenerateTrianglesIntersect.glsl:80-121).
__device__ __forceinline__ void project_tet(const ViewUniforms& U, const int32_t* __restrict__ cells,
                                            const float* __restrict__ pos, int64_t tet, ProjectedTet& T) {
    const int4 idx = *reinterpret_cast<const int4*>(cells + 4 * tet);
    const int32_t vid[4] = {idx.x, idx.y, idx.z, idx.w};
    T.culled = false;
    T.visible = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float3_ p = load_pos(pos, vid[i]);
        T.p[i] = p;
        float c[4];
#pragma unroll
        for (int r = 0; r < 4; r++)
            c[r] = ((U.viewProj[0 * 4 + r] * p.x + U.viewProj[1 * 4 + r] * p.y) + U.viewProj[2 * 4 + r] * p.z) +
                   U.viewProj[3 * 4 + r] * 1.0f;
        if (!(c[3] >= U.znear)) T.culled = true;
        T.ndc[i][0] = c[0] / c[3];
        T.ndc[i][1] = c[1] / c[3];
        T.ndc[i][2] = c[2] / c[3];
        if (!(fabsf(T.ndc[i][0]) <= kGuard) || !(fabsf(T.ndc[i][1]) <= kGuard)) T.culled = true;
    }
    if (T.culled) return;
    const float3_ cam = f3(U.camPos[0], U.camPos[1], U.camPos[2]);
#pragma unroll
    for (int f = 0; f < 4; f++) {
        float3_ a = T.p[kFaceTable[f][0]], b = T.p[kFaceTable[f][1]], c = T.p[kFaceTable[f][2]];
        float3_ o = T.p[kFaceTable[f][3]];
        float3_ n = cross(sub(b, a), sub(c, a));
        float d = -dot(n, a);
        float sc = signf_(dot(cam, n) + d);
        float so = signf_(dot(o, n) + d);
        if (sc != 0.0f && sc != so) T.visible |= 1u << f;
    }
}

// NDC x, y of face f's three vertices only, with project_tet's arithmetic (bit-identical): for a record whose
// tet already passed culling.  T.ndc of the fourth vertex is left 0; face_screen does not read it.
__device__ __forceinline__ void project_face(const ViewUniforms& U, const int4 idx, const float* __restrict__ pos,
                                             int f, ProjectedTet& T) {
    const int32_t vid[4] = {idx.x, idx.y, idx.z, idx.w};
#pragma unroll
    for (int i = 0; i < 4; i++) T.ndc[i][0] = T.ndc[i][1] = T.ndc[i][2] = 0.f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int v = face_vertex(f, k);
        const float3_ p = load_pos(pos, sel4i(v, vid[0], vid[1], vid[2], vid[3]));
        float c[4];
#pragma unroll
        for (int r = 0; r < 4; r++)
            c[r] = ((U.viewProj[0 * 4 + r] * p.x + U.viewProj[1 * 4 + r] * p.y) + U.viewProj[2 * 4 + r] * p.z) +
                   U.viewProj[3 * 4 + r] * 1.0f;
        const float nx = c[0] / c[3], ny = c[1] / c[3];
        // scatter into slot v (register selects; v is 0..3)
        T.ndc[0][0] = v == 0 ? nx : T.ndc[0][0];
        T.ndc[0][1] = v == 0 ? ny : T.ndc[0][1];
        T.ndc[1][0] = v == 1 ? nx : T.ndc[1][0];
        T.ndc[1][1] = v == 1 ? ny : T.ndc[1][1];
        T.ndc[2][0] = v == 2 ? nx : T.ndc[2][0];
        T.ndc[2][1] = v == 2 ? ny : T.ndc[2][1];
        T.ndc[3][0] = v == 3 ? nx : T.ndc[3][0];
        T.ndc[3][1] = v == 3 ? ny : T.ndc[3][1];
    }
}

// Depth key of face f: NDC z of the face centroid (ComputeTriangleDepths.glsl:99-108), as raw float bits
// (keys are >= 0 after the near-plane cull, so bit order equals value order).
__device__ __forceinline__ uint32_t face_key(const ProjectedTet& T, int f) {
    const float z0 = sel4(face_vertex(f, 0), T.ndc[0][2], T.ndc[1][2], T.ndc[2][2], T.ndc[3][2]);
    const float z1 = sel4(face_vertex(f, 1), T.ndc[0][2], T.ndc[1][2], T.ndc[2][2], T.ndc[3][2]);
    const float z2 = sel4(face_vertex(f, 2), T.ndc[0][2], T.ndc[1][2], T.ndc[2][2], T.ndc[3][2]);
    float z = ((z0 + z1) + z2) / 3.0f;
    return __float_as_uint(z);
}

// Fixed-point screen triangle of face f, oriented to positive area.
struct ScreenTri {
    int32_t X[3], Y[3];
    int64_t area2;
};

__device__ __forceinline__ void face_screen(const ViewUniforms& U, const ProjectedTet& T, int f, ScreenTri& S) {
    const float hw = 0.5f * (float)U.width, hh = 0.5f * (float)U.height;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int v = face_vertex(f, k);
        const float qx = sel4(v, T.ndc[0][0], T.ndc[1][0], T.ndc[2][0], T.ndc[3][0]);
        const float qy = sel4(v, T.ndc[0][1], T.ndc[1][1], T.ndc[2][1], T.ndc[3][1]);
        float xf = qx * hw + hw;
        float yf = qy * hh + hh;
        S.X[k] = (int32_t)rintf(xf * 256.0f);
        S.Y[k] = (int32_t)rintf(yf * 256.0f);
    }
    int64_t area = (int64_t)(S.X[1] - S.X[0]) * (S.Y[2] - S.Y[0]) - (int64_t)(S.Y[1] - S.Y[0]) * (S.X[2] - S.X[0]);
    if (area < 0) {
        int32_t tx = S.X[1], ty = S.Y[1];
        S.X[1] = S.X[2];
        S.Y[1] = S.Y[2];
        S.X[2] = tx;
        S.Y[2] = ty;
        area = -area;
    }
    S.area2 = area;
}

__device__ __forceinline__ int64_t floordiv256(int64_t a) { return a >= 0 ? a / 256 : -((-a + 255) / 256); }

// Inclusive pixel range [x0,x1]x[y0,y1] whose centres lie in the triangle's bounding box, clipped to the view.
// Empty (x0 > x1) for zero-area triangles.
__device__ __forceinline__ void tri_pixel_bbox(const ViewUniforms& U, const ScreenTri& S, int32_t& x0, int32_t& x1,
                                               int32_t& y0, int32_t& y1) {
    int32_t mnx = min(S.X[0], min(S.X[1], S.X[2])), mxx = max(S.X[0], max(S.X[1], S.X[2]));
    int32_t mny = min(S.Y[0], min(S.Y[1], S.Y[2])), mxy = max(S.Y[0], max(S.Y[1], S.Y[2]));
    int64_t a0 = -floordiv256(-(int64_t)mnx + 128), a1 = floordiv256((int64_t)mxx - 128);
    int64_t b0 = -floordiv256(-(int64_t)mny + 128), b1 = floordiv256((int64_t)mxy - 128);
    a0 = a0 < 0 ? 0 : a0;
    b0 = b0 < 0 ? 0 : b0;
    a1 = a1 > U.width - 1 ? U.width - 1 : a1;
    b1 = b1 > U.height - 1 ? U.height - 1 : b1;
    if (S.area2 == 0) {
        a0 = 1;
        a1 = 0;
    }
    x0 = (int32_t)a0;
    x1 = (int32_t)a1;
    y0 = (int32_t)b0;
    y1 = (int32_t)b1;
}

// Camera ray direction through the centre of pixel (x, y) (IntersectRasterization.glsl:141-144).
__device__ __forceinline__ float3_ ray_direction(const ViewUniforms& U, int x, int y) {
    float fx = (float)x + 0.5f, fy = (float)y + 0.5f;
    float nx = 2.0f * (fx / (float)U.width) - 1.0f;
    float ny = 2.0f * (fy / (float)U.height) - 1.0f;
    float t[3];
#pragma unroll
    for (int r = 0; r < 3; r++)
        t[r] = ((U.invProj[0 * 4 + r] * nx + U.invProj[1 * 4 + r] * ny) + U.invProj[2 * 4 + r] * 1.0f) +
               U.invProj[3 * 4 + r] * 1.0f;
    float l = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    float n0 = t[0] / l, n1 = t[1] / l, n2 = t[2] / l;
    float d[3];
#pragma unroll
    for (int r = 0; r < 3; r++)
        d[r] = ((U.invView[0 * 4 + r] * n0 + U.invView[1 * 4 + r] * n1) + U.invView[2 * 4 + r] * n2) +
               U.invView[3 * 4 + r] * 0.0f;
    return f3(d[0], d[1], d[2]);
}

}  // namespace dtv
