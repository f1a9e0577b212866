// Per-view tet projection, front-face culling and screen-tile binning on gfx950.
//
// Replaces GenerateTrianglesIntersect.glsl:74-132 (per-tet projection + front faces, one atomic append each),
// ComputeTriangleDepths.glsl:72-110 (centroid-depth keys) and the rasteriser's fragment generation: each front
// face becomes one (tile, face) record per screen tile its pixel-centre bounding box touches, keyed by
// (tile << 32 | depth-key bits).  Counting and emission are two passes over the tets separated by a
// device-wide scan, so records are written in (tet, face) order and the stable sort keeps that order for
// equal depth keys.
#include "dtv_kernels.h"
#include "dtv_prims.h"

namespace dtv {

struct FaceBins {
    int32_t tx0, tx1, ty0, ty1;
    uint32_t count;
};

__device__ __forceinline__ FaceBins face_bins(const ViewUniforms& U, const ProjectedTet& T, int f) {
    FaceBins b{0, -1, 0, -1, 0u};
    ScreenTri S;
    face_screen(U, T, f, S);
    int32_t x0, x1, y0, y1;
    tri_pixel_bbox(U, S, x0, x1, y0, y1);
    if (x0 > x1 || y0 > y1) return b;
    b.tx0 = x0 >> U.tile_shift;
    b.tx1 = x1 >> U.tile_shift;
    b.ty0 = y0 >> U.tile_shift;
    b.ty1 = y1 >> U.tile_shift;
    b.count = (uint32_t)((b.tx1 - b.tx0 + 1) * (b.ty1 - b.ty0 + 1));
    return b;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_and_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v &= __shfl_xor(v, d, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v |= __shfl_xor(v, d, 64);
    return v;
}

// One thread per tet slot of the batch's block grid; counts[slot] = records of the tet (all its faces).
__global__ __launch_bounds__(256) void k_count_pairs(const BatchItem* __restrict__ items, int nitems,
                                                      uint32_t* __restrict__ counts, uint4* __restrict__ block_stats) {
    const BatchItem& B = items[item_of_block(items, nitems, blockIdx.x)];
    const ViewUniforms& U = B.U;
    const int64_t tet = (int64_t)(blockIdx.x - B.block0) * 256 + threadIdx.x;
    uint32_t c = 0, ntri = 0, kand = 0xFFFFFFFFu, kor = 0u;
    if (tet < B.T) {
        ProjectedTet PT;
        project_tet(U, B.cells, B.pos, tet, PT);
        if (!PT.culled) {
#pragma unroll
            for (int f = 0; f < 4; f++) {
                if (!(PT.visible & (1u << f))) continue;
                ntri++;
                FaceBins b = face_bins(U, PT, f);
                c += b.count;
                if (b.count) {
                    uint32_t k = face_key(PT, f);
                    kand &= k;
                    kor |= k;
                }
            }
        }
    }
    counts[(int64_t)blockIdx.x * 256 + threadIdx.x] = c;
    // Per-block partials (no global atomics: one address hit by every wave serialises at the memory side).
    __shared__ uint4 s_part[4];
    const uint32_t wtri = wave_sum_u32(ntri);
    const uint32_t wand = wave_and_u32(kand);
    const uint32_t wor = wave_or_u32(kor);
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

// Reduces the per-block partials of k_count_pairs; the record total comes from the scan.
__global__ __launch_bounds__(256) void k_reduce_bin_stats(const uint4* __restrict__ block_stats, int64_t nblocks,
                                                           const uint64_t* __restrict__ scan_total,
                                                           BinStats* __restrict__ stats) {
    __shared__ uint4 s_part[4];
    uint32_t ntri = 0, kand = 0xFFFFFFFFu, kor = 0u;
    for (int64_t i = threadIdx.x; i < nblocks; i += blockDim.x) {
        const uint4 b = block_stats[i];
        ntri += b.x;
        kand &= b.y;
        kor |= b.z;
    }
    const uint32_t wtri = wave_sum_u32(ntri), wand = wave_and_u32(kand), wor = wave_or_u32(kor);
    if (lane_id() == 0) s_part[threadIdx.x / 64] = make_uint4(wtri, wand, wor, 0u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint4 r = s_part[0];
        for (int w = 1; w < 4; w++) {
            r.x += s_part[w].x;
            r.y &= s_part[w].y;
            r.z |= s_part[w].z;
        }
        stats->total_pairs = *scan_total;
        stats->num_tris = r.x;
        stats->key_and = r.y;
        stats->key_or = r.z;
    }
}

// Same grid as k_count_pairs: writes the tet's records at offsets[slot], faces in order, tiles row-major.
__global__ __launch_bounds__(256) void k_emit_pairs(const BatchItem* __restrict__ items, int nitems,
                                                     const uint32_t* __restrict__ offsets,
                                                     const uint32_t* __restrict__ counts,
                                                     uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const BatchItem& B = items[item_of_block(items, nitems, blockIdx.x)];
    const ViewUniforms& U = B.U;
    const int64_t tet = (int64_t)(blockIdx.x - B.block0) * 256 + threadIdx.x;
    if (tet >= B.T) return;
    const int64_t slot = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t left = counts[slot];
    if (left == 0) return;
    uint32_t o = offsets[slot];
    ProjectedTet PT;
    project_tet(U, B.cells, B.pos, tet, PT);
#pragma unroll
    for (int f = 0; f < 4; f++) {
        if (!(PT.visible & (1u << f)) || left == 0) continue;
        FaceBins b = face_bins(U, PT, f);
        if (b.count == 0) continue;
        const uint64_t k = (uint64_t)face_key(PT, f);
        const uint32_t id = (uint32_t)(tet * 4 + f);
        const int bw = b.tx1 - b.tx0 + 1;
        // Bounded by the stored count: the pass recomputes the same bins, but never writes past its slots.
        const uint32_t n = min(left, b.count);
        for (uint32_t i = 0; i < n; i++) {
            const int tx = b.tx0 + (int)(i % (uint32_t)bw), ty = b.ty0 + (int)(i / (uint32_t)bw);
            keys[o + i] = ((uint64_t)(B.tile0 + (uint32_t)(ty * U.tiles_x + tx)) << 32) | k;
            vals[o + i] = id;
        }
        o += n;
        left -= n;
    }
}

__global__ void k_tile_ranges(const uint64_t* __restrict__ keys, int64_t n, uint2* __restrict__ ranges) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t t = (uint32_t)(keys[i] >> 32);
        if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != t) ranges[t].x = (uint32_t)i;
        if (i == n - 1 || (uint32_t)(keys[i + 1] >> 32) != t) ranges[t].y = (uint32_t)(i + 1);
    }
}

int64_t count_pairs_blocks(int64_t T) { return T > 0 ? (T + 255) / 256 : 0; }

void launch_count_pairs(const BatchItem* items, int nitems, int64_t nblocks, uint32_t* counts, uint32_t* block_stats,
                        hipStream_t s) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(k_count_pairs, dim3((unsigned)nblocks), dim3(256), 0, s, items, nitems, counts,
                       reinterpret_cast<uint4*>(block_stats));
}

void launch_reduce_bin_stats(const uint32_t* block_stats, int64_t nblocks, const uint64_t* scan_total,
                             BinStats* stats, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_bin_stats, dim3(1), dim3(256), 0, s, reinterpret_cast<const uint4*>(block_stats),
                       nblocks, scan_total, stats);
}

void launch_emit_pairs(const BatchItem* items, int nitems, int64_t nblocks, const uint32_t* offsets,
                       const uint32_t* counts, uint64_t* keys, uint32_t* vals, hipStream_t s) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(k_emit_pairs, dim3((unsigned)nblocks), dim3(256), 0, s, items, nitems, offsets, counts, keys,
                       vals);
}

void launch_tile_ranges(const uint64_t* sorted_keys, int64_t n, int64_t num_tiles, uint2* ranges, hipStream_t s) {
    fill_u32(reinterpret_cast<uint32_t*>(ranges), 0u, 2 * num_tiles, s);
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_tile_ranges, dim3((unsigned)blocks), dim3(256), 0, s, sorted_keys, n, ranges);
}

}  // namespace dtv
