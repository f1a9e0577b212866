// Per-view tet projection, front-face culling and screen-tile binning on gfx950.
//
// Replaces GenerateTrianglesIntersect.glsl:74-132 (per-tet projection + front faces, one atomic append each),
// ComputeTriangleDepths.glsl:72-110 (centroid-depth keys) and the rasteriser's fragment generation.  Each front
// face becomes one (tile, face) record per 8x8 screen tile of its pixel-centre bounding box that the face's edge
// functions do not exclude (tile_may_cover), keyed by (tile << 32 | depth-key bits).  Counting and emission are
// two passes over the tets separated by a device-wide scan, so records are written in (tet, face, tile) order
// and the stable sort keeps that order for equal depth keys.  The sort moves face ids; k_gather_records then
// lays the sorted records out for the tile renderer as {face id, exact 64-bit coverage mask of the tile}
// (one thread per record, so the per-pixel edge tests are spread evenly over the GPU).
#include "dtv_kernels.h"
#include "dtv_prims.h"
#include "raster.h"

namespace dtv {

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
                const FaceBins b = face_bins(U, PT, f);
                const uint32_t cf = face_record_count(b);
                c += cf;
                if (cf) {
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
// keys/vals are the sort input, vals = the record's face id tet * 4 + face (the stable sort keeps equal keys in
// emit order, so the payload need not encode it).
// Records at index >= cap (the capacity of keys/vals) are not written; *overflow is then set to 1.
__global__ __launch_bounds__(256) void k_emit_pairs(const BatchItem* __restrict__ items, int nitems,
                                                     const uint32_t* __restrict__ offsets,
                                                     const uint32_t* __restrict__ counts,
                                                     uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                     uint64_t cap, uint32_t* __restrict__ overflow) {
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
        const FaceBins b = face_bins(U, PT, f);
        const uint64_t k = (uint64_t)face_key(PT, f);
        const uint32_t id = (uint32_t)(tet * 4 + f);
        const bool test = test_tiles(b);
        // Bounded by the stored count: the pass recomputes the same records, but never writes past its slots.
        for (int ty = b.ty0; ty <= b.ty1 && left > 0; ty++)
            for (int tx = b.tx0; tx <= b.tx1 && left > 0; tx++) {
                if (test && !tile_may_cover(b.S, tx, ty, b.shift)) continue;
                if (o < cap) {
                    keys[o] = ((uint64_t)(B.tile0 + (uint32_t)(ty * U.tiles_x + tx)) << 32) | k;
                    vals[o] = id;
                } else {
                    *overflow = 1u;
                }
                o++;
                left--;
            }
    }
}

__global__ void k_tile_ranges(const SortRoute* __restrict__ route, int64_t n_cap, const uint64_t* __restrict__ n_dev,
                              uint2* __restrict__ ranges) {
    const uint64_t* __restrict__ keys = route->keys_out;
    const int64_t n = run_count(n_cap, n_dev);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t t = (uint32_t)(keys[i] >> 32);
        if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != t) ranges[t].x = (uint32_t)i;
        if (i == n - 1 || (uint32_t)(keys[i + 1] >> 32) != t) ranges[t].y = (uint32_t)(i + 1);
    }
}

// Tile launch order, heaviest first.  A tile's rendering time grows with its records and fragments, and tiles are
// handed to the CUs in workgroup order: in row-major order the heavy tiles of the image centre start mid-launch
// and the last of them finish long after the rest of the GPU has gone idle.  Started first, the light tiles fill
// in behind them.  Costs are ranked coarsely (8 buckets per octave, 256 buckets for 32 bits).
__device__ __forceinline__ uint32_t cost_bucket(uint32_t c) {
    const uint32_t v = c < 0xFFFFFFFFu ? c + 1u : c;  // >= 1
    const int e = 31 - __clz((int)v);
    const uint32_t m = e >= 3 ? (v >> (e - 3)) & 7u : (v << (3 - e)) & 7u;
    return 8u * (uint32_t)e + m;
}

// The lanes of the wave whose bucket equals this lane's (eq, live lanes only) and this lane's rank among them.
// Every lane of the wave must take part.
__device__ __forceinline__ void bucket_match(bool live, uint32_t b, uint64_t& eq, uint32_t& rank) {
    eq = ballot(live);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint64_t bk = ballot(((b >> k) & 1u) != 0u);
        eq &= ((b >> k) & 1u) ? bk : ~bk;
    }
    rank = (uint32_t)popc64(eq & lanemask_lt());
}

// Tiles per thread per step: their costs are loaded together, so the loads overlap (one step per pass up to
// 16K tiles, a 1024x1024 view).
constexpr int kOrderPer = 16;

__device__ __forceinline__ void load_buckets(const uint2* __restrict__ ranges, const uint32_t* __restrict__ cost,
                                             uint32_t i0, uint32_t n, uint32_t b[kOrderPer]) {
    uint32_t c[kOrderPer];
#pragma unroll
    for (int q = 0; q < kOrderPer; q++) {
        const uint32_t i = i0 + 1024u * q + threadIdx.x;
        c[q] = 0u;
        if (i < n) c[q] = cost != nullptr ? cost[i] : ranges[i].y - ranges[i].x;
    }
#pragma unroll
    for (int q = 0; q < kOrderPer; q++) b[q] = cost_bucket(c[q]);
}

__global__ __launch_bounds__(1024) void k_tile_order(const uint2* __restrict__ ranges,
                                                     const uint32_t* __restrict__ cost, uint32_t n,
                                                     uint32_t* __restrict__ order) {
    __shared__ uint32_t count[256], start[256];
    const uint32_t t = threadIdx.x;
    if (t < 256) count[t] = 0u;
    __syncthreads();
    // trip counts are uniform over the workgroup, so every lane of a wave reaches the ballots
    for (uint32_t i0 = 0; i0 < n; i0 += 1024u * kOrderPer) {
        uint32_t b[kOrderPer];
        load_buckets(ranges, cost, i0, n, b);
#pragma unroll
        for (int q = 0; q < kOrderPer; q++) {
            const bool live = i0 + 1024u * q + t < n;
            uint64_t eq;
            uint32_t rank;
            bucket_match(live, b[q], eq, rank);
            if (live && rank == 0u) atomicAdd(&count[b[q]], (uint32_t)popc64(eq));
        }
    }
    __syncthreads();
    if (t < 64) {
        // start[b] = tiles in the buckets above b; lane l takes buckets 255 - 4l down to 252 - 4l
        uint32_t c[4], sum = 0u;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            c[k] = count[255 - 4 * t - k];
            sum += c[k];
        }
        uint32_t run = wave_inclusive_scan(sum) - sum;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            start[255 - 4 * t - k] = run;
            run += c[k];
        }
    }
    __syncthreads();
    const int lane = (int)(t & 63u);
    for (uint32_t i0 = 0; i0 < n; i0 += 1024u * kOrderPer) {
        uint32_t b[kOrderPer];
        load_buckets(ranges, cost, i0, n, b);
#pragma unroll
        for (int q = 0; q < kOrderPer; q++) {
            const uint32_t i = i0 + 1024u * q + t;
            const bool live = i < n;
            uint64_t eq;
            uint32_t rank;
            bucket_match(live, b[q], eq, rank);
            uint32_t base = 0u;
            if (live && rank == 0u) base = atomicAdd(&start[b[q]], (uint32_t)popc64(eq));
            const int leader = eq != 0ull ? __ffsll((unsigned long long)eq) - 1 : lane;
            base = __shfl(base, leader, 64);
            if (live) order[base + rank] = i;
        }
    }
}

// Sorted record i = out[2i], out[2i + 1] = {the tet's 4 vertex indices}, {face id, coverage mask low, high, 0}:
// the face sorted_ids[i] is projected again (the same arithmetic as binning) and its exact coverage of the
// record's tile computed.  Carrying the vertex indices saves the renderer one dependent load per record.
__global__ __launch_bounds__(256) void k_gather_records(const BatchItem* __restrict__ items, int nitems,
                                                         const SortRoute* __restrict__ route, int64_t n_cap,
                                                         const uint64_t* __restrict__ n_dev,
                                                         uint4* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= run_count(n_cap, n_dev)) return;
    const uint64_t* __restrict__ sorted_keys = route->keys_out;
    const uint32_t* __restrict__ sorted_ids = route->vals_out;
    const uint32_t id = sorted_ids[i];
    const uint32_t gtile = (uint32_t)(sorted_keys[i] >> 32);
    const BatchItem& B = items[item_of_tile(items, nitems, gtile)];
    const ViewUniforms& U = B.U;
    const int tile = (int)(gtile - B.tile0);
    const int64_t tet = id >> 2;
    const int4 vid = *reinterpret_cast<const int4*>(B.cells + 4 * tet);
    ProjectedTet PT;
    project_face(U, vid, B.pos, (int)(id & 3), PT);  // the tet passed culling when its records were emitted
    ScreenTri S;
    face_screen(U, PT, id & 3, S);
    // larger binning tiles: each 8x8 sub-tile's renderer tests its own pixels (no mask here)
    const uint64_t m =
        U.tile_shift == 3 ? tile_coverage(S, tile % U.tiles_x, tile / U.tiles_x, U.width, U.height) : 0ull;
    out[2 * i] = make_uint4((uint32_t)vid.x, (uint32_t)vid.y, (uint32_t)vid.z, (uint32_t)vid.w);
    out[2 * i + 1] = make_uint4(id, (uint32_t)m, (uint32_t)(m >> 32), 0u);
}

// cost[g] = records of the binning tile that holds 8x8 sub-tile g (the forward's launch order for binning tiles
// larger than 8x8, whose records every sub-tile of the tile walks).
__global__ __launch_bounds__(256) void k_subtile_cost(const BatchItem* __restrict__ items, int nitems, int64_t nsub,
                                                       const uint2* __restrict__ ranges, uint32_t* __restrict__ cost) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= nsub) return;
    const BatchItem& B = items[item_of_sub(items, nitems, (uint32_t)g)];
    const ViewUniforms& U = B.U;
    const int sub = (int)((uint32_t)g - B.sub0);
    const int k = U.tile_shift - 3;
    const uint32_t t = B.tile0 + (uint32_t)(((sub / U.sub_x) >> k) * U.tiles_x + ((sub % U.sub_x) >> k));
    const uint2 r = ranges[t];
    cost[g] = r.y - r.x;
}

void launch_subtile_cost(const BatchItem* items, int nitems, int64_t nsub, const uint2* ranges, uint32_t* cost,
                         hipStream_t s) {
    if (nsub <= 0) return;
    hipLaunchKernelGGL(k_subtile_cost, dim3((unsigned)((nsub + 255) / 256)), dim3(256), 0, s, items, nitems, nsub,
                       ranges, cost);
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
                       const uint32_t* counts, uint64_t* keys, uint32_t* vals, int64_t cap, uint32_t* overflow,
                       hipStream_t s) {
    if (nblocks <= 0) return;
    hipLaunchKernelGGL(k_emit_pairs, dim3((unsigned)nblocks), dim3(256), 0, s, items, nitems, offsets, counts, keys,
                       vals, (uint64_t)cap, overflow);
}

void launch_tile_ranges(const SortRoute* route, int64_t n, const uint64_t* n_dev, int64_t num_tiles,
                        uint2* ranges, hipStream_t s) {
    fill_u32(reinterpret_cast<uint32_t*>(ranges), 0u, 2 * num_tiles, s);
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_tile_ranges, dim3((unsigned)blocks), dim3(256), 0, s, route, n, n_dev, ranges);
}

void launch_tile_order(const uint2* ranges, const uint32_t* cost, int64_t num_tiles, uint32_t* order,
                       hipStream_t s) {
    if (num_tiles <= 0) return;
    hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(1024), 0, s, ranges, cost, (uint32_t)num_tiles, order);
}

void launch_gather_records(const BatchItem* items, int nitems, const SortRoute* route, int64_t n,
                           const uint64_t* n_dev, uint4* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_records, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, items, nitems, route,
                       n, n_dev, out);
}

}  // namespace dtv
