// Host-side launchers of the difftetvr_amd kernels (internal; the public surface is include/difftetvr_amd.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "dtv_device.h"

namespace dtv {

// sort.hip
void fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
int64_t scan_partitions(int64_t n);
void scan_exclusive_u32(const uint32_t* in, uint32_t* out, int64_t n, uint64_t* status, uint32_t* counter,
                        uint64_t* total, hipStream_t s);
int64_t sort_partitions(int64_t n);
// Where the sorted pairs end up, decided on the device: a selected digit that every key shares is skipped
// (its pass exits at once), so the ping-pong parity is known only there.  pass_src[i]: 0 = pass i reads the
// first buffers, 1 = the alternates, 2 = skipped.
struct SortRoute {
    uint32_t pass_src[8];
    const uint64_t* keys_out;
    const uint32_t* vals_out;
};
// n_dev (may be NULL): the element count in device memory, clamped to n (the capacity the launches are sized
// for); NULL sorts exactly n.  route (device) receives the result buffers; returns the passes launched.
int radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, int64_t n,
                     const uint64_t* n_dev, uint32_t digit_mask, uint32_t* hist, uint32_t* lookback,
                     uint32_t* counters, SortRoute* route, hipStream_t s);

// binning.hip
struct BinStats {
    uint64_t total_pairs;   // sum of (tile, face) records
    uint32_t num_tris;      // visible faces
    uint32_t key_and;       // AND of depth keys of all faces that produce records
    uint32_t key_or;        // OR of the same keys
    uint32_t pad;
};
int64_t count_pairs_blocks(int64_t T);
// Grid of nblocks 256-tet blocks over all items (item i owns blocks [block0_i, block0_{i+1})).  counts[256 * nblocks]
// gets the (tile, face) record count of every tet slot (0 for padding); block_stats the per-block partials.
void launch_count_pairs(const BatchItem* items, int nitems, int64_t nblocks, uint32_t* counts, uint32_t* block_stats,
                        hipStream_t s);
// stats.total_pairs = *scan_total; the other fields reduce the per-block partials of launch_count_pairs.
void launch_reduce_bin_stats(const uint32_t* block_stats, int64_t nblocks, const uint64_t* scan_total,
                             BinStats* stats, hipStream_t s);
// keys/vals: sort input, vals = face id (tet * 4 + face) of the record.
// Writes only records below cap (the buffers' capacity) and sets *overflow = 1 if any record lies beyond.
void launch_emit_pairs(const BatchItem* items, int nitems, int64_t nblocks, const uint32_t* offsets,
                       const uint32_t* counts, uint64_t* keys, uint32_t* vals, int64_t cap, uint32_t* overflow,
                       hipStream_t s);
// n: capacity the launch is sized for; n_dev (may be NULL): actual count in device memory (see run_count).
// The sorted keys and face ids are route->keys_out / vals_out (device pointers, set by the sort).
void launch_tile_ranges(const SortRoute* route, int64_t n, const uint64_t* n_dev, int64_t num_tiles,
                        uint2* ranges, hipStream_t s);
// out[2i], out[2i + 1] = {tet vertex indices}, {face id, 8x8 coverage mask low, high, 0} of sorted record i: the
// tile renderer's record stream.
void launch_gather_records(const BatchItem* items, int nitems, const SortRoute* route, int64_t n,
                           const uint64_t* n_dev, uint4* out, hipStream_t s);
// cost[g] = record count of the binning tile holding 8x8 sub-tile g, for g < nsub.
void launch_subtile_cost(const BatchItem* items, int nitems, int64_t nsub, const uint2* ranges, uint32_t* cost,
                         hipStream_t s);
// Launch order of the tile renderers: order[0, num_tiles) lists every tile once, heaviest first, by the tile's
// cost: its record count (cost == NULL: ranges[t].y - ranges[t].x) or cost[t].  Costs are ranked in buckets of
// 1/8 octave; the order within a bucket is unspecified.  One workgroup; needs no cleared memory.
void launch_tile_order(const uint2* ranges, const uint32_t* cost, int64_t num_tiles, uint32_t* order,
                       hipStream_t s);

// render.hip
// Per-pixel state of a rendered batch, concatenated over its items (P = all pixels of the batch); item i's
// checkpoint planes start at max_ckpt * pixel0_i.
struct PixelState {
    float4* ray_color;      // [P] composite before the clear-colour blend: premultiplied RGB, transmittance
    uint32_t* frag_count;   // [P] fragments composited
    uint32_t* last_rec;     // [P] index (in the sorted record list) one past the last composited record
    uint32_t* num_ckpt;     // [P]
    float* ckpt_t;          // [max_ckpt][P] transmittance after ckpt_n fragments
    uint32_t* ckpt_n;       // [max_ckpt][P]
    unsigned long long* counters;  // optional (NULL): [0] += gradient float atomics issued by the adjoint
    uint32_t* tile_frags;   // [tiles of the batch] fragments composited per tile (written by the forward)
};
// order (may be NULL = tile t in workgroup t): the tile each workgroup renders, from launch_tile_order.
// colour_kinds: which colour storages the batch's items use (kColoursPerVertex | kColoursPerCell); one launch
// per kind present, each rendering the tiles of its items.
constexpr uint32_t kColoursPerVertex = 1, kColoursPerCell = 2;
// round: fragments shaded per round of the forward (64, 128 or 256; other values take 128).  ntiles counts
// 8x8 sub-tiles (one workgroup each).
// sub: the batch bins tiles larger than 8x8 (the kernels then test coverage per 8x8 sub-tile).
// queue: NULL launches one workgroup per tile; else a persistent grid (the workgroups the GPU holds at once)
// takes the tiles in launch order from this device counter, which each launch zeroes first.
void launch_render_forward(const BatchItem* items, int nitems, int64_t ntiles, const uint32_t* order,
                           const uint4* recs, const uint2* ranges, PixelState ps, uint32_t colour_kinds,
                           hipStream_t s, int round = 128, bool sub = false, uint32_t* queue = nullptr);
void launch_render_adjoint(const BatchItem* items, const AdjointTarget* targets, int nitems, int64_t ntiles,
                           const uint32_t* order, const uint4* recs, const uint2* ranges, PixelState ps, int use_abs,
                           int mode, uint32_t colour_kinds, hipStream_t s, bool sub = false,
                           uint32_t* queue = nullptr);
// Adds the tile renderer adjoint's interleaved accumulator rows acc [V][8] into grad_pos [V][3] and, unless
// NULL, grad_col [V][4].
void launch_fold_vertex_grads(const float* acc, int64_t V, float* grad_pos, float* grad_col, hipStream_t s);
void launch_export_fragments(const BatchItem* items, int nitems, int64_t ntiles, const uint4* recs,
                             const uint2* ranges, PixelState ps, int max_per_pixel, int32_t* out_ids, hipStream_t s);

// proj.hip: the Projection renderer.  Per-tet buffers are indexed by the tet's slot in the batch's block grid
// (block0 * 256 + tet): pt_info [slots], pt_vert / pt_col [5 * slots] (the five vertices' NDC xyz + thickness,
// colour); tri_grad [21 * 4 * slots] (the adjoint's per-triangle-corner sums).  Records: recs[2i + 1] =
// {triangle slot id (tet * 4 + k, item-local tet), coverage mask low, high, 0}.
void launch_pt_count(const BatchItem* items, int nitems, int64_t nblocks, uint32_t* pt_info, float4* pt_vert,
                     float4* pt_col, uint32_t* counts, uint32_t* block_stats, hipStream_t s);
void launch_pt_emit(const BatchItem* items, int nitems, int64_t nblocks, const uint32_t* pt_info,
                    const float4* pt_vert, const uint32_t* offsets, const uint32_t* counts, uint64_t* keys,
                    uint32_t* vals, int64_t cap, uint32_t* overflow, hipStream_t s);
void launch_pt_gather(const BatchItem* items, int nitems, const uint32_t* pt_info, const float4* pt_vert,
                      const SortRoute* route, int64_t n, const uint64_t* n_dev, uint4* out, hipStream_t s);
void launch_pt_forward(const BatchItem* items, int nitems, int64_t ntiles, const uint32_t* order,
                       const uint32_t* pt_info, const float4* pt_vert, const float4* pt_col, const uint4* recs,
                       const uint2* ranges, PixelState ps, hipStream_t s);
void launch_pt_adjoint(const BatchItem* items, const AdjointTarget* targets, int nitems, int64_t ntiles,
                       int64_t nblocks, const uint32_t* order, const uint32_t* pt_info, const float4* pt_vert,
                       const float4* pt_col, const uint4* recs, const uint2* ranges, PixelState ps, float* tri_grad,
                       int64_t tri_grad_floats, int use_abs, hipStream_t s);

// optim.hip
struct AdamGroupDev {
    float* param;
    float* grad;
    float* m;
    float* v;
    int64_t n;
    float step_size;    // lr / (1 - beta1^t)
    float bc2_sqrt;     // sqrt(1 - beta2^t)
    float w1;           // 1 - beta1
    float beta2;
    float w2;           // 1 - beta2
    float eps;
    int32_t clamp_nonneg;
    int32_t zero_grad;
    double lr, beta1, beta2d;  // for bias corrections computed on the device
};
// step_dev (may be NULL): steps taken so far in device memory; then the bias corrections use step_dev + 1
// (step_size / bc2_sqrt of the groups are ignored) and a second launch increments it.
void launch_adam(const AdamGroupDev* groups, int ngroups, int64_t* step_dev, hipStream_t s);
void launch_mse(const float* img, const float* tgt, int64_t n, int64_t denom, float* dimg, float* loss,
                hipStream_t s);

// regularizer.hip
void launch_tet_regularizer(const int32_t* cells, const float* pos, int64_t T, float lambda, float beta,
                            float* grad_pos, hipStream_t s);

// api.hip: sets the message dtv_last_error() returns and passes the status code through.
int report_error(int code, const std::string& msg);

}  // namespace dtv
