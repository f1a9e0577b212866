// C ABI of difftetvr_amd (include/difftetvr_amd.h): views, planning of the per-view binning/sort/render
// pipeline, and the optimiser entry points.  Host C++; all device work is enqueued on the caller's stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/difftetvr_amd.h"
#include "dtv_kernels.h"

using namespace dtv;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

}  // namespace

// Error reporting for the other host sources of the library (dtv_last_error).
int dtv::report_error(int code, const std::string& msg) { return fail(code, msg); }

namespace {

#define DTV_HIP(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail(DTV_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device buffer that grows geometrically and is never shrunk.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p != nullptr) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
        }
        p = nullptr;
        cap = 0;
        size_t nb = bytes + bytes / 4 + 256;
        hipError_t e = hipMalloc(&p, nb);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = nb;
        return hipSuccess;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

constexpr int kMaxCheckpoints = 32;

// Camera uniforms: perspective RH, depth in [0,1], y flipped; built in double, rounded once to float.
// Column-major 4x4 (m[c*4 + r]).
void mul4(const double* A, const double* B, double* C) {
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++)
            C[c * 4 + r] = ((A[0 * 4 + r] * B[c * 4 + 0] + A[1 * 4 + r] * B[c * 4 + 1]) + A[2 * 4 + r] * B[c * 4 + 2]) +
                           A[3 * 4 + r] * B[c * 4 + 3];
}

bool inv4(const double* m, double* out) {
    double a[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            a[r][c] = m[c * 4 + r];
            a[r][c + 4] = (r == c) ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; col++) {
        int piv = col;
        for (int r = col + 1; r < 4; r++)
            if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
        if (a[piv][col] == 0.0) return false;
        if (piv != col)
            for (int k = 0; k < 8; k++) std::swap(a[col][k], a[piv][k]);
        const double inv = 1.0 / a[col][col];
        for (int k = 0; k < 8; k++) a[col][k] *= inv;
        for (int r = 0; r < 4; r++) {
            if (r == col) continue;
            const double f = a[r][col];
            if (f == 0.0) continue;
            for (int k = 0; k < 8; k++) a[r][k] -= f * a[col][k];
        }
    }
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) out[c * 4 + r] = a[r][c + 4];
    return true;
}

bool camera_uniforms(const dtv_camera& cam, double u[68]) {
    double V[16], P[16] = {0}, iP[16] = {0}, VP[16], iV[16];
    for (int i = 0; i < 16; i++) V[i] = (double)cam.view[i];
    const double th = std::tan(0.5 * (double)cam.fovy);
    const double aspect = (double)cam.width / (double)cam.height;
    P[0] = 1.0 / (aspect * th);
    P[5] = -1.0 / th;
    P[10] = (double)cam.zfar / ((double)cam.znear - (double)cam.zfar);
    P[11] = -1.0;
    P[14] = -((double)cam.zfar * (double)cam.znear) / ((double)cam.zfar - (double)cam.znear);
    iP[0] = 1.0 / P[0];
    iP[5] = 1.0 / P[5];
    iP[11] = 1.0 / P[14];
    iP[14] = -1.0;
    iP[15] = P[10] / P[14];
    mul4(P, V, VP);
    if (!inv4(V, iV)) return false;
    std::memcpy(u + 0, VP, sizeof(VP));
    std::memcpy(u + 16, V, sizeof(V));
    std::memcpy(u + 32, iP, sizeof(iP));
    std::memcpy(u + 48, iV, sizeof(iV));
    u[64] = iV[12];
    u[65] = iV[13];
    u[66] = iV[14];
    u[67] = 0.0;
    return true;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Sizes, buffers and alignment of a mesh; exactly one colour storage (per vertex or per cell).
int check_mesh(const dtv_mesh& m, const std::string& f) {
    if (m.num_cells < 0 || m.num_vertices < 0 || m.num_cells > (int64_t)(1u << 28))
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": invalid mesh size");
    if ((m.colors != nullptr) == (m.cell_colors != nullptr))
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": set exactly one of colors (per vertex) and cell_colors (per cell)");
    if (m.num_cells > 0 && (!m.cells || !m.positions)) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": NULL mesh buffer");
    if (!aligned16(m.cells) || !aligned16(m.colors) || !aligned16(m.cell_colors))
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": cells and colours must be 16-byte aligned");
    return DTV_OK;
}

uint32_t colour_kind(const dtv_mesh& m) { return m.cell_colors ? kColoursPerCell : kColoursPerVertex; }

}  // namespace

// ---- allocation-free rendering (dtv_render_noalloc) ---------------------------------------------------------
// Every intermediate buffer of dtv_render_noalloc lives in this one array, which each GPU the library runs on
// holds until the process ends.  Limits: kNaMaxTets tets, kNaMaxPixels pixels (kNaMaxTiles 8x8 tiles) and
// kNaMaxRecords (tile, face) records per render.  Every call on a GPU reuses that GPU's copy (the buffers the
// pipeline needs cleared are cleared by its own fill kernels), so calls that overlap on one GPU race.
constexpr int64_t kNaMaxTets = 1 << 18;          // 262,144
constexpr int64_t kNaMaxPixels = 1 << 18;        // e.g. 512 x 512
constexpr int64_t kNaMaxTiles = 1 << 13;         // 8,192
constexpr int64_t kNaMaxRecords = 3 << 18;       // 786,432
constexpr size_t kNaScratchBytes = 56ull << 20;  // 56 MiB, >= the layout below at the limits (checked per call)
static __device__ __align__(256) unsigned char scratch[kNaScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each device (set up once per device)

// Must run with `device` current.  The host-side name `scratch` is not a device address: kernels and copies
// use the address looked up here.
int setup_noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(DTV_ERR_INVALID_ARGUMENT, "device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return DTV_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        return fail(DTV_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return DTV_OK;
}

__global__ void k_store_item(BatchItem item, BatchItem* __restrict__ dst) { *dst = item; }
__global__ void k_copy_flag(const uint32_t* __restrict__ src, int32_t* __restrict__ dst) { *dst = (int32_t)*src; }

}  // namespace

struct dtv_ctx {
    int device = 0;
};

struct dtv_view {
    dtv_ctx* ctx = nullptr;
    // the batch of the last render (host copy; the device copy is dItems)
    std::vector<BatchItem> items;
    std::vector<int64_t> item_V;
    uint32_t colour_kinds = 0;             // colour storages of the batch's items (kColoursPerVertex | PerCell)
    BatchItem* pinnedItems = nullptr;
    int pinned_items_cap = 0;
    hipEvent_t items_copied = nullptr;     // the last upload out of pinnedItems
    bool items_pending = false;
    AdjointTarget* pinnedTargets = nullptr;
    int pinned_targets_cap = 0;
    hipEvent_t targets_copied = nullptr;   // the last upload out of pinnedTargets
    bool targets_pending = false;
    BinStats* pinnedStats = nullptr;
    DevBuf dItems, dTargets, counts, offsets, scan_status, small, block_stats, keys, vals, keys2, vals2,
        recs, hist,
        lookback, ranges;
    DevBuf ray_color, frag_count, last_rec, num_ckpt, ckpt_t, ckpt_n;
    DevBuf tile_frags, tile_order;  // per tile: forward's fragments; launch orders of the forward and adjoint
    DevBuf pt_info, pt_vert, pt_col, tri_grad;  // Projection renderer: per-tet triangles, per-corner gradients
    DevBuf vacc;  // adjoint: interleaved per-vertex gradient rows of the distinct targets
    bool projection = false;        // the last render ran the Projection renderer
    bool sub_tiles = false;         // the last render binned tiles larger than 8x8
    bool queued = false;            // the last render took its tiles from a work queue (the adjoint does too)
    bool cost_ordered = true;       // the last render launched its tiles heaviest first
    int64_t planned_cap = 0;  // > 0: planned (synchronisation-free) rendering with this record capacity
    // state of the last render
    bool rendered = false;
    int64_t total_blocks = 0, total_tiles = 0, total_subtiles = 0, total_pixels = 0;
    int64_t num_pairs = 0, num_tris = 0;
    int sort_passes = 0;
    // per-stage timers
    bool timing = false;
    hipEvent_t ev[DTV_NUM_STAGES][2] = {};
    bool ev_recorded[DTV_NUM_STAGES] = {};
    void mark(int stage, int which, hipStream_t s) {
        if (!timing) return;
        (void)hipEventRecord(ev[stage][which], s);
        if (which == 1) ev_recorded[stage] = true;
    }
    int nitems() const { return (int)items.size(); }
    const uint4* sorted_recs() const { return recs.as<uint4>(); }
    uint32_t* tile_queue() const { return reinterpret_cast<uint32_t*>(small.as<char>() + 384); }
    PixelState pixel_state() const {
        PixelState ps;
        ps.ray_color = ray_color.as<float4>();
        ps.frag_count = frag_count.as<uint32_t>();
        ps.last_rec = last_rec.as<uint32_t>();
        ps.num_ckpt = num_ckpt.as<uint32_t>();
        ps.ckpt_t = ckpt_t.as<float>();
        ps.ckpt_n = ckpt_n.as<uint32_t>();
        ps.counters = nullptr;
        ps.tile_frags = tile_frags.as<uint32_t>();
        return ps;
    }
    bool atomics_counted = false;  // the last adjoint ran with timing on and counted its atomics
};

namespace {

constexpr int kMaxBatchItems = 4096;

bool valid_camera(const dtv_camera& c) {
    return c.width > 0 && c.height > 0 && c.width <= 16384 && c.height <= 16384 && c.znear > 0.0f &&
           c.zfar > c.znear && c.fovy > 0.0f && c.fovy < 3.14159f;
}

// Uniforms of one camera; false for a singular view matrix.
bool make_uniforms(const dtv_camera& camera, const dtv_render_params& params, ViewUniforms& U) {
    double u[68];
    if (!camera_uniforms(camera, u)) return false;
    for (int i = 0; i < 16; i++) {
        U.viewProj[i] = (float)u[i];
        U.invProj[i] = (float)u[32 + i];
        U.invView[i] = (float)u[48 + i];
    }
    U.camPos[0] = (float)u[64];
    U.camPos[1] = (float)u[65];
    U.camPos[2] = (float)u[66];
    U.znear = camera.znear;
    U.width = camera.width;
    U.height = camera.height;
    U.tile_shift = params.tile_size == 16 ? 4 : params.tile_size == 32 ? 5 : 3;
    const int S = 1 << U.tile_shift;
    U.tiles_x = (camera.width + S - 1) / S;
    U.tiles_y = (camera.height + S - 1) / S;
    U.sub_x = (camera.width + 7) / 8;
    U.sub_y = (camera.height + 7) / 8;
    U.attenuation = params.attenuation;
    U.ert_alpha = params.early_ray_out_alpha;
    U.use_ert = params.use_early_ray_termination ? 1 : 0;
    for (int i = 0; i < 4; i++) U.clear[i] = params.clear_color[i];
    U.ckpt_interval = params.checkpoint_interval > 0 ? params.checkpoint_interval : 64;
    U.max_ckpt = kMaxCheckpoints;
    // the view axis: minus row 2 of the view matrix (column-major), i.e. the direction of increasing linear depth
    U.depth_axis[0] = -camera.view[2];
    U.depth_axis[1] = -camera.view[6];
    U.depth_axis[2] = -camera.view[10];
    U.linear_depth = params.segment_length == DTV_SEGMENT_LINEAR_DEPTH ? 1 : 0;
    return true;
}

// Binning, sort and forward render of a batch (one host synchronisation, after the record count).
int render_batch(dtv_view* view, const dtv_render_item* in, int n, const dtv_render_params* params, hipStream_t s,
                 const char* fn) {
    const std::string f(fn);
    if (!view || !in || !params) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": NULL argument");
    if (n <= 0 || n > kMaxBatchItems) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": 1..4096 items required");
    const int tile = params->tile_size == 0 ? 8 : params->tile_size;
    if (tile != 8 && tile != 16 && tile != 32)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": tile_size must be 8, 16 or 32");
    if (params->tile_order != 0 && params->tile_order != 1)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": tile_order must be 0 (heaviest first) or 1 (row-major)");
    if (params->segment_length != DTV_SEGMENT_EUCLIDEAN && params->segment_length != DTV_SEGMENT_LINEAR_DEPTH)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": segment_length must be 0 (Euclidean) or 1 (linear depth)");
    if (params->renderer != DTV_RENDERER_TILE && params->renderer != DTV_RENDERER_PROJECTION)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": renderer must be 0 (tile) or 1 (projection)");
    const int spp = params->segments_per_pass == 0 ? 128 : params->segments_per_pass;
    if (spp != 64 && spp != 128 && spp != 256)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": segments_per_pass must be 64, 128 or 256");
    if (params->tile_launch != DTV_TILE_LAUNCH_GRID && params->tile_launch != DTV_TILE_LAUNCH_QUEUE)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": tile_launch must be 0 (grid) or 1 (queue)");
    const bool proj = params->renderer == DTV_RENDERER_PROJECTION;
    if (proj && tile != 8) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": the Projection renderer bins 8x8 tiles");
    if (proj && params->tile_launch != DTV_TILE_LAUNCH_GRID)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": the Projection renderer launches a workgroup per tile");
    dtv_render_params prm = *params;
    prm.tile_size = tile;
    params = &prm;
    std::vector<BatchItem> items((size_t)n);
    std::vector<int64_t> item_V((size_t)n);
    int64_t blocks = 0, tiles = 0, subtiles = 0, pixels = 0;
    uint32_t kinds = 0;
    for (int i = 0; i < n; i++) {
        const dtv_render_item& it = in[i];
        const dtv_mesh& m = it.mesh;
        if (!it.out_image) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": NULL argument");
        if (!valid_camera(it.camera)) {
            if (it.camera.width <= 0 || it.camera.height <= 0 || it.camera.width > 16384 || it.camera.height > 16384)
                return fail(DTV_ERR_INVALID_ARGUMENT, f + ": viewport must be 1..16384 pixels per side");
            return fail(DTV_ERR_INVALID_ARGUMENT, f + ": invalid fovy / clip planes");
        }
        if (const int rc = check_mesh(m, f); rc != DTV_OK) return rc;
        if (!aligned16(it.out_image)) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": out_image must be 16-byte aligned");
        BatchItem& B = items[(size_t)i];
        if (!make_uniforms(it.camera, *params, B.U)) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": singular view matrix");
        if (proj) {  // no early ray termination, no segment lengths
            B.U.use_ert = 0;
            B.U.linear_depth = 0;
        }
        B.cells = m.cells;
        B.pos = m.positions;
        B.col = m.colors;
        B.cell_col = m.cell_colors;
        kinds |= colour_kind(m);
        B.image = it.out_image;
        B.T = m.num_cells;
        B.block0 = (uint32_t)blocks;
        B.tile0 = (uint32_t)tiles;
        B.sub0 = (uint32_t)subtiles;
        B.pixel0 = pixels;
        item_V[(size_t)i] = m.num_vertices;
        blocks += count_pairs_blocks(m.num_cells);
        tiles += (int64_t)B.U.tiles_x * B.U.tiles_y;
        subtiles += (int64_t)B.U.sub_x * B.U.sub_y;
        pixels += (int64_t)B.U.width * B.U.height;
        if (blocks >= (1ll << 31) / 256 || subtiles >= (1ll << 31) || pixels >= (1ll << 31))
            return fail(DTV_ERR_INVALID_ARGUMENT, f + ": batch too large");
    }
    DTV_HIP(hipSetDevice(view->ctx->device));
    view->rendered = false;
    // The items reach the device through pinned staging.  Under stream capture the upload becomes a graph node
    // that reads the staging at every replay; otherwise the previous render's upload must have left the staging
    // before it is overwritten (in planned mode nothing else waits for the device between renders).
    if (!view->items_copied) DTV_HIP(hipEventCreateWithFlags(&view->items_copied, hipEventDisableTiming));
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    DTV_HIP(hipStreamIsCapturing(s, &capture));
    const bool capturing = capture != hipStreamCaptureStatusNone;
    if (view->items_pending && !capturing) DTV_HIP(hipEventSynchronize(view->items_copied));
    view->items_pending = false;
    if (view->pinned_items_cap < n) {
        if (view->pinnedItems) DTV_HIP(hipHostFree(view->pinnedItems));
        view->pinnedItems = nullptr;
        view->pinned_items_cap = 0;
        DTV_HIP(hipHostMalloc((void**)&view->pinnedItems, sizeof(BatchItem) * (size_t)n, hipHostMallocDefault));
        view->pinned_items_cap = n;
    }
    std::memcpy(view->pinnedItems, items.data(), sizeof(BatchItem) * (size_t)n);
    DTV_HIP(view->dItems.ensure(sizeof(BatchItem) * (size_t)n));
    DTV_HIP(hipMemcpyAsync(view->dItems.p, view->pinnedItems, sizeof(BatchItem) * (size_t)n, hipMemcpyHostToDevice,
                           s));
    if (!capturing) {
        DTV_HIP(hipEventRecord(view->items_copied, s));
        view->items_pending = true;
    }
    const BatchItem* dI = view->dItems.as<BatchItem>();

    // 1. count (tile, face) records per tet, 2. scan
    const int64_t nslots = 256 * (blocks > 0 ? blocks : 1);
    DTV_HIP(view->counts.ensure(sizeof(uint32_t) * nslots));
    DTV_HIP(view->offsets.ensure(sizeof(uint32_t) * nslots));
    DTV_HIP(view->scan_status.ensure(sizeof(uint64_t) * (scan_partitions(nslots) + 1)));
    // stats 0, scan counter 64, total 128, overflow 192, atomics 256, sort route 320 (48 B), tile queue 384
    DTV_HIP(view->small.ensure(448));
    BinStats* dstats = view->small.as<BinStats>();
    uint32_t* scan_counter = reinterpret_cast<uint32_t*>(view->small.as<char>() + 64);
    uint64_t* scan_total = reinterpret_cast<uint64_t*>(view->small.as<char>() + 128);
    DTV_HIP(view->block_stats.ensure(sizeof(uint32_t) * 4 * (blocks + 1)));
    if (proj) {
        DTV_HIP(view->pt_info.ensure(sizeof(uint32_t) * nslots));
        DTV_HIP(view->pt_vert.ensure(sizeof(float4) * 5 * nslots));
        DTV_HIP(view->pt_col.ensure(sizeof(float4) * 5 * nslots));
    }
    view->mark(DTV_STAGE_COUNT_PAIRS, 0, s);
    if (proj)
        launch_pt_count(dI, n, blocks, view->pt_info.as<uint32_t>(), view->pt_vert.as<float4>(),
                        view->pt_col.as<float4>(), view->counts.as<uint32_t>(), view->block_stats.as<uint32_t>(), s);
    else
        launch_count_pairs(dI, n, blocks, view->counts.as<uint32_t>(), view->block_stats.as<uint32_t>(), s);
    view->mark(DTV_STAGE_COUNT_PAIRS, 1, s);
    view->mark(DTV_STAGE_SCAN, 0, s);
    scan_exclusive_u32(view->counts.as<uint32_t>(), view->offsets.as<uint32_t>(), 256 * blocks,
                       view->scan_status.as<uint64_t>(), scan_counter, scan_total, s);
    launch_reduce_bin_stats(view->block_stats.as<uint32_t>(), blocks, scan_total, dstats, s);
    view->mark(DTV_STAGE_SCAN, 1, s);
    DTV_HIP(hipGetLastError());
    // The record count: read back after one synchronisation, or, in planned mode, a fixed capacity with the
    // count left in device memory (scan_total) for the kernels to read.
    const bool planned = view->planned_cap > 0;
    BinStats st{};
    int64_t K;
    if (!planned) {
        DTV_HIP(hipMemcpyAsync(view->pinnedStats, dstats, sizeof(BinStats), hipMemcpyDeviceToHost, s));
        DTV_HIP(hipStreamSynchronize(s));
        st = *view->pinnedStats;
        if (st.total_pairs >= (1ull << 30))
            return fail(DTV_ERR_OUT_OF_MEMORY, f + ": more than 2^30 (tile, face) records");
        K = (int64_t)st.total_pairs;
        view->num_pairs = K;
        view->num_tris = st.num_tris;
    } else {
        K = view->planned_cap;
        view->num_pairs = -1;
        view->num_tris = -1;
    }
    const uint64_t* n_dev = planned ? scan_total : nullptr;
    uint32_t* overflow = reinterpret_cast<uint32_t*>(view->small.as<char>() + 192);
    fill_u32(overflow, 0u, 1, s);

    // 3. emit, 4. sort, 5. tile ranges
    DTV_HIP(view->keys.ensure(sizeof(uint64_t) * (K > 0 ? K : 1)));
    DTV_HIP(view->vals.ensure(sizeof(uint32_t) * (K > 0 ? K : 1)));
    DTV_HIP(view->keys2.ensure(sizeof(uint64_t) * (K > 0 ? K : 1)));
    DTV_HIP(view->vals2.ensure(sizeof(uint32_t) * (K > 0 ? K : 1)));
    DTV_HIP(view->recs.ensure(2 * sizeof(uint4) * (K > 0 ? K : 1)));
    view->mark(DTV_STAGE_EMIT, 0, s);
    if (proj)
        launch_pt_emit(dI, n, blocks, view->pt_info.as<uint32_t>(), view->pt_vert.as<float4>(),
                       view->offsets.as<uint32_t>(), view->counts.as<uint32_t>(), view->keys.as<uint64_t>(),
                       view->vals.as<uint32_t>(), K, overflow, s);
    else
        launch_emit_pairs(dI, n, blocks, view->offsets.as<uint32_t>(), view->counts.as<uint32_t>(),
                          view->keys.as<uint64_t>(), view->vals.as<uint32_t>(), K, overflow, s);
    view->mark(DTV_STAGE_EMIT, 1, s);
    uint32_t mask = 0;
    // planned: a pass per depth digit (which of them vary is not known on the host; constant ones exit at once)
    const uint32_t varying = planned ? 0xFFFFFFFFu : (st.key_and ^ st.key_or);
    for (int p = 0; p < 4; p++)
        if ((varying >> (8 * p)) & 255u) mask |= 1u << p;
    int tile_bits = 0;
    while ((1ll << tile_bits) < tiles) tile_bits++;
    for (int p = 0; 8 * p < tile_bits; p++) mask |= 1u << (4 + p);
    int npass = 0;
    for (int p = 0; p < 8; p++) npass += (mask >> p) & 1;
    DTV_HIP(view->hist.ensure(sizeof(uint32_t) * 8 * 256 + 64));
    DTV_HIP(view->lookback.ensure(sizeof(uint32_t) * 256 * (size_t)(npass > 0 ? npass : 1) *
                                  (size_t)(sort_partitions(K) + 1)));
    // planned: a depth digit all keys share costs one launch that exits at once (the sort routes on the device)
    SortRoute* route = reinterpret_cast<SortRoute*>(view->small.as<char>() + 320);
    view->mark(DTV_STAGE_SORT, 0, s);
    view->sort_passes = radix_sort_pairs(view->keys.as<uint64_t>(), view->vals.as<uint32_t>(),
                                         view->keys2.as<uint64_t>(), view->vals2.as<uint32_t>(), K, n_dev, mask,
                                         view->hist.as<uint32_t>(), view->lookback.as<uint32_t>(),
                                         view->hist.as<uint32_t>() + 8 * 256, route, s);
    view->mark(DTV_STAGE_SORT, 1, s);
    DTV_HIP(view->ranges.ensure(sizeof(uint2) * tiles));
    DTV_HIP(view->tile_frags.ensure(sizeof(uint32_t) * subtiles));
    DTV_HIP(view->tile_order.ensure(2 * sizeof(uint32_t) * subtiles));
    view->cost_ordered = params->tile_order == 0;
    // forward launch order: tiles by record count, heaviest first
    const uint32_t* fwd_order = view->cost_ordered ? view->tile_order.as<uint32_t>() : nullptr;
    view->mark(DTV_STAGE_RANGES, 0, s);
    launch_tile_ranges(route, K, n_dev, tiles, view->ranges.as<uint2>(), s);
    if (fwd_order) {
        if (tile == 8) {
            launch_tile_order(view->ranges.as<uint2>(), nullptr, tiles, view->tile_order.as<uint32_t>(), s);
        } else {  // sub-tiles by their binning tile's records (the cost goes where the adjoint's order will be)
            uint32_t* cost = view->tile_order.as<uint32_t>() + subtiles;
            launch_subtile_cost(dI, n, subtiles, view->ranges.as<uint2>(), cost, s);
            launch_tile_order(nullptr, cost, subtiles, view->tile_order.as<uint32_t>(), s);
        }
    }
    if (proj)
        launch_pt_gather(dI, n, view->pt_info.as<uint32_t>(), view->pt_vert.as<float4>(), route, K, n_dev,
                         view->recs.as<uint4>(), s);
    else
        launch_gather_records(dI, n, route, K, n_dev, view->recs.as<uint4>(), s);
    view->mark(DTV_STAGE_RANGES, 1, s);

    // 6. render
    DTV_HIP(view->ray_color.ensure(sizeof(float4) * pixels));
    DTV_HIP(view->frag_count.ensure(sizeof(uint32_t) * pixels));
    DTV_HIP(view->last_rec.ensure(sizeof(uint32_t) * pixels));
    DTV_HIP(view->num_ckpt.ensure(sizeof(uint32_t) * pixels));
    DTV_HIP(view->ckpt_t.ensure(sizeof(float) * pixels * kMaxCheckpoints));
    DTV_HIP(view->ckpt_n.ensure(sizeof(uint32_t) * pixels * kMaxCheckpoints));
    view->mark(DTV_STAGE_FORWARD, 0, s);
    if (proj)
        launch_pt_forward(dI, n, tiles, fwd_order, view->pt_info.as<uint32_t>(), view->pt_vert.as<float4>(),
                          view->pt_col.as<float4>(), view->sorted_recs(), view->ranges.as<uint2>(),
                          view->pixel_state(), s);
    else
        launch_render_forward(dI, n, subtiles, fwd_order, view->sorted_recs(), view->ranges.as<uint2>(),
                              view->pixel_state(), kinds, s, spp, tile != 8,
                              params->tile_launch == DTV_TILE_LAUNCH_QUEUE ? view->tile_queue() : nullptr);
    view->mark(DTV_STAGE_FORWARD, 1, s);
    DTV_HIP(hipGetLastError());
    view->colour_kinds = kinds;
    view->projection = proj;
    view->items.swap(items);
    view->item_V.swap(item_V);
    view->total_blocks = blocks;
    view->total_tiles = tiles;
    view->total_subtiles = subtiles;
    view->sub_tiles = tile != 8;
    view->queued = params->tile_launch == DTV_TILE_LAUNCH_QUEUE;
    view->total_pixels = pixels;
    view->rendered = true;
    return DTV_OK;
}

// Adjoint of the last render_batch; tg[i] belongs to its item i.
int adjoint_batch(dtv_view* view, const AdjointTarget* tg, int n, int32_t use_abs_grad, int32_t adjoint_mode,
                  hipStream_t s, const char* fn) {
    const std::string f(fn);
    if (adjoint_mode != DTV_ADJOINT_EXACT && adjoint_mode != DTV_ADJOINT_REFERENCE)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": unsupported adjoint_mode");
    if (n != view->nitems()) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": item count differs from the render");
    if (adjoint_mode == DTV_ADJOINT_REFERENCE && view->items[0].U.linear_depth)
        return fail(DTV_ERR_INVALID_ARGUMENT,
                    f + ": the reference adjoint mode restates the INTERSECTION renderer's adjoint and needs "
                        "Euclidean segment lengths");
    if (adjoint_mode == DTV_ADJOINT_REFERENCE && view->projection)
        return fail(DTV_ERR_INVALID_ARGUMENT,
                    f + ": the Projection renderer's adjoint is the exact derivative of its forward "
                        "(DTV_ADJOINT_EXACT)");
    if (adjoint_mode == DTV_ADJOINT_REFERENCE && (view->colour_kinds & kColoursPerCell))
        return fail(DTV_ERR_INVALID_ARGUMENT,
                    f + ": the reference adjoint mode restates the per-vertex colour adjoint; per-cell colours "
                        "take the exact adjoint");
    for (int i = 0; i < n; i++) {
        if (!tg[i].dimage) continue;  // item not differentiated
        if (!tg[i].grad_pos || !tg[i].grad_col) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": NULL argument");
        if (!aligned16(tg[i].dimage)) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": dimage must be 16-byte aligned");
    }
    DTV_HIP(hipSetDevice(view->ctx->device));
    if (!view->targets_copied) DTV_HIP(hipEventCreateWithFlags(&view->targets_copied, hipEventDisableTiming));
    // Under stream capture the upload becomes a graph node that reads the staging at every replay (nothing may
    // wait on the device then); otherwise the staging of the previous call must have been consumed first.
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    DTV_HIP(hipStreamIsCapturing(s, &capture));
    const bool capturing = capture != hipStreamCaptureStatusNone;
    if (view->targets_pending && !capturing) DTV_HIP(hipEventSynchronize(view->targets_copied));
    view->targets_pending = false;
    if (view->pinned_targets_cap < n) {
        if (view->pinnedTargets) DTV_HIP(hipHostFree(view->pinnedTargets));
        view->pinnedTargets = nullptr;
        view->pinned_targets_cap = 0;
        DTV_HIP(hipHostMalloc((void**)&view->pinnedTargets, sizeof(AdjointTarget) * (size_t)n, hipHostMallocDefault));
        view->pinned_targets_cap = n;
    }
    // Each distinct gradient target (the items of one mesh share theirs) accumulates its vertex gradients in
    // zeroed interleaved rows (8 floats per vertex), folded into its gradients after the adjoint.
    struct Fold {
        float* acc;
        int64_t V;
        float* grad_pos;
        float* grad_col;
    };
    std::vector<Fold> folds;
    std::memcpy(view->pinnedTargets, tg, sizeof(AdjointTarget) * (size_t)n);
    {
        std::unordered_map<const float*, int> first;  // grad_pos -> first item
        std::vector<int64_t> row0((size_t)n, -1);
        int64_t rows = 0;
        for (int i = 0; i < n; i++) {
            if (!tg[i].dimage) continue;
            const bool cell = view->items[(size_t)i].cell_col != nullptr;
            auto it = first.find(tg[i].grad_pos);
            if (it != first.end()) {
                const int k = it->second;
                if (view->item_V[(size_t)k] != view->item_V[(size_t)i] || tg[k].grad_col != tg[i].grad_col ||
                    (view->items[(size_t)k].cell_col != nullptr) != cell)
                    return fail(DTV_ERR_INVALID_ARGUMENT,
                                f + ": items that share a gradient buffer must render the same mesh");
                row0[(size_t)i] = row0[(size_t)k];
                continue;
            }
            first.emplace(tg[i].grad_pos, i);
            row0[(size_t)i] = rows;
            folds.push_back({nullptr, view->item_V[(size_t)i], tg[i].grad_pos, cell ? nullptr : tg[i].grad_col});
            rows += view->item_V[(size_t)i];
        }
        DTV_HIP(view->vacc.ensure(sizeof(float) * 8 * (size_t)(rows > 0 ? rows : 1)));
        float* base = view->vacc.as<float>();
        int64_t r = 0;
        for (Fold& fo : folds) {
            fo.acc = base + 8 * r;
            r += fo.V;
        }
        for (int i = 0; i < n; i++)
            view->pinnedTargets[i].acc = row0[(size_t)i] >= 0 ? base + 8 * row0[(size_t)i] : nullptr;
        if (rows > 0) fill_u32(reinterpret_cast<uint32_t*>(base), 0u, 8 * rows, s);
    }
    DTV_HIP(view->dTargets.ensure(sizeof(AdjointTarget) * (size_t)n));
    DTV_HIP(hipMemcpyAsync(view->dTargets.p, view->pinnedTargets, sizeof(AdjointTarget) * (size_t)n,
                           hipMemcpyHostToDevice, s));
    if (!capturing) {
        DTV_HIP(hipEventRecord(view->targets_copied, s));
        view->targets_pending = true;
    }
    PixelState ps = view->pixel_state();
    view->atomics_counted = view->timing;
    if (view->timing) {  // count the gradient atomics along with the stage timers
        ps.counters = reinterpret_cast<unsigned long long*>(view->small.as<char>() + 256);
        fill_u32(reinterpret_cast<uint32_t*>(ps.counters), 0u, 2, s);
    }
    // adjoint launch order: tiles by the forward's composited fragments, heaviest first
    const uint32_t* adj_order = nullptr;
    if (view->cost_ordered) {
        uint32_t* o = view->tile_order.as<uint32_t>() + view->total_subtiles;
        launch_tile_order(nullptr, view->tile_frags.as<uint32_t>(), view->total_subtiles, o, s);
        adj_order = o;
    }
    view->mark(DTV_STAGE_ADJOINT, 0, s);
    if (view->projection) {
        const int64_t tg_floats = 21 * 4 * 256 * view->total_blocks;
        DTV_HIP(view->tri_grad.ensure(sizeof(float) * (size_t)(tg_floats > 0 ? tg_floats : 1)));
        launch_pt_adjoint(view->dItems.as<BatchItem>(), view->dTargets.as<AdjointTarget>(), n, view->total_tiles,
                          view->total_blocks, adj_order, view->pt_info.as<uint32_t>(), view->pt_vert.as<float4>(),
                          view->pt_col.as<float4>(), view->sorted_recs(), view->ranges.as<uint2>(), ps,
                          view->tri_grad.as<float>(), tg_floats, use_abs_grad ? 1 : 0, s);
    } else {
        launch_render_adjoint(view->dItems.as<BatchItem>(), view->dTargets.as<AdjointTarget>(), n,
                              view->total_subtiles, adj_order, view->sorted_recs(), view->ranges.as<uint2>(), ps,
                              use_abs_grad ? 1 : 0, adjoint_mode, view->colour_kinds, s, view->sub_tiles,
                              view->queued ? view->tile_queue() : nullptr);
    }
    for (const Fold& fo : folds) launch_fold_vertex_grads(fo.acc, fo.V, fo.grad_pos, fo.grad_col, s);
    view->mark(DTV_STAGE_ADJOINT, 1, s);
    DTV_HIP(hipGetLastError());
    return DTV_OK;
}

// The adjoint reads the cells, positions and colours the render recorded and scatters into gradients sized by
// the vertex count: the caller's mesh must be that one, unchanged in its buffers and sizes.
bool same_mesh(const dtv_view* view, int i, const dtv_mesh& m) {
    const BatchItem& B = view->items[(size_t)i];
    return m.cells == B.cells && m.positions == B.pos && m.colors == B.col && m.cell_colors == B.cell_col &&
           m.num_cells == B.T && m.num_vertices == view->item_V[(size_t)i];
}

}  // namespace

extern "C" {

int dtv_get_version(void) { return 1; }

const char* dtv_last_error(void) { return g_last_error.c_str(); }

int dtv_ctx_create(int32_t device, dtv_ctx** out_ctx) {
    if (!out_ctx) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_ctx_create: out_ctx is NULL");
    int n = 0;
    DTV_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_ctx_create: invalid device index");
    DTV_HIP(hipSetDevice(device));
    const int rc = setup_noalloc_scratch(device);
    if (rc != DTV_OK) return rc;
    dtv_ctx* c = new dtv_ctx();
    c->device = device;
    *out_ctx = c;
    return DTV_OK;
}

int dtv_ctx_destroy(dtv_ctx* ctx) {
    delete ctx;
    return DTV_OK;
}

int dtv_ctx_get_device(const dtv_ctx* ctx, int32_t* out_device) {
    if (!ctx || !out_device) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_ctx_get_device: NULL argument");
    *out_device = ctx->device;
    return DTV_OK;
}

int dtv_view_create(dtv_ctx* ctx, dtv_view** out_view) {
    if (!ctx || !out_view) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_view_create: NULL argument");
    DTV_HIP(hipSetDevice(ctx->device));
    dtv_view* v = new dtv_view();
    v->ctx = ctx;
    hipError_t e = hipHostMalloc((void**)&v->pinnedStats, sizeof(BinStats) + 16, hipHostMallocDefault);
    if (e != hipSuccess) {
        delete v;
        return fail(DTV_ERR_HIP, std::string("dtv_view_create: hipHostMalloc: ") + hipGetErrorString(e));
    }
    *out_view = v;
    return DTV_OK;
}

int dtv_view_destroy(dtv_view* v) {
    if (!v) return DTV_OK;
    (void)hipSetDevice(v->ctx->device);
    (void)hipDeviceSynchronize();
    DevBuf* bufs[] = {&v->dItems, &v->dTargets, &v->counts, &v->offsets, &v->scan_status, &v->small, &v->block_stats, &v->keys,
                      &v->vals, &v->keys2,
                      &v->vals2, &v->recs, &v->hist, &v->lookback, &v->ranges, &v->ray_color, &v->frag_count, &v->last_rec,
                      &v->num_ckpt, &v->ckpt_t, &v->ckpt_n, &v->tile_frags, &v->tile_order, &v->pt_info,
                      &v->pt_vert, &v->pt_col, &v->tri_grad, &v->vacc};
    for (DevBuf* b : bufs) b->release();
    for (int i = 0; i < DTV_NUM_STAGES; i++)
        for (int k = 0; k < 2; k++)
            if (v->ev[i][k]) (void)hipEventDestroy(v->ev[i][k]);
    if (v->targets_copied) (void)hipEventDestroy(v->targets_copied);
    if (v->items_copied) (void)hipEventDestroy(v->items_copied);
    if (v->pinnedItems) (void)hipHostFree(v->pinnedItems);
    if (v->pinnedTargets) (void)hipHostFree(v->pinnedTargets);
    if (v->pinnedStats) (void)hipHostFree(v->pinnedStats);
    delete v;
    return DTV_OK;
}

int dtv_camera_uniforms(const dtv_camera* camera, float* out68) {
    if (!camera || !out68) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_camera_uniforms: NULL argument");
    double u[68];
    if (!camera_uniforms(*camera, u)) return fail(DTV_ERR_INVALID_ARGUMENT, "singular view matrix");
    for (int i = 0; i < 68; i++) out68[i] = (float)u[i];
    return DTV_OK;
}

int dtv_render(dtv_view* view, const dtv_mesh* mesh, const dtv_camera* camera, const dtv_render_params* params,
               float* out_image, void* stream) {
    if (!view || !mesh || !camera || !params || !out_image)
        return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_render: NULL argument");
    dtv_render_item it;
    it.mesh = *mesh;
    it.camera = *camera;
    it.out_image = out_image;
    return render_batch(view, &it, 1, params, reinterpret_cast<hipStream_t>(stream), "dtv_render");
}

int dtv_render_noalloc(dtv_ctx* ctx, const dtv_mesh* mesh, const dtv_camera* camera,
                       const dtv_render_params* params, float* out_image, int32_t* out_overflow, void* stream) {
    const std::string f = "dtv_render_noalloc";
    if (!ctx || !mesh || !camera || !params || !out_image) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": NULL argument");
    const int tile = params->tile_size == 0 ? 8 : params->tile_size;
    if (tile != 8) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": tile_size must be 8");
    if (params->tile_order != 0 && params->tile_order != 1)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": tile_order must be 0 (heaviest first) or 1 (row-major)");
    if (params->segment_length != DTV_SEGMENT_EUCLIDEAN && params->segment_length != DTV_SEGMENT_LINEAR_DEPTH)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": segment_length must be 0 (Euclidean) or 1 (linear depth)");
    if (params->renderer != DTV_RENDERER_TILE)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": renders with the tile renderer only (renderer 0)");
    if (params->tile_launch != DTV_TILE_LAUNCH_GRID)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": launches a workgroup per tile only (tile_launch 0)");
    if (!valid_camera(*camera)) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": invalid viewport, fovy or clip planes");
    const dtv_mesh& m = *mesh;
    if (const int rc = check_mesh(m, f); rc != DTV_OK) return rc;
    if (!aligned16(out_image)) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": out_image must be 16-byte aligned");
    const int64_t P = (int64_t)camera->width * camera->height;
    const int64_t tiles = (int64_t)((camera->width + 7) / 8) * ((camera->height + 7) / 8);
    if (m.num_cells > kNaMaxTets || P > kNaMaxPixels || tiles > kNaMaxTiles)
        return fail(DTV_ERR_INVALID_ARGUMENT,
                    f + ": beyond the allocation-free limits (262,144 tets, 262,144 pixels, 8,192 8x8 tiles); "
                        "use dtv_render");
    BatchItem B{};
    // forward only: no checkpoint slots (the forward stores none; the render cannot be differentiated)
    dtv_render_params prm = *params;
    if (!make_uniforms(*camera, prm, B.U)) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": singular view matrix");
    B.U.max_ckpt = 0;
    B.cells = m.cells;
    B.pos = m.positions;
    B.col = m.colors;
    B.cell_col = m.cell_colors;
    B.image = out_image;
    B.T = m.num_cells;
    B.block0 = 0;
    B.tile0 = 0;
    B.sub0 = 0;
    B.pixel0 = 0;
    const int64_t blocks = count_pairs_blocks(m.num_cells);
    const int64_t nslots = 256 * (blocks > 0 ? blocks : 1);
    const int64_t K = kNaMaxRecords;
    int tile_bits = 0;
    while ((1ll << tile_bits) < tiles) tile_bits++;
    uint32_t mask = 0xFu;  // every depth digit: which of them vary is known only on the device
    for (int p = 0; 8 * p < tile_bits; p++) mask |= 1u << (4 + p);
    int npass = 0;
    for (int p = 0; p < 8; p++) npass += (mask >> p) & 1;
    // carve the scratch (256-byte aligned pieces)
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_item = take(sizeof(BatchItem)), o_small = take(384), o_counts = take(4 * (size_t)nslots),
                 o_offsets = take(4 * (size_t)nslots), o_status = take(8 * (size_t)(scan_partitions(nslots) + 1)),
                 o_bstats = take(16 * (size_t)(blocks + 1)), o_keys = take(8 * (size_t)K), o_keys2 = take(8 * (size_t)K),
                 o_vals = take(4 * (size_t)K), o_vals2 = take(4 * (size_t)K), o_recs = take(32 * (size_t)K),
                 o_hist = take(4 * 8 * 256 + 64), o_lookback = take(4 * 256 * (size_t)npass * (sort_partitions(K) + 1)),
                 o_ranges = take(8 * (size_t)tiles), o_tfrags = take(4 * (size_t)tiles), o_order = take(4 * (size_t)tiles),
                 o_color = take(16 * (size_t)P), o_count = take(4 * (size_t)P), o_last = take(4 * (size_t)P),
                 o_nck = take(4 * (size_t)P);
    if (off > kNaScratchBytes) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": scratch layout exceeds 56 MiB");
    DTV_HIP(hipSetDevice(ctx->device));
    char* base = static_cast<char*>(g_scratch_ptr[ctx->device]);
    if (!base) return fail(DTV_ERR_HIP, f + ": scratch not set up on this device");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    BatchItem* dI = reinterpret_cast<BatchItem*>(base + o_item);
    char* small = base + o_small;
    uint32_t* scan_counter = reinterpret_cast<uint32_t*>(small + 64);
    uint64_t* scan_total = reinterpret_cast<uint64_t*>(small + 128);
    uint32_t* overflow = reinterpret_cast<uint32_t*>(small + 192);
    SortRoute* route = reinterpret_cast<SortRoute*>(small + 320);
    uint32_t* counts = reinterpret_cast<uint32_t*>(base + o_counts);
    uint32_t* hist = reinterpret_cast<uint32_t*>(base + o_hist);
    uint2* ranges = reinterpret_cast<uint2*>(base + o_ranges);
    uint32_t* order = params->tile_order == 0 ? reinterpret_cast<uint32_t*>(base + o_order) : nullptr;
    uint4* recs = reinterpret_cast<uint4*>(base + o_recs);
    // the item reaches the device as a kernel argument: no host staging
    hipLaunchKernelGGL(k_store_item, dim3(1), dim3(1), 0, s, B, dI);
    launch_count_pairs(dI, 1, blocks, counts, reinterpret_cast<uint32_t*>(base + o_bstats), s);
    scan_exclusive_u32(counts, reinterpret_cast<uint32_t*>(base + o_offsets), 256 * blocks,
                       reinterpret_cast<uint64_t*>(base + o_status), scan_counter, scan_total, s);
    fill_u32(overflow, 0u, 1, s);
    launch_emit_pairs(dI, 1, blocks, reinterpret_cast<uint32_t*>(base + o_offsets), counts,
                      reinterpret_cast<uint64_t*>(base + o_keys), reinterpret_cast<uint32_t*>(base + o_vals), K,
                      overflow, s);
    radix_sort_pairs(reinterpret_cast<uint64_t*>(base + o_keys), reinterpret_cast<uint32_t*>(base + o_vals),
                     reinterpret_cast<uint64_t*>(base + o_keys2), reinterpret_cast<uint32_t*>(base + o_vals2), K,
                     scan_total, mask, hist, reinterpret_cast<uint32_t*>(base + o_lookback), hist + 8 * 256, route, s);
    launch_tile_ranges(route, K, scan_total, tiles, ranges, s);
    if (order) launch_tile_order(ranges, nullptr, tiles, order, s);
    launch_gather_records(dI, 1, route, K, scan_total, recs, s);
    PixelState ps;
    ps.ray_color = reinterpret_cast<float4*>(base + o_color);
    ps.frag_count = reinterpret_cast<uint32_t*>(base + o_count);
    ps.last_rec = reinterpret_cast<uint32_t*>(base + o_last);
    ps.num_ckpt = reinterpret_cast<uint32_t*>(base + o_nck);
    ps.ckpt_t = nullptr;  // max_ckpt = 0: never written
    ps.ckpt_n = nullptr;
    ps.counters = nullptr;
    ps.tile_frags = reinterpret_cast<uint32_t*>(base + o_tfrags);
    launch_render_forward(dI, 1, tiles, order, recs, ranges, ps, colour_kind(m), s);
    if (out_overflow) hipLaunchKernelGGL(k_copy_flag, dim3(1), dim3(1), 0, s, overflow, out_overflow);
    DTV_HIP(hipGetLastError());
    return DTV_OK;
}

int dtv_render_batch(dtv_view* view, const dtv_render_item* items, int32_t num_items,
                     const dtv_render_params* params, void* stream) {
    return render_batch(view, items, num_items, params, reinterpret_cast<hipStream_t>(stream), "dtv_render_batch");
}

int dtv_autotune(dtv_view* view, const dtv_render_item* items, int32_t num_items, const dtv_render_params* params,
                 int32_t with_adjoint, int32_t repeats, dtv_tune_result* out, void* stream) {
    const std::string f = "dtv_autotune";
    if (!view || !items || !params || !out) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": NULL argument");
    if (num_items <= 0 || num_items > kMaxBatchItems || repeats < 1 || repeats > 100)
        return fail(DTV_ERR_INVALID_ARGUMENT, f + ": 1..4096 items and 1..100 repeats required");
    if (params->renderer != DTV_RENDERER_TILE) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": tile renderer only");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    DTV_HIP(hipSetDevice(view->ctx->device));
    // scratch adjoint inputs: a constant image adjoint per item, gradients per distinct mesh
    std::vector<dtv_adjoint_item> adj((size_t)num_items);
    std::vector<void*> scratch;
    auto release = [&]() {
        for (void* p : scratch) (void)hipFree(p);
        scratch.clear();
    };
    if (with_adjoint) {
        for (int i = 0; i < num_items; i++) {
            const dtv_render_item& it = items[i];
            const int64_t P = (int64_t)it.camera.width * it.camera.height;
            const int64_t ncol = it.mesh.cell_colors ? it.mesh.num_cells : it.mesh.num_vertices;
            void *dimg = nullptr, *gp = nullptr, *gc = nullptr;
            if (hipMalloc(&dimg, sizeof(float) * 4 * (size_t)(P > 0 ? P : 1)) != hipSuccess ||
                (scratch.push_back(dimg), hipMalloc(&gp, sizeof(float) * 3 * (size_t)(it.mesh.num_vertices + 1))) !=
                    hipSuccess ||
                (scratch.push_back(gp), hipMalloc(&gc, sizeof(float) * 4 * (size_t)(ncol + 1))) != hipSuccess) {
                release();
                return fail(DTV_ERR_OUT_OF_MEMORY, f + ": scratch gradients");
            }
            scratch.push_back(gc);
            fill_u32(reinterpret_cast<uint32_t*>(dimg), 0x3A83126Fu /* 1e-3f */, 4 * P, s);
            fill_u32(reinterpret_cast<uint32_t*>(gp), 0u, 3 * it.mesh.num_vertices, s);
            fill_u32(reinterpret_cast<uint32_t*>(gc), 0u, 4 * ncol, s);
            adj[(size_t)i].mesh = it.mesh;
            adj[(size_t)i].dimage = static_cast<const float*>(dimg);
            adj[(size_t)i].grad_positions = static_cast<float*>(gp);
            adj[(size_t)i].grad_colors = static_cast<float*>(gc);
        }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        release();
        return fail(DTV_ERR_HIP, f + ": hipEventCreate failed");
    }
    const int tiles[3] = {8, 16, 32}, spps[3] = {64, 128, 256};
    int nc = 0, rc = DTV_OK;
    float best = 0.f;
    out->ms_default = -1.f;
    for (int ti = 0; ti < 3 && rc == DTV_OK; ti++)
        for (int si = 0; si < 3 && rc == DTV_OK; si++) {
            dtv_render_params prm = *params;
            prm.tile_size = tiles[ti];
            prm.segments_per_pass = spps[si];
            float total = 0.f;
            for (int r = 0; r <= repeats && rc == DTV_OK; r++) {
                if (r == 1) (void)hipEventRecord(e0, s);
                rc = render_batch(view, items, num_items, &prm, s, "dtv_autotune");
                if (rc == DTV_OK && with_adjoint)
                    rc = dtv_render_adjoint_batch(view, adj.data(), num_items, 0, DTV_ADJOINT_EXACT, stream);
            }
            if (rc != DTV_OK) break;
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&total, e0, e1) != hipSuccess) {
                rc = fail(DTV_ERR_HIP, f + ": timing failed");
                break;
            }
            const float ms = total / (float)repeats;
            out->cand_tile_size[nc] = tiles[ti];
            out->cand_segments_per_pass[nc] = spps[si];
            out->cand_ms[nc] = ms;
            if (tiles[ti] == 8 && spps[si] == 128) out->ms_default = ms;
            if (nc == 0 || ms < best) {
                best = ms;
                out->tile_size = tiles[ti];
                out->segments_per_pass = spps[si];
                out->ms = ms;
            }
            nc++;
        }
    out->num_candidates = nc;
    (void)hipStreamSynchronize(s);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    release();
    return rc;
}

int dtv_render_adjoint(dtv_view* view, const dtv_mesh* mesh, const float* dimage, int32_t use_abs_grad,
                       int32_t adjoint_mode, float* grad_positions, float* grad_colors, void* stream) {
    if (!view || !mesh || !dimage || !grad_positions || !grad_colors)
        return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_render_adjoint: NULL argument");
    if (!view->rendered) return fail(DTV_ERR_NOT_RENDERED, "dtv_render_adjoint: view has not been rendered");
    if (view->nitems() != 1 || !same_mesh(view, 0, *mesh))
        return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_render_adjoint: mesh differs from the rendered one");
    AdjointTarget t{};
    t.dimage = dimage;
    t.grad_pos = grad_positions;
    t.grad_col = grad_colors;
    return adjoint_batch(view, &t, 1, use_abs_grad, adjoint_mode, reinterpret_cast<hipStream_t>(stream),
                         "dtv_render_adjoint");
}

int dtv_render_adjoint_batch(dtv_view* view, const dtv_adjoint_item* items, int32_t num_items, int32_t use_abs_grad,
                             int32_t adjoint_mode, void* stream) {
    if (!view || !items) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_render_adjoint_batch: NULL argument");
    if (!view->rendered) return fail(DTV_ERR_NOT_RENDERED, "dtv_render_adjoint_batch: view has not been rendered");
    if (num_items <= 0 || num_items > kMaxBatchItems)
        return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_render_adjoint_batch: item count differs from the render");
    if (num_items != view->nitems())
        return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_render_adjoint_batch: item count differs from the render");
    std::vector<AdjointTarget> tg((size_t)num_items);
    for (int i = 0; i < num_items; i++) {
        if (items[i].dimage && !same_mesh(view, i, items[i].mesh))
            return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_render_adjoint_batch: item " + std::to_string(i) +
                                                      ": mesh differs from the rendered one");
        tg[(size_t)i].dimage = items[i].dimage;
        tg[(size_t)i].grad_pos = items[i].grad_positions;
        tg[(size_t)i].grad_col = items[i].grad_colors;
    }
    return adjoint_batch(view, tg.data(), num_items, use_abs_grad, adjoint_mode, reinterpret_cast<hipStream_t>(stream),
                         "dtv_render_adjoint_batch");
}

int dtv_view_fragment_counts(dtv_view* view, int32_t* out_counts, void* stream) {
    if (!view || !out_counts) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_view_fragment_counts: NULL argument");
    if (!view->rendered) return fail(DTV_ERR_NOT_RENDERED, "dtv_view_fragment_counts: view has not been rendered");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    DTV_HIP(hipSetDevice(view->ctx->device));
    DTV_HIP(hipMemcpyAsync(out_counts, view->frag_count.p, sizeof(uint32_t) * view->total_pixels,
                           hipMemcpyDeviceToDevice, s));
    return DTV_OK;
}

int dtv_view_export_fragments(dtv_view* view, const dtv_mesh* mesh, int32_t max_per_pixel, int32_t* out_ids,
                              void* stream) {
    if (!view || !mesh || !out_ids || max_per_pixel <= 0)
        return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_view_export_fragments: invalid argument");
    if (!view->rendered) return fail(DTV_ERR_NOT_RENDERED, "dtv_view_export_fragments: view has not been rendered");
    if (view->projection)
        return fail(DTV_ERR_INVALID_ARGUMENT,
                    "dtv_view_export_fragments: the Projection renderer's fragments are triangles, not tet faces");
    if (view->nitems() != 1 || mesh->cells != view->items[0].cells || mesh->num_cells != view->items[0].T)
        return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_view_export_fragments: mesh differs from the rendered one");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    DTV_HIP(hipSetDevice(view->ctx->device));
    launch_export_fragments(view->dItems.as<BatchItem>(), view->nitems(), view->total_subtiles, view->sorted_recs(),
                            view->ranges.as<uint2>(), view->pixel_state(), max_per_pixel, out_ids, s);
    DTV_HIP(hipGetLastError());
    return DTV_OK;
}

int dtv_view_set_record_capacity(dtv_view* view, int64_t record_capacity) {
    if (!view || record_capacity < 0 || record_capacity >= (int64_t)(1u << 30))
        return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_view_set_record_capacity: capacity must be 0..2^30-1");
    view->planned_cap = record_capacity;
    return DTV_OK;
}

int dtv_view_records_overflowed(dtv_view* view, int32_t* out_overflowed, int64_t* out_records) {
    if (!view || !out_overflowed) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_view_records_overflowed: NULL argument");
    if (!view->rendered) return fail(DTV_ERR_NOT_RENDERED, "dtv_view_records_overflowed: view has not been rendered");
    DTV_HIP(hipSetDevice(view->ctx->device));
    DTV_HIP(hipDeviceSynchronize());
    uint64_t total = 0;
    DTV_HIP(hipMemcpy(&total, view->small.as<char>() + 128, sizeof(total), hipMemcpyDeviceToHost));
    BinStats st{};
    DTV_HIP(hipMemcpy(&st, view->small.p, sizeof(st), hipMemcpyDeviceToHost));
    view->num_pairs = (int64_t)total;
    view->num_tris = st.num_tris;
    const int64_t cap = view->planned_cap > 0 ? view->planned_cap : (int64_t)total;
    *out_overflowed = (int64_t)total > cap ? 1 : 0;
    if (out_records) *out_records = (int64_t)total;
    return DTV_OK;
}

int dtv_view_get_stats(dtv_view* view, dtv_view_stats* out) {
    if (!view || !out) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_view_get_stats: NULL argument");
    if (view->rendered && view->num_pairs < 0) {  // planned render: the counts are still on the device
        int32_t o = 0;
        const int rc = dtv_view_records_overflowed(view, &o, nullptr);
        if (rc != DTV_OK) return rc;
    }
    out->num_triangles = view->num_tris;
    out->num_pairs = view->num_pairs;
    out->num_tiles = view->total_tiles;
    out->sort_passes = view->sort_passes;
    out->width = view->rendered ? view->items[0].U.width : 0;
    out->height = view->rendered ? view->items[0].U.height : 0;
    out->num_items = view->rendered ? view->nitems() : 0;
    out->num_pixels = view->total_pixels;
    out->num_gradient_atomics = -1;
    if (view->atomics_counted) {
        DTV_HIP(hipSetDevice(view->ctx->device));
        DTV_HIP(hipDeviceSynchronize());
        unsigned long long c = 0;
        DTV_HIP(hipMemcpy(&c, view->small.as<char>() + 256, sizeof(c), hipMemcpyDeviceToHost));
        out->num_gradient_atomics = (int64_t)c;
    }
    return DTV_OK;
}

int dtv_view_set_timing(dtv_view* view, int32_t enabled) {
    if (!view) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_view_set_timing: NULL view");
    DTV_HIP(hipSetDevice(view->ctx->device));
    if (enabled && !view->ev[0][0]) {
        for (int i = 0; i < DTV_NUM_STAGES; i++)
            for (int k = 0; k < 2; k++) DTV_HIP(hipEventCreate(&view->ev[i][k]));
    }
    view->timing = enabled != 0;
    for (int i = 0; i < DTV_NUM_STAGES; i++) view->ev_recorded[i] = false;
    return DTV_OK;
}

int dtv_view_get_timings(dtv_view* view, float* out_ms) {
    if (!view || !out_ms) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_view_get_timings: NULL argument");
    DTV_HIP(hipSetDevice(view->ctx->device));
    for (int i = 0; i < DTV_NUM_STAGES; i++) {
        out_ms[i] = -1.0f;
        if (!view->ev_recorded[i]) continue;
        DTV_HIP(hipEventSynchronize(view->ev[i][1]));
        float ms = 0.f;
        DTV_HIP(hipEventElapsedTime(&ms, view->ev[i][0], view->ev[i][1]));
        out_ms[i] = ms;
    }
    return DTV_OK;
}

namespace {

int adam(const dtv_adam_group* groups, int32_t num_groups, int64_t* step_dev, hipStream_t s, const char* fn) {
    const std::string f(fn);
    if (!groups || num_groups <= 0 || num_groups > 8) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": 1..8 groups required");
    AdamGroupDev g[8];
    for (int i = 0; i < num_groups; i++) {
        const dtv_adam_group& G = groups[i];
        if (G.numel < 0 || (G.numel > 0 && (!G.param || !G.grad || !G.exp_avg || !G.exp_avg_sq)))
            return fail(DTV_ERR_INVALID_ARGUMENT, f + ": invalid group");
        if (!aligned16(G.param) || !aligned16(G.grad) || !aligned16(G.exp_avg) || !aligned16(G.exp_avg_sq))
            return fail(DTV_ERR_INVALID_ARGUMENT, f + ": buffers must be 16-byte aligned");
        const int64_t step = step_dev ? 1 : G.step;  // with a device step count the host value is unused
        if (step < 1) return fail(DTV_ERR_INVALID_ARGUMENT, f + ": step must be >= 1");
        // torch.optim.Adam: bias corrections in double from the double hyper-parameters, the update's constants
        // (1 - beta1, beta2, 1 - beta2, lr / bc1, sqrt(bc2), eps) rounded once to float
        const double bc1 = 1.0 - std::pow(G.beta1, (double)step);
        const double bc2 = 1.0 - std::pow(G.beta2, (double)step);
        g[i].lr = G.lr;
        g[i].beta1 = G.beta1;
        g[i].beta2d = G.beta2;
        g[i].param = G.param;
        g[i].grad = G.grad;
        g[i].m = G.exp_avg;
        g[i].v = G.exp_avg_sq;
        g[i].n = G.numel;
        g[i].step_size = (float)(G.lr / bc1);
        g[i].bc2_sqrt = (float)std::sqrt(bc2);
        g[i].w1 = (float)(1.0 - G.beta1);
        g[i].beta2 = (float)G.beta2;
        g[i].w2 = (float)(1.0 - G.beta2);
        g[i].eps = (float)G.eps;
        g[i].clamp_nonneg = G.clamp_nonneg;
        g[i].zero_grad = G.zero_grad;
    }
    launch_adam(g, num_groups, step_dev, s);
    DTV_HIP(hipGetLastError());
    return DTV_OK;
}

}  // namespace

int dtv_adam_step(const dtv_adam_group* groups, int32_t num_groups, void* stream) {
    return adam(groups, num_groups, nullptr, reinterpret_cast<hipStream_t>(stream), "dtv_adam_step");
}

int dtv_adam_step_device(const dtv_adam_group* groups, int32_t num_groups, int64_t* step_state, void* stream) {
    if (!step_state) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_adam_step_device: NULL step_state");
    return adam(groups, num_groups, step_state, reinterpret_cast<hipStream_t>(stream), "dtv_adam_step_device");
}

int dtv_tet_regularizer_grad(const dtv_mesh* mesh, float lambda, float beta, float* grad_positions, void* stream) {
    if (!mesh || !grad_positions) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_tet_regularizer_grad: NULL argument");
    if (mesh->num_cells < 0 || (mesh->num_cells > 0 && (!mesh->cells || !mesh->positions)))
        return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_tet_regularizer_grad: invalid mesh");
    if (!aligned16(mesh->cells)) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_tet_regularizer_grad: cells must be 16-byte aligned");
    launch_tet_regularizer(mesh->cells, mesh->positions, mesh->num_cells, lambda, beta, grad_positions,
                           reinterpret_cast<hipStream_t>(stream));
    DTV_HIP(hipGetLastError());
    return DTV_OK;
}

int dtv_mse_loss_grad(const float* image, const float* target, int64_t n, float* dimage, float* loss_out,
                      void* stream) {
    if (!image || !target || !dimage || n <= 0) return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_mse_loss_grad: invalid argument");
    launch_mse(image, target, n, n, dimage, loss_out, reinterpret_cast<hipStream_t>(stream));
    DTV_HIP(hipGetLastError());
    return DTV_OK;
}

int dtv_mse_loss_grad_batch(const float* image, const float* target, int64_t n_per_item, int32_t num_items,
                            float* dimage, float* loss_out, void* stream) {
    if (!image || !target || !dimage || n_per_item <= 0 || num_items <= 0)
        return fail(DTV_ERR_INVALID_ARGUMENT, "dtv_mse_loss_grad_batch: invalid argument");
    launch_mse(image, target, n_per_item * num_items, n_per_item, dimage, loss_out,
               reinterpret_cast<hipStream_t>(stream));
    DTV_HIP(hipGetLastError());
    return DTV_OK;
}

}  // extern "C"
