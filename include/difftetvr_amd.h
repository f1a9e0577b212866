/*
 * difftetvr_amd: C ABI of the MI355X (gfx950) differentiable tetrahedral volume renderer.
 *
 * Plain pointers and sizes only.  Every device pointer is caller-owned and borrowed for the duration of the call;
 * every call is asynchronous on the given HIP stream (hipStream_t passed as void*; NULL = default stream) unless
 * documented otherwise.  Errors return a negative status; dtv_last_error() returns the message of the calling
 * thread's last failure.  No C++ exceptions cross this boundary.
 *
 * Each entry point cites the reference interface it replaces (paths relative to the DiffTetVR source tree).
 */
#ifndef DIFFTETVR_AMD_H
#define DIFFTETVR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DTV_OK 0
#define DTV_ERR_INVALID_ARGUMENT (-1)
#define DTV_ERR_HIP (-2)
#define DTV_ERR_OUT_OF_MEMORY (-3)
#define DTV_ERR_NOT_RENDERED (-4)

/* Adjoint flavours accepted by dtv_render_adjoint. */
#define DTV_ADJOINT_EXACT 0     /* derivative of the forward (default) */
#define DTV_ADJOINT_REFERENCE 1 /* AdjointIntersectRasterization.glsl as written (skips sub-segment 0) */

typedef struct dtv_ctx dtv_ctx;   /* per-device scratch and settings */
typedef struct dtv_view dtv_view; /* state of one rendered view, kept for its adjoint */

/* Tet mesh.  Layouts of TetMesh::getVertexPositionTensor/getVertexColorTensor/getCellColorTensor
 * (src/Tet/TetMesh.cpp:236-367); cells are int32 vertex indices, 4 per tet, with the reference's orientation
 * convention dot(cross(p1-p0, p2-p0), p3-p0) < 0 (src/Tet/TetMesh.cpp:532-541) (other orientations render the
 * same). Colour alpha is the density before attenuation.  Exactly one of colors (ColorStorage.PER_VERTEX:
 * colours interpolated over each segment, 4 sub-segments) and cell_colors (ColorStorage.PER_CELL: one colour per
 * tet, accumulateConst over the segment, IntersectRasterization.glsl:250-255) is set; the adjoint's colour
 * gradient then has the same layout. */
typedef struct dtv_mesh {
    const int32_t* cells;    /* device [num_cells * 4] */
    const float* positions;  /* device [num_vertices * 3] */
    const float* colors;     /* device [num_vertices * 4], or NULL */
    int64_t num_cells;
    int64_t num_vertices;
    const float* cell_colors; /* device [num_cells * 4], or NULL */
} dtv_mesh;

/* Camera: view matrix as Renderer.set_view_matrix (src/Module/Module.cpp:904-911, 16 floats column-major),
 * fovy as set_camera_fovy (:897-903), clip planes as the module's defaults (:457-463), viewport as
 * set_viewport_size (:874-887).  Projection: right-handed, depth in [0,1], y flipped (top-left origin). */
typedef struct dtv_camera {
    float view[16];
    float fovy;
    float znear;
    float zfar;
    int32_t width;
    int32_t height;
} dtv_camera;

/* Render settings: TetMeshVolumeRenderer attenuation/clear colour/early-ray-termination
 * (src/Renderer/TetMeshVolumeRenderer.hpp:237-244, Module.cpp:864-873). */
typedef struct dtv_render_params {
    float attenuation;
    float clear_color[4];
    int32_t use_early_ray_termination; /* applied only when nonzero (the module passes use_ert && use_gradients) */
    float early_ray_out_alpha;         /* stop once accumulated alpha > this; reference default 1 - 1e-4 */
    int32_t tile_size;                 /* binning tile edge in pixels: 8 (0 = default), 16 or 32; rendered as
                                          8x8 sub-tiles (no reference counterpart: the reference rasterises) */
    int32_t checkpoint_interval;       /* fragments between adjoint checkpoints; 0 = 64 */
    int32_t tile_order;                /* launch order of screen tiles (no reference counterpart: the reference
                                          rasterises): 0 = heaviest first (default; forward by record count,
                                          adjoint by the forward's fragment count), 1 = row-major */
    int32_t segment_length;            /* DTV_SEGMENT_EUCLIDEAN (INTERSECTION renderer) or
                                          DTV_SEGMENT_LINEAR_DEPTH (PPLL renderer) */
    int32_t renderer;                  /* DTV_RENDERER_TILE (0) or DTV_RENDERER_PROJECTION (1) */
    int32_t segments_per_pass;         /* ray segments the forward shades per round: 64, 128 (0 = default) or
                                          256 (no reference counterpart) */
    int32_t tile_launch;               /* DTV_TILE_LAUNCH_GRID (0, default): one workgroup per 8x8 tile;
                                          DTV_TILE_LAUNCH_QUEUE (1): a persistent grid of the workgroups the GPU
                                          holds at once takes the tiles in launch order from a device work queue
                                          (forward and adjoint; no reference counterpart) */
} dtv_render_params;

#define DTV_TILE_LAUNCH_GRID 0
#define DTV_TILE_LAUNCH_QUEUE 1

/* Which renderer dtv_render / dtv_render_batch runs.  TILE: per-pixel ray-tet segments of the INTERSECTION and
 * PPLL renderers (segment_length chooses the segment length).  PROJECTION: the Projection renderer's projected
 * tetrahedra (TetMeshRendererProjection::render, src/Renderer/TetMeshRendererProjection.cpp:811-903): each tet
 * in view becomes 3 or 4 screen triangles with a thickness at the thick vertex
 * (Data/Shaders/Projection/GenerateTrianglesVTK.glsl), sorted by centroid depth and blended with
 * alpha = 1 - exp(-density * thickness) of linearly interpolated colour and thickness
 * (ProjectedRasterization.glsl).  It has no early ray termination (use_early_ray_termination and
 * segment_length are ignored), and its adjoint (dtv_render_adjoint, DTV_ADJOINT_EXACT only) replaces
 * TetMeshRendererProjection::renderAdjoint (:905-971) with the exact derivative of that forward. */
#define DTV_RENDERER_TILE 0
#define DTV_RENDERER_PROJECTION 1

/* Length of a ray's segment through a tet.  EUCLIDEAN: the distance between the entry and exit points, as the
 * INTERSECTION renderer (Data/Shaders/Intersect/IntersectRasterization.glsl:211).  LINEAR_DEPTH: the difference
 * of the two faces' linear eye depths, as the PPLL renderer (Data/Shaders/PPLL/Forward.glsl:53,188), i.e. the
 * Euclidean length times the cosine between the ray and the view axis. */
#define DTV_SEGMENT_EUCLIDEAN 0
#define DTV_SEGMENT_LINEAR_DEPTH 1

typedef struct dtv_view_stats {
    int64_t num_triangles;    /* generated front faces (GenerateTrianglesIntersect.glsl:124) */
    int64_t num_pairs;        /* (tile, face) records after binning */
    int64_t num_tiles;
    int32_t sort_passes;      /* radix passes actually run */
    int32_t width, height;    /* viewport of the first item */
    int32_t num_items;        /* renders in the last batch (1 after dtv_render) */
    int64_t num_pixels;       /* pixels of all items */
    int64_t num_gradient_atomics;  /* float atomics of the last adjoint if it ran with timing on, else -1 */
} dtv_view_stats;

/* One render of a batch: a mesh seen through a camera into out_image [height, width, 4]. */
typedef struct dtv_render_item {
    dtv_mesh mesh;
    dtv_camera camera;
    float* out_image;
} dtv_render_item;

/* Adjoint input and gradient targets of one batch item: dimage [height, width, 4]; gradients accumulate (+=) into
 * grad_positions [V,3] and grad_colors ([V,4], or [T,4] for a mesh with cell_colors) (items that share a mesh may
 * share them).  mesh must equal the item's mesh as rendered (same buffers and sizes; checked). */
typedef struct dtv_adjoint_item {
    dtv_mesh mesh;
    const float* dimage;
    float* grad_positions;
    float* grad_colors;
} dtv_adjoint_item;

/* Adam over one parameter tensor: torch.optim.Adam as used by pytests/train.py:301-319,351, optional colour
 * clamp c -= min(c, 0) (train.py:352-356), optional fused zero_grad for the next step (train.py:342). */
typedef struct dtv_adam_group {
    float* param;
    float* grad;
    float* exp_avg;
    float* exp_avg_sq;
    int64_t numel;
    double lr;               /* hyper-parameters in double, as torch.optim.Adam keeps them: the bias corrections */
    double beta1, beta2, eps;/* are formed in double, the update's constants rounded once to float */
    int64_t step;            /* 1-based step count after this update */
    int32_t clamp_nonneg;    /* 1: param = max(param, 0) after the update */
    int32_t zero_grad;       /* 1: grad = 0 after it is consumed */
} dtv_adam_group;

/* ---- library ---------------------------------------------------------------------------------------- */
int dtv_get_version(void);
const char* dtv_last_error(void);

/* Replaces ApplicationState creation (src/Module/Module.cpp:192-489, ensureStateExists :646-650). */
int dtv_ctx_create(int32_t device, dtv_ctx** out_ctx);
int dtv_ctx_destroy(dtv_ctx* ctx);
int dtv_ctx_get_device(const dtv_ctx* ctx, int32_t* out_device);

/* Replaces the renderer object's intermediate buffers (TetMeshRendererIntersection.cpp:424-487). */
int dtv_view_create(dtv_ctx* ctx, dtv_view** out_view);
int dtv_view_destroy(dtv_view* view);

/* Host-only: the 68 camera uniform floats (viewProj, view, invProj, invView column-major, camera position),
 * the counterpart of the uniform upload in TetMeshRendererIntersection::render (:604-616). */
int dtv_camera_uniforms(const dtv_camera* camera, float* out68);

/* Forward render into out_image [height, width, 4] (premultiplied RGBA blended onto the clear colour).
 * Replaces Renderer.render (Module.cpp:912-921) -> TetMeshRendererIntersection::render (:603-707).
 * Synchronises the stream once to size the tile-pair buffers. */
int dtv_render(dtv_view* view, const dtv_mesh* mesh, const dtv_camera* camera, const dtv_render_params* params,
               float* out_image, void* stream);

/* Forward render of one (mesh, camera) into out_image, allocating nothing: every intermediate buffer lives in one
 * static device array of 56 MiB per GPU (held until the process ends), whose limits are 262,144 tets, 262,144
 * pixels (8,192 8x8 tiles) and 786,432 (tile, face) records; sizes beyond them fail with DTV_ERR_INVALID_ARGUMENT.
 * No host synchronisation and no host staging (capturable in a HIP graph).  *out_overflow (device int32, may be
 * NULL) receives 1 when the render produced more records than the limit (that image is incomplete), else 0.
 * The image equals dtv_render's.  Forward only: it keeps no state for an adjoint.  Every call on a GPU shares that
 * GPU's array, so calls that overlap on one GPU (other streams or threads) race.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (HIP graph capture, serving loops, harnesses
 * that forbid allocation). */
int dtv_render_noalloc(dtv_ctx* ctx, const dtv_mesh* mesh, const dtv_camera* camera, const dtv_render_params* params,
                       float* out_image, int32_t* out_overflow, void* stream);

/* Renders num_items (1..4096) (mesh, camera) items as one batch: one binning grid over all their tets, one sort of
 * all (tile, face) records and one tile-renderer launch over all their tiles, so that many small or uneven views
 * fill the GPU together.  Each image equals what dtv_render gives for its item alone.  No reference counterpart:
 * the reference renders one view per call (Module.cpp:912-921); dtv_render is the one-item case.
 * Synchronises the stream once to size the record buffers. */
int dtv_render_batch(dtv_view* view, const dtv_render_item* items, int32_t num_items,
                     const dtv_render_params* params, void* stream);

/* Adjoint of the last dtv_render_batch of this view, items[i] belonging to its render item i (same count).
 * One launch over all tiles of the batch; an item with dimage == NULL is not differentiated.  An item whose mesh
 * differs from the one rendered (buffers or sizes) fails with DTV_ERR_INVALID_ARGUMENT before any work. */
int dtv_render_adjoint_batch(dtv_view* view, const dtv_adjoint_item* items, int32_t num_items, int32_t use_abs_grad,
                             int32_t adjoint_mode, void* stream);

/* Adjoint of the last dtv_render of this view: accumulates (+=) into grad_positions [V,3] and
 * grad_colors [V,4] ([T,4] for cell colours).  Replaces Renderer.render_adjoint (Module.cpp:922-959) ->
 * TetMeshRendererIntersection::renderAdjoint (:709-715).  The mesh must be the one rendered (same cells,
 * positions and colours buffers and sizes; else DTV_ERR_INVALID_ARGUMENT). */
int dtv_render_adjoint(dtv_view* view, const dtv_mesh* mesh, const float* dimage, int32_t use_abs_grad,
                       int32_t adjoint_mode, float* grad_positions, float* grad_colors, void* stream);

/* Per-pixel number of fragments composited by the last render (depth complexity; the reference's
 * SHOW_DEPTH_COMPLEXITY counter, IntersectRasterization.glsl:267-270).  out: device int32 [H*W], or for a batch
 * the items' [H_i*W_i] one after another. */
int dtv_view_fragment_counts(dtv_view* view, int32_t* out_counts, void* stream);

/* Per-pixel ordered fragment ids (tet*4 + face), front to back, up to max_per_pixel each, padded with -1.
 * Counterpart of the PPLL export (Module.cpp:960-981).  out: device int32 [H*W*max_per_pixel].  Single-item
 * renders only. */
int dtv_view_export_fragments(dtv_view* view, const dtv_mesh* mesh, int32_t max_per_pixel, int32_t* out_ids,
                              void* stream);

int dtv_view_get_stats(dtv_view* view, dtv_view_stats* out);

/* Autotuner (no reference counterpart; the nearest is the PPLL renderer's per-mesh-size fragment budget,
 * src/Renderer/TetMeshRendererPPLL.cpp:290-318, .hpp:48-54).  Times dtv_render_batch (and, with_adjoint, the
 * exact adjoint into scratch gradients) of these items for every tile_size in {8, 16, 32} and segments_per_pass
 * in {64, 128, 256}: one untimed run, then `repeats` timed runs each, hipEvents on `stream`.  params supplies
 * the other settings.  The items' out_image buffers are written; the view is left rendered with the last
 * candidate.  The tile renderer only (params->renderer 0). */
typedef struct dtv_tune_result {
    int32_t tile_size;          /* the fastest candidate */
    int32_t segments_per_pass;
    float ms;                   /* its mean time per run */
    float ms_default;           /* 8x8 tiles, 128 segments per pass */
    int32_t num_candidates;
    int32_t cand_tile_size[9];
    int32_t cand_segments_per_pass[9];
    float cand_ms[9];
} dtv_tune_result;
int dtv_autotune(dtv_view* view, const dtv_render_item* items, int32_t num_items, const dtv_render_params* params,
                 int32_t with_adjoint, int32_t repeats, dtv_tune_result* out, void* stream);

/* Planned (synchronisation-free) rendering.  With record_capacity > 0, dtv_render / dtv_render_batch size the
 * (tile, face) record buffers for that many records instead of synchronising on the count, sort all four
 * depth-key digits, and let the kernels read the record count from device memory: nothing in a render (or an
 * adjoint) waits for the device, so a whole training step can be captured in a HIP graph and replayed (after
 * one ordinary call has sized every buffer).  If a render produces more records than the capacity, the extra
 * records are dropped (that image is wrong; nothing outside the buffers is touched) and
 * dtv_view_records_overflowed says so.  0 (default) = size exactly, with one synchronisation per render.
 * No counterpart in the reference (its GPU work is recorded in Vulkan command buffers). */
int dtv_view_set_record_capacity(dtv_view* view, int64_t record_capacity);
/* Synchronises the device.  *out_overflowed = 1 if the last render produced more records than the capacity;
 * *out_records (may be NULL) = the records it produced. */
int dtv_view_records_overflowed(dtv_view* view, int32_t* out_overflowed, int64_t* out_records);

/* Per-stage GPU timers (hipEvents recorded on the caller's stream around each stage of dtv_render /
 * dtv_render_adjoint).  No reference counterpart (the reference only has wall-clock opt_time,
 * pytests/train.py:366,434).  dtv_view_get_timings synchronises on the recorded events and returns the
 * milliseconds of the most recent call of each stage in out_ms[DTV_NUM_STAGES] (negative if not run). */
#define DTV_STAGE_COUNT_PAIRS 0 /* projection, culling and record counting */
#define DTV_STAGE_SCAN 1
#define DTV_STAGE_EMIT 2
#define DTV_STAGE_SORT 3
#define DTV_STAGE_RANGES 4 /* tile ranges and the record gather (exact coverage masks) */
#define DTV_STAGE_FORWARD 5
#define DTV_STAGE_ADJOINT 6
#define DTV_NUM_STAGES 7
int dtv_view_set_timing(dtv_view* view, int32_t enabled);
int dtv_view_get_timings(dtv_view* view, float* out_ms);

/* One fused launch over up to 8 parameter groups.  Replaces optimizer.step() + clamp (train.py:351-356)
 * and the GLSL Adam pass (Optimizer.glsl:136-169, OptimizerPass.cpp:120-177). */
int dtv_adam_step(const dtv_adam_group* groups, int32_t num_groups, void* stream);

/* The same update with the step count in device memory, for captured graphs: the groups' step fields are
 * ignored; *step_state (device int64, caller-owned, 0 before the first step) counts the steps taken, the bias
 * corrections use *step_state + 1, and the call increments it on the stream.  No reference counterpart. */
int dtv_adam_step_device(const dtv_adam_group* groups, int32_t num_groups, int64_t* step_state, void* stream);

/* Tet-quality regulariser: accumulates (+=) the gradient of sum over tets of the mirrored softplus
 * (lambda / beta) log(1 + exp(-beta q)), q = 6 sqrt(2) V / l_rms^3, into grad_positions [V,3].
 * Replaces TetRegularizer.compute_grad (Module.cpp:984-995) -> TetRegularizerPass (TetRegularizerPass.cpp:91-95,
 * Data/Shaders/Optimizer/TetRegularizer.glsl:108-166). */
int dtv_tet_regularizer_grad(const dtv_mesh* mesh, float lambda, float beta, float* grad_positions, void* stream);

/* ---- Coarse-to-fine refinement (host arrays; no device work) ----
 * Subdivides, one after another, the ceil(splits_ratio * V) vertices whose gradient rows ([V,3] position or
 * [V,4] colour gradients, gradient_components 3 or 4) have the largest length, ties to the larger index: each
 * incident edge gets its midpoint (position and colour at t = 0.5), each incident tet becomes a corner tet and
 * a prism split into three tets, the prism diagonals chosen by the smallest-vertex-index rule.  The result is
 * read with dtv_split_result_sizes / dtv_split_result_copy and freed with dtv_split_result_destroy.
 * Replaces TetMesh::splitByLargestGradientMagnitudes (src/Tet/TetMesh.cpp:1786-1879, bound at
 * Module.cpp:820-832) and TetMesh::subdivideVertices (TetMesh.cpp:1757-1784). */
typedef struct dtv_split_result dtv_split_result;
int dtv_split_by_largest_gradient_magnitudes(const int32_t* cells, int64_t num_cells, const float* positions,
                                             const float* colors, int64_t num_vertices, const float* gradients,
                                             int32_t gradient_components, float splits_ratio,
                                             dtv_split_result** out_result);
/* The same with the magnitudes given and the number of splits explicit (TetMesh::subdivideVertices). */
int dtv_subdivide_vertices(const int32_t* cells, int64_t num_cells, const float* positions, const float* colors,
                           int64_t num_vertices, const float* gradient_magnitudes, int64_t num_splits,
                           dtv_split_result** out_result);
int dtv_split_result_sizes(const dtv_split_result* result, int64_t* num_cells, int64_t* num_vertices);
/* cells [num_cells,4] int32, positions [num_vertices,3], colors [num_vertices,4], host memory. */
int dtv_split_result_copy(const dtv_split_result* result, int32_t* cells, float* positions, float* colors);
int dtv_split_result_destroy(dtv_split_result* result);

/* MSE loss and its image adjoint: dimage = 2 (image - target) / n, *loss_out (device float, may be NULL)
 * = mean squared error.  Replaces torch.nn.MSELoss in train.py:325,346 and Loss.glsl:55-95. */
int dtv_mse_loss_grad(const float* image, const float* target, int64_t n, float* dimage, float* loss_out,
                      void* stream);

/* MSE summed over num_items equally sized images laid out one after another (n_per_item floats each):
 * *loss_out = sum over items of the item's mean squared error, dimage = 2 (image - target) / n_per_item.
 * The multi-view loss of a batched step (the reference sums per-view losses, train.py:325-346). */
int dtv_mse_loss_grad_batch(const float* image, const float* target, int64_t n_per_item, int32_t num_items,
                            float* dimage, float* loss_out, void* stream);

/* ---- render server ------------------------------------------------------------------------------------
 * Meshes resident in device memory and a request queue served by worker threads, each owning a view and a
 * stream: a free worker takes the pending requests in arrival order (up to max_batch, within memory_budget
 * bytes of estimated transient device memory for all batches in flight; 0 = unbounded) and renders them as
 * one dtv_render_batch, so requests that arrive while other batches are on the GPU are admitted as workers
 * and memory free up, and their work overlaps on the device.  A request with a target image is also
 * differentiated: MSE loss against the target, then the adjoint accumulating into its gradient buffers.
 * Every result equals what dtv_render gives for the request alone.  No counterpart in the reference (it
 * renders one view per Python call); errors of these calls are reported by dtv_server_last_error(). */
typedef struct dtv_server dtv_server;

typedef struct dtv_server_request {
    int32_t mesh_id;            /* from dtv_server_add_mesh */
    dtv_camera camera;
    float* out_image;           /* device [height, width, 4], written when the request completes */
    const float* target_image;  /* optional device [height, width, 4]: also run loss + adjoint */
    float* grad_positions;      /* with target_image: device [V,3], accumulated */
    float* grad_colors;         /* with target_image: device [V,4] ([T,4] for cell colours), accumulated */
} dtv_server_request;

typedef struct dtv_server_stats {
    int64_t submitted, completed, batches, max_batch_seen;
} dtv_server_stats;

const char* dtv_server_last_error(void);
int dtv_server_create(dtv_ctx* ctx, int32_t num_workers, int32_t max_batch, int64_t memory_budget,
                      const dtv_render_params* params, dtv_server** out_server);
/* Waits for every worker to finish its batch; requests still queued are rendered first. */
int dtv_server_destroy(dtv_server* server);
/* The mesh's device buffers are borrowed and must outlive the server. */
int dtv_server_add_mesh(dtv_server* server, const dtv_mesh* mesh, int32_t* out_mesh_id);
/* Thread-safe, non-blocking. */
int dtv_server_submit(dtv_server* server, const dtv_server_request* request, uint64_t* out_ticket);
/* Blocks until the request completed; latency = submit to completion, queue = submit to batch start (ms). */
int dtv_server_wait(dtv_server* server, uint64_t ticket, double* out_latency_ms, double* out_queue_ms);
int dtv_server_get_stats(dtv_server* server, dtv_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* DIFFTETVR_AMD_H */
