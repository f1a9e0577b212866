// Render server: meshes resident in device memory, a queue of render requests, and worker threads that batch
// the pending requests (across meshes and resolutions) into one dtv_render_batch each.  Continuous batching:
// every worker owns a view and a stream, so while one batch runs on the GPU the next worker admits the
// requests that arrived meanwhile and launches them; admission is bounded by a device-memory budget for the
// batches in flight.  Requests with a target image are also differentiated (MSE loss, then the batched
// adjoint accumulating into the request's gradient buffers); they are batched apart from renders, one such
// batch at a time, so that renders never wait behind an adjoint.  No counterpart in the reference, which
// renders one view per call from Python.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/difftetvr_amd.h"

namespace {

thread_local std::string g_server_error;

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    dtv_server_request req{};
    int64_t mem = 0;
    Clock::time_point t_submit, t_start, t_done;
    int status = 1;  // 1 pending, 0 done, < 0 error
    std::string error;
};

}  // namespace

struct dtv_server {
    dtv_ctx* ctx = nullptr;
    int device = 0;
    dtv_render_params params{};
    int32_t max_batch = 16;
    int64_t budget = 0;
    std::vector<dtv_mesh> meshes;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending_render, pending_train;  // forward-only requests; requests with a target
    bool train_in_flight = false;                        // one differentiated batch on the GPU at a time
    std::unordered_map<uint64_t, Request*> requests;
    uint64_t next_ticket = 1;
    int64_t in_flight_mem = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    dtv_server_stats stats{};
};

namespace {

// Transient device memory of one request: per-pixel state, its share of the record and binning buffers.
int64_t request_memory(const dtv_mesh& m, const dtv_camera& c) {
    const int64_t P = (int64_t)c.width * c.height;
    return P * 560 + m.num_cells * 80;
}

void worker_main(dtv_server* S) {
    (void)hipSetDevice(S->device);
    dtv_view* view = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done_ev = nullptr;
    float* scratch = nullptr;  // dimage of differentiated requests
    int64_t scratch_cap = 0;
    const int init_ok = dtv_view_create(S->ctx, &view) == DTV_OK &&
                        hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) == hipSuccess &&
                        hipEventCreateWithFlags(&done_ev, hipEventDisableTiming) == hipSuccess;
    std::vector<Request*> batch;
    std::vector<dtv_render_item> items;
    std::vector<dtv_adjoint_item> adj;
    bool quit = false;
    while (!quit) {
        batch.clear();
        int64_t mem = 0;
        bool train = false;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            // Renders and differentiated requests form separate batches.  At most one differentiated batch runs
            // at a time (its adjoint is several times a render's work), so the other workers keep rendering
            // beside it; between the two classes the older request goes first.
            for (;;) {
                const bool t_ok = !S->pending_train.empty() && !S->train_in_flight;
                const bool r_ok = !S->pending_render.empty();
                if (t_ok || r_ok) {
                    train = t_ok && (!r_ok || S->pending_train.front()->ticket < S->pending_render.front()->ticket);
                    break;
                }
                if (S->stop && S->pending_render.empty() && S->pending_train.empty()) {
                    quit = true;
                    break;
                }
                S->cv_work.wait(lk);
            }
            if (quit) break;
            std::deque<Request*>& queue = train ? S->pending_train : S->pending_render;
            // admit in arrival order while the batch has room and the budget allows (one request always fits)
            while (!queue.empty() && (int32_t)batch.size() < S->max_batch) {
                Request* r = queue.front();
                if (!batch.empty() || S->in_flight_mem > 0) {
                    if (S->budget > 0 && S->in_flight_mem + mem + r->mem > S->budget) break;
                }
                queue.pop_front();
                batch.push_back(r);
                mem += r->mem;
            }
            if (batch.empty()) {
                // budget exhausted by batches in flight: wait for one to finish
                S->cv_work.wait(lk, [&] { return S->stop || S->in_flight_mem == 0 || queue.empty(); });
                continue;
            }
            if (train) S->train_in_flight = true;
            S->in_flight_mem += mem;
            S->stats.batches++;
            S->stats.max_batch_seen = std::max<int64_t>(S->stats.max_batch_seen, (int64_t)batch.size());
        }
        const Clock::time_point t_start = Clock::now();
        int rc = init_ok ? DTV_OK : DTV_ERR_HIP;
        std::string err = init_ok ? "" : "server worker initialisation failed";
        items.assign(batch.size(), dtv_render_item{});
        int64_t adj_pixels = 0;
        size_t n_adj = 0;
        for (size_t i = 0; i < batch.size(); i++) {
            const dtv_server_request& q = batch[i]->req;
            items[i].mesh = S->meshes[(size_t)q.mesh_id];
            items[i].camera = q.camera;
            items[i].out_image = q.out_image;
            if (q.target_image) {
                n_adj++;
                adj_pixels += (int64_t)q.camera.width * q.camera.height;
            }
        }
        const bool any_adj = n_adj > 0;
        if (rc == DTV_OK) {
            rc = dtv_render_batch(view, items.data(), (int32_t)items.size(), &S->params, stream);
            if (rc != DTV_OK) err = dtv_last_error();
        }
        if (rc == DTV_OK && any_adj) {
            if (scratch_cap < 4 * adj_pixels) {
                if (scratch) (void)hipFree(scratch);
                scratch = nullptr;
                scratch_cap = 0;
                if (hipMalloc((void**)&scratch, sizeof(float) * 4 * adj_pixels) == hipSuccess) scratch_cap = 4 * adj_pixels;
            }
            if (scratch_cap < 4 * adj_pixels) {
                rc = DTV_ERR_OUT_OF_MEMORY;
                err = "server: out of device memory for image adjoints";
            }
            adj.assign(batch.size(), dtv_adjoint_item{});
            int64_t off = 0;
            for (size_t i = 0; rc == DTV_OK && i < batch.size(); i++) {
                const dtv_server_request& q = batch[i]->req;
                if (!q.target_image) continue;
                const int64_t n = 4LL * q.camera.width * q.camera.height;
                rc = dtv_mse_loss_grad(q.out_image, q.target_image, n, scratch + off, nullptr, stream);
                adj[i].mesh = items[i].mesh;
                adj[i].dimage = scratch + off;
                adj[i].grad_positions = q.grad_positions;
                adj[i].grad_colors = q.grad_colors;
                off += n;
            }
            if (rc == DTV_OK) rc = dtv_render_adjoint_batch(view, adj.data(), (int32_t)adj.size(), 0, DTV_ADJOINT_EXACT, stream);
            if (rc != DTV_OK && err.empty()) err = dtv_last_error();
        }
        if (init_ok) {
            (void)hipEventRecord(done_ev, stream);
            if (hipEventSynchronize(done_ev) != hipSuccess && rc == DTV_OK) {
                rc = DTV_ERR_HIP;
                err = "server: batch failed on the device";
            }
        }
        const Clock::time_point t_done = Clock::now();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            for (Request* r : batch) {
                r->t_start = t_start;
                r->t_done = t_done;
                r->status = rc == DTV_OK ? 0 : rc;
                r->error = err;
                S->stats.completed++;
            }
            S->in_flight_mem -= mem;
            if (train) S->train_in_flight = false;
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
    if (scratch) (void)hipFree(scratch);
    if (done_ev) (void)hipEventDestroy(done_ev);
    if (view) (void)dtv_view_destroy(view);
    if (stream) (void)hipStreamDestroy(stream);
}

int sfail(int code, const std::string& m) {
    g_server_error = m;
    return code;
}

}  // namespace

extern "C" {

const char* dtv_server_last_error(void) { return g_server_error.c_str(); }

int dtv_server_create(dtv_ctx* ctx, int32_t num_workers, int32_t max_batch, int64_t memory_budget,
                      const dtv_render_params* params, dtv_server** out) {
    if (!ctx || !params || !out) return sfail(DTV_ERR_INVALID_ARGUMENT, "dtv_server_create: NULL argument");
    int32_t device = 0;
    if (dtv_ctx_get_device(ctx, &device) != DTV_OK) return sfail(DTV_ERR_INVALID_ARGUMENT, "dtv_server_create: bad ctx");
    if (num_workers < 1 || num_workers > 8 || max_batch < 1 || max_batch > 4096 || memory_budget < 0)
        return sfail(DTV_ERR_INVALID_ARGUMENT, "dtv_server_create: 1..8 workers, 1..4096 requests per batch");
    dtv_server* S = new dtv_server();
    S->ctx = ctx;
    S->device = device;
    S->params = *params;
    S->max_batch = max_batch;
    S->budget = memory_budget;
    for (int i = 0; i < num_workers; i++) S->workers.emplace_back(worker_main, S);
    *out = S;
    return DTV_OK;
}

int dtv_server_destroy(dtv_server* S) {
    if (!S) return DTV_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->workers) t.join();
    for (auto& kv : S->requests) delete kv.second;
    delete S;
    return DTV_OK;
}

int dtv_server_add_mesh(dtv_server* S, const dtv_mesh* mesh, int32_t* out_mesh_id) {
    if (!S || !mesh || !out_mesh_id) return sfail(DTV_ERR_INVALID_ARGUMENT, "dtv_server_add_mesh: NULL argument");
    if (mesh->num_cells < 0 || mesh->num_vertices < 0 || (mesh->colors != nullptr) == (mesh->cell_colors != nullptr) ||
        (mesh->num_cells > 0 && (!mesh->cells || !mesh->positions)))
        return sfail(DTV_ERR_INVALID_ARGUMENT, "dtv_server_add_mesh: invalid mesh");
    std::lock_guard<std::mutex> lk(S->mu);
    S->meshes.push_back(*mesh);
    *out_mesh_id = (int32_t)S->meshes.size() - 1;
    return DTV_OK;
}

int dtv_server_submit(dtv_server* S, const dtv_server_request* req, uint64_t* out_ticket) {
    if (!S || !req || !out_ticket) return sfail(DTV_ERR_INVALID_ARGUMENT, "dtv_server_submit: NULL argument");
    if (!req->out_image) return sfail(DTV_ERR_INVALID_ARGUMENT, "dtv_server_submit: NULL out_image");
    if (req->target_image && (!req->grad_positions || !req->grad_colors))
        return sfail(DTV_ERR_INVALID_ARGUMENT, "dtv_server_submit: a target image needs gradient buffers");
    Request* r = new Request();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (req->mesh_id < 0 || req->mesh_id >= (int32_t)S->meshes.size()) {
            delete r;
            return sfail(DTV_ERR_INVALID_ARGUMENT, "dtv_server_submit: unknown mesh_id");
        }
        r->ticket = S->next_ticket++;
        r->req = *req;
        r->mem = request_memory(S->meshes[(size_t)req->mesh_id], req->camera);
        r->t_submit = Clock::now();
        S->requests[r->ticket] = r;
        (req->target_image ? S->pending_train : S->pending_render).push_back(r);
        S->stats.submitted++;
        *out_ticket = r->ticket;
    }
    S->cv_work.notify_one();
    return DTV_OK;
}

int dtv_server_wait(dtv_server* S, uint64_t ticket, double* out_latency_ms, double* out_queue_ms) {
    if (!S) return sfail(DTV_ERR_INVALID_ARGUMENT, "dtv_server_wait: NULL server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->requests.find(ticket);
    if (it == S->requests.end()) return sfail(DTV_ERR_INVALID_ARGUMENT, "dtv_server_wait: unknown ticket");
    Request* r = it->second;
    S->cv_done.wait(lk, [&] { return r->status != 1; });
    const int status = r->status;
    const std::string err = r->error;
    if (out_latency_ms)
        *out_latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (out_queue_ms)
        *out_queue_ms = std::chrono::duration<double, std::milli>(r->t_start - r->t_submit).count();
    S->requests.erase(it);
    delete r;
    if (status != 0) return sfail(status, "dtv_server_wait: " + err);
    return DTV_OK;
}

int dtv_server_get_stats(dtv_server* S, dtv_server_stats* out) {
    if (!S || !out) return sfail(DTV_ERR_INVALID_ARGUMENT, "dtv_server_get_stats: NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->stats;
    return DTV_OK;
}

}  // extern "C"
