// serve.cpp -- continuous-batching request loop for surface-wave likelihoods.
//
// Many chains submit likelihood() requests (their Voronoi sites and noise);
// a dispatcher thread packs whatever is pending into a batch for the first
// free lane and launches it.  A lane is one surface context (its own HIP
// stream and device arrays for max_batch chains), so batches on different
// lanes overlap on the device, and requests that arrive while batches are
// running are admitted as soon as a lane frees.  Device memory is fixed at
// create time: admission is bounded by lanes x max_batch slots.  Results are
// the same as a plain call's, bit for bit (every chain in a batch is
// independent).
#include "../../include/mctomo_hip.h"
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;

struct request {
    int64_t ticket;
    int ncells;
    std::vector<double> points, params, noise;   // (ncells,3), (ncells,3), (2,np)
    clk::time_point t_submit;
    bool done = false;
    int status = 0;
    double out[3] = {0, 0, 0};
    double latency_ms = 0;
};

struct lane {
    mct_surf_ctx *ctx = nullptr;
    double *d_points = nullptr, *d_params = nullptr, *d_noise = nullptr, *d_out = nullptr;
    int32_t *d_ncells = nullptr, *d_status = nullptr;
    double *h_points = nullptr, *h_params = nullptr, *h_noise = nullptr, *h_out = nullptr;
    int32_t *h_ncells = nullptr, *h_status = nullptr;
    hipStream_t stream = nullptr;    // the context's stream (copies go on it too)
    std::vector<request *> batch;
    bool busy = false;
};

} // namespace

struct mct_server {
    int max_batch = 0, max_cells = 0, np = 0;
    std::vector<lane> lanes;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<request *> pending;
    std::unordered_map<int64_t, request *> all;
    int64_t next_ticket = 1;
    bool stop = false;
    int error = 0;
    std::thread dispatcher;
    int64_t batches = 0, evals = 0;
};

static int lane_alloc(lane &L, int B, int C, int np)
{
    auto dm = [](void **p, size_t n) { return hipMalloc(p, n ? n : 16) == hipSuccess; };
    auto hm = [](void **p, size_t n) { return hipHostMalloc(p, n ? n : 16, hipHostMallocDefault) == hipSuccess; };
    if (!dm((void **)&L.d_points, sizeof(double) * 3 * C * B) || !dm((void **)&L.d_params, sizeof(double) * 3 * C * B) ||
        !dm((void **)&L.d_noise, sizeof(double) * 2 * np * B) || !dm((void **)&L.d_out, sizeof(double) * 3 * B) ||
        !dm((void **)&L.d_ncells, sizeof(int32_t) * B) || !dm((void **)&L.d_status, sizeof(int32_t) * B))
        return MCT_ERR_NOMEM;
    if (!hm((void **)&L.h_points, sizeof(double) * 3 * C * B) || !hm((void **)&L.h_params, sizeof(double) * 3 * C * B) ||
        !hm((void **)&L.h_noise, sizeof(double) * 2 * np * B) || !hm((void **)&L.h_out, sizeof(double) * 3 * B) ||
        !hm((void **)&L.h_ncells, sizeof(int32_t) * B) || !hm((void **)&L.h_status, sizeof(int32_t) * B))
        return MCT_ERR_NOMEM;
    return MCT_OK;
}

static void lane_free(lane &L)
{
    if (L.ctx) mct_surf_destroy(L.ctx);
    void *d[] = {L.d_points, L.d_params, L.d_noise, L.d_out, L.d_ncells, L.d_status};
    for (void *p : d) if (p) (void)hipFree(p);
    void *h[] = {L.h_points, L.h_params, L.h_noise, L.h_out, L.h_ncells, L.h_status};
    for (void *p : h) if (p) (void)hipHostFree(p);
    L = lane();
}

// Pack the lane's batch into its pinned buffers and launch it (no wait).
static int lane_launch(mct_server *S, lane &L)
{
    const int B = (int)L.batch.size(), C = S->max_cells, np = S->np;
    for (int i = 0; i < B; i++) {
        request *r = L.batch[i];
        memcpy(L.h_points + (size_t)3 * C * i, r->points.data(), sizeof(double) * 3 * r->ncells);
        memcpy(L.h_params + (size_t)3 * C * i, r->params.data(), sizeof(double) * 3 * r->ncells);
        memcpy(L.h_noise + (size_t)2 * np * i, r->noise.data(), sizeof(double) * 2 * np);
        L.h_ncells[i] = r->ncells;
    }
    hipStream_t s = L.stream;
    if (hipMemcpyAsync(L.d_points, L.h_points, sizeof(double) * 3 * C * B, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(L.d_params, L.h_params, sizeof(double) * 3 * C * B, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(L.d_noise, L.h_noise, sizeof(double) * 2 * np * B, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(L.d_ncells, L.h_ncells, sizeof(int32_t) * B, hipMemcpyHostToDevice, s) != hipSuccess)
        return MCT_ERR_HIP;
    int rc = mct_surf_likelihood_batch_dev(L.ctx, B, L.d_points, L.d_params, L.d_ncells, L.d_noise, L.d_out,
                                           L.d_status);
    if (rc) return rc;
    if (hipMemcpyAsync(L.h_out, L.d_out, sizeof(double) * 3 * B, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(L.h_status, L.d_status, sizeof(int32_t) * B, hipMemcpyDeviceToHost, s) != hipSuccess)
        return MCT_ERR_HIP;
    return MCT_OK;
}

static void dispatcher_main(mct_server *S)
{
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        // complete finished lanes
        bool any_busy = false;
        for (lane &L : S->lanes) {
            if (!L.busy) continue;
            lk.unlock();
            hipError_t q = hipStreamQuery(L.stream);
            lk.lock();
            if (q == hipErrorNotReady) { any_busy = true; continue; }
            const auto now = clk::now();
            for (size_t i = 0; i < L.batch.size(); i++) {
                request *r = L.batch[i];
                r->status = q == hipSuccess ? L.h_status[i] : MCT_ERR_HIP;
                for (int k = 0; k < 3; k++) r->out[k] = L.h_out[3 * i + k];
                r->latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
                r->done = true;
            }
            S->evals += (int64_t)L.batch.size();
            L.batch.clear();
            L.busy = false;
            S->cv_done.notify_all();
        }
        // admit pending requests into free lanes
        for (lane &L : S->lanes) {
            if (L.busy || S->pending.empty()) continue;
            while (!S->pending.empty() && (int)L.batch.size() < S->max_batch) {
                L.batch.push_back(S->pending.front());
                S->pending.pop_front();
            }
            L.busy = true;
            any_busy = true;
            lk.unlock();
            int rc = lane_launch(S, L);
            lk.lock();
            S->batches++;
            if (rc) {
                for (request *r : L.batch) { r->status = rc; r->done = true; }
                L.batch.clear();
                L.busy = false;
                S->error = rc;
                S->cv_done.notify_all();
            }
        }
        if (S->stop && S->pending.empty() && !any_busy) return;
        if (any_busy) {
            // poll the running lanes; new requests are picked up on the next pass
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(200));
            lk.lock();
        } else {
            S->cv_work.wait(lk, [&] { return S->stop || !S->pending.empty(); });
        }
    }
}

extern "C" {

int mct_serve_create(const mct_grid *grid, const mct_surf_setup *setup, int max_batch, int nlanes,
                     mct_server **out)
{
    if (!grid || !setup || !out || max_batch < 1 || nlanes < 1 || nlanes > 16) return MCT_ERR_ARG;
    *out = nullptr;
    mct_server *S = new (std::nothrow) mct_server();
    if (!S) return MCT_ERR_NOMEM;
    S->max_batch = max_batch;
    S->max_cells = setup->max_cells;
    S->np = setup->np;
    S->lanes.resize(nlanes);
    mct_surf_setup su = *setup;
    su.max_chains = max_batch;
    for (lane &L : S->lanes) {
        int rc = mct_surf_create(grid, &su, &L.ctx);
        if (!rc) rc = lane_alloc(L, max_batch, setup->max_cells, setup->np);
        if (!rc) rc = mct_surf_stream(L.ctx, (void **)&L.stream);
        if (rc) {
            for (lane &M : S->lanes) lane_free(M);
            delete S;
            return rc;
        }
    }
    S->dispatcher = std::thread(dispatcher_main, S);
    *out = S;
    return MCT_OK;
}

int mct_serve_submit(mct_server *S, const double *points, const double *params, int ncells,
                     const double *snoise0, const double *snoise1, int64_t *ticket)
{
    if (!S || !points || !params || !ticket || ncells < 1 || ncells > S->max_cells) return MCT_ERR_ARG;
    request *r = new (std::nothrow) request();
    if (!r) return MCT_ERR_NOMEM;
    r->ncells = ncells;
    r->points.assign(points, points + 3 * (size_t)ncells);
    r->params.assign(params, params + 3 * (size_t)ncells);
    r->noise.assign(2 * (size_t)S->np, 0.0);
    if (snoise0) memcpy(r->noise.data(), snoise0, sizeof(double) * S->np);
    if (snoise1) memcpy(r->noise.data() + S->np, snoise1, sizeof(double) * S->np);
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (S->stop) { delete r; return MCT_ERR_ARG; }
        r->ticket = S->next_ticket++;
        r->t_submit = clk::now();
        S->all[r->ticket] = r;
        S->pending.push_back(r);
    }
    S->cv_work.notify_one();
    *ticket = r->ticket;
    return MCT_OK;
}

int mct_serve_wait(mct_server *S, int64_t ticket, double *out, int32_t *status, double *latency_ms)
{
    if (!S || !out) return MCT_ERR_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) return MCT_ERR_ARG;
    request *r = it->second;
    S->cv_done.wait(lk, [&] { return r->done; });
    for (int k = 0; k < 3; k++) out[k] = r->out[k];
    if (status) *status = r->status;
    if (latency_ms) *latency_ms = r->latency_ms;
    S->all.erase(it);
    delete r;
    return MCT_OK;
}

int mct_serve_stats(mct_server *S, int64_t *batches, int64_t *evals)
{
    if (!S) return MCT_ERR_ARG;
    std::lock_guard<std::mutex> g(S->mu);
    if (batches) *batches = S->batches;
    if (evals) *evals = S->evals;
    return S->error;
}

int mct_serve_destroy(mct_server *S)
{
    if (!S) return MCT_OK;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->dispatcher.joinable()) S->dispatcher.join();
    for (auto &kv : S->all) delete kv.second;
    for (lane &L : S->lanes) lane_free(L);
    delete S;
    return MCT_OK;
}

} // extern "C"
