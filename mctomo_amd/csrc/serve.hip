// serve.hip -- continuous-batching request loop for surface-wave likelihoods.
//
// Many chains submit likelihood() requests (their Voronoi sites and noise);
// a dispatcher thread packs whatever is pending into a batch for a free lane
// and launches it.  A lane is one surface context (its own HIP stream and
// device arrays for max_batch chains), so batches on different lanes overlap
// on the device, and requests that arrive while batches are running are
// admitted as soon as a lane frees.  Results are the same as a plain call's,
// bit for bit (every chain in a batch is independent).
//
// A server can serve several models (data sets on their own grids: the
// served mix); a request names its model and goes to a lane of that model.
// Lanes are admitted by device-memory footprint: with a budget, a lane is
// created only while the bytes of all lanes stay within it, and when a model
// has waiting requests and no room, idle lanes of models with nothing pending
// are released first (least recently used), so memory follows the mix.
//
// Batch fill: under steady arrivals, launching whatever is pending the moment
// a lane frees gives small batches that share the device, and a small batch
// costs nearly as much as a full one.  So a partial batch waits while another
// lane is running, until the oldest of its requests has waited fill_ms (or the
// batch is full); with no lane running it goes at once, so the device never
// idles while requests wait.  fill_ms < 0 turns the wait off.
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
    int model;
    int ncells;
    std::vector<double> points, params, noise;   // (ncells,3), (ncells,3), (2,np)
    clk::time_point t_submit;
    bool done = false;
    int status = 0;
    double out[3] = {0, 0, 0};
    double latency_ms = 0;
};

struct lane {
    int model = 0;
    mct_surf_ctx *ctx = nullptr;
    double *d_points = nullptr, *d_params = nullptr, *d_noise = nullptr, *d_out = nullptr;
    int32_t *d_ncells = nullptr, *d_status = nullptr;
    double *h_points = nullptr, *h_params = nullptr, *h_noise = nullptr, *h_out = nullptr;
    int32_t *h_ncells = nullptr, *h_status = nullptr;
    hipStream_t stream = nullptr;    // the context's stream (copies go on it too)
    std::vector<request *> batch;
    bool busy = false;
    int64_t bytes = 0;               // device bytes of the context and the lane's arrays
    clk::time_point last_used;
};

// One data set of the mix: its grid and setup (with the host arrays the setup
// points to owned here, since lanes are created later), its queue and lanes.
struct model {
    mct_grid grid;
    mct_surf_setup setup;
    std::vector<double> freqs, src, rev, ttime;
    std::vector<int32_t> raystat;
    int max_batch = 0;
    std::deque<request *> pending;
    int nlanes = 0;
    int64_t lane_bytes = 0;          // footprint of one lane (known after the first)
};

} // namespace

struct mct_server {
    std::vector<model> models;
    std::vector<lane *> lanes;
    int max_lanes = 1;               // per model
    int64_t budget = 0;              // device bytes for all lanes (0: no limit)
    int64_t used = 0, peak = 0, releases = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::unordered_map<int64_t, request *> all;
    int64_t next_ticket = 1;
    bool stop = false;
    int error = 0;
    std::thread dispatcher;
    int64_t batches = 0, evals = 0;
    int rr = 0;                      // round-robin start over models
    double fill_ms = 25.0;           // longest wait of a partial batch while a lane runs (< 0: none)
};

static void lane_free(lane *L)
{
    if (L->ctx) mct_surf_destroy(L->ctx);
    void *d[] = {L->d_points, L->d_params, L->d_noise, L->d_out, L->d_ncells, L->d_status};
    for (void *p : d) if (p) (void)hipFree(p);
    void *h[] = {L->h_points, L->h_params, L->h_noise, L->h_out, L->h_ncells, L->h_status};
    for (void *p : h) if (p) (void)hipHostFree(p);
    delete L;
}

// A lane of model M: its surface context and its staging arrays.
static int lane_create(const model &M, lane **out)
{
    *out = nullptr;
    lane *L = new (std::nothrow) lane();
    if (!L) return MCT_ERR_NOMEM;
    const int B = M.max_batch, C = M.setup.max_cells, np = M.setup.np;
    mct_surf_setup su = M.setup;
    su.max_chains = B;
    int rc = mct_surf_create(&M.grid, &su, &L->ctx);
    auto dm = [](void **p, size_t n) { return hipMalloc(p, n ? n : 16) == hipSuccess; };
    auto hm = [](void **p, size_t n) { return hipHostMalloc(p, n ? n : 16, hipHostMallocDefault) == hipSuccess; };
    const size_t dbytes[6] = {sizeof(double) * 3 * C * B, sizeof(double) * 3 * C * B, sizeof(double) * 2 * np * B,
                              sizeof(double) * 3 * B, sizeof(int32_t) * B, sizeof(int32_t) * B};
    if (!rc && (!dm((void **)&L->d_points, dbytes[0]) || !dm((void **)&L->d_params, dbytes[1]) ||
                !dm((void **)&L->d_noise, dbytes[2]) || !dm((void **)&L->d_out, dbytes[3]) ||
                !dm((void **)&L->d_ncells, dbytes[4]) || !dm((void **)&L->d_status, dbytes[5])))
        rc = MCT_ERR_NOMEM;
    if (!rc && (!hm((void **)&L->h_points, dbytes[0]) || !hm((void **)&L->h_params, dbytes[1]) ||
                !hm((void **)&L->h_noise, dbytes[2]) || !hm((void **)&L->h_out, dbytes[3]) ||
                !hm((void **)&L->h_ncells, dbytes[4]) || !hm((void **)&L->h_status, dbytes[5])))
        rc = MCT_ERR_NOMEM;
    if (!rc) rc = mct_surf_stream(L->ctx, (void **)&L->stream);
    int64_t cb = 0;
    if (!rc) rc = mct_surf_device_bytes(L->ctx, &cb);
    if (rc) { lane_free(L); return rc; }
    L->bytes = cb;
    for (size_t b : dbytes) L->bytes += (int64_t)b;
    L->last_used = clk::now();
    *out = L;
    return MCT_OK;
}

// Pack the lane's batch into its pinned buffers and launch it (no wait).
static int lane_launch(const model &M, lane *L)
{
    const int B = (int)L->batch.size(), C = M.setup.max_cells, np = M.setup.np;
    for (int i = 0; i < B; i++) {
        request *r = L->batch[i];
        memcpy(L->h_points + (size_t)3 * C * i, r->points.data(), sizeof(double) * 3 * r->ncells);
        memcpy(L->h_params + (size_t)3 * C * i, r->params.data(), sizeof(double) * 3 * r->ncells);
        memcpy(L->h_noise + (size_t)2 * np * i, r->noise.data(), sizeof(double) * 2 * np);
        L->h_ncells[i] = r->ncells;
    }
    hipStream_t s = L->stream;
    if (hipMemcpyAsync(L->d_points, L->h_points, sizeof(double) * 3 * C * B, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(L->d_params, L->h_params, sizeof(double) * 3 * C * B, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(L->d_noise, L->h_noise, sizeof(double) * 2 * np * B, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(L->d_ncells, L->h_ncells, sizeof(int32_t) * B, hipMemcpyHostToDevice, s) != hipSuccess)
        return MCT_ERR_HIP;
    // a lane's buffers never move, so each batch size replays its own captured graph
    int rc = mct_surf_likelihood_batch_graph(L->ctx, B, L->d_points, L->d_params, L->d_ncells, L->d_noise, L->d_out,
                                             L->d_status);
    if (rc == MCT_ERR_ARG)   // a batch in parts has no graph form
        rc = mct_surf_likelihood_batch_dev(L->ctx, B, L->d_points, L->d_params, L->d_ncells, L->d_noise, L->d_out,
                                           L->d_status);
    if (rc) return rc;
    if (hipMemcpyAsync(L->h_out, L->d_out, sizeof(double) * 3 * B, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(L->h_status, L->d_status, sizeof(int32_t) * B, hipMemcpyDeviceToHost, s) != hipSuccess)
        return MCT_ERR_HIP;
    return MCT_OK;
}

// Room for a lane of `need` bytes: release idle lanes of models with nothing
// pending, least recently used first, until it fits.  Called with the lock.
static bool make_room(mct_server *S, int64_t need)
{
    if (S->budget <= 0) return true;
    while (S->used + need > S->budget) {
        lane *victim = nullptr;
        size_t vi = 0;
        for (size_t i = 0; i < S->lanes.size(); i++) {
            lane *L = S->lanes[i];
            if (L->busy || !S->models[L->model].pending.empty()) continue;
            if (!victim || L->last_used < victim->last_used) { victim = L; vi = i; }
        }
        if (!victim) return false;
        S->lanes.erase(S->lanes.begin() + (long)vi);
        S->models[victim->model].nlanes--;
        S->used -= victim->bytes;
        S->releases++;
        lane_free(victim);
    }
    return true;
}

static void dispatcher_main(mct_server *S)
{
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        // complete finished lanes
        bool any_busy = false;
        for (lane *L : S->lanes) {
            if (!L->busy) continue;
            lk.unlock();
            hipError_t q = hipStreamQuery(L->stream);
            lk.lock();
            if (q == hipErrorNotReady) { any_busy = true; continue; }
            const auto now = clk::now();
            for (size_t i = 0; i < L->batch.size(); i++) {
                request *r = L->batch[i];
                r->status = q == hipSuccess ? L->h_status[i] : MCT_ERR_HIP;
                for (int k = 0; k < 3; k++) r->out[k] = L->h_out[3 * i + k];
                r->latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
                r->done = true;
            }
            S->evals += (int64_t)L->batch.size();
            L->batch.clear();
            L->busy = false;
            L->last_used = now;
            S->cv_done.notify_all();
        }
        // admit pending requests: free lanes of their model first, then new
        // lanes while the footprint allows
        const int nm = (int)S->models.size();
        bool waiting = false;
        for (int k = 0; k < nm; k++) {
            const int m = (S->rr + k) % nm;
            model &M = S->models[m];
            while (!M.pending.empty()) {
                if ((int)M.pending.size() < M.max_batch && any_busy && S->fill_ms >= 0.0) {
                    // a partial batch waits for more requests while the device is busy
                    const double waited = std::chrono::duration<double, std::milli>(
                                              clk::now() - M.pending.front()->t_submit).count();
                    if (waited < S->fill_ms) { waiting = true; break; }
                }
                lane *L = nullptr;
                for (lane *c : S->lanes)
                    if (c->model == m && !c->busy) { L = c; break; }
                if (!L) {
                    if (M.nlanes >= S->max_lanes || !make_room(S, M.lane_bytes)) { waiting = true; break; }
                    M.nlanes++;                       // reserved while the lane is built
                    S->used += M.lane_bytes;
                    lk.unlock();
                    lane *n = nullptr;
                    int rc = lane_create(M, &n);
                    lk.lock();
                    S->used -= M.lane_bytes;
                    if (rc == MCT_ERR_NOMEM && M.nlanes > 1) {
                        // the device is fuller than the budget says: make do with the
                        // lanes the model has, and stop asking for more
                        M.nlanes--;
                        if (S->max_lanes > M.nlanes) S->max_lanes = M.nlanes;
                        waiting = true;
                        break;
                    }
                    if (rc) {
                        M.nlanes--;
                        S->error = rc;
                        for (request *r : M.pending) { r->status = rc; r->done = true; }
                        M.pending.clear();
                        S->cv_done.notify_all();
                        break;
                    }
                    n->model = m;
                    M.lane_bytes = n->bytes;
                    S->used += n->bytes;
                    if (S->used > S->peak) S->peak = S->used;
                    S->lanes.push_back(n);
                    L = n;
                }
                while (!M.pending.empty() && (int)L->batch.size() < M.max_batch) {
                    L->batch.push_back(M.pending.front());
                    M.pending.pop_front();
                }
                L->busy = true;
                any_busy = true;
                lk.unlock();
                int rc = lane_launch(M, L);
                lk.lock();
                S->batches++;
                if (rc) {
                    for (request *r : L->batch) { r->status = rc; r->done = true; }
                    L->batch.clear();
                    L->busy = false;
                    S->error = rc;
                    S->cv_done.notify_all();
                }
            }
        }
        S->rr = (S->rr + 1) % (nm > 0 ? nm : 1);
        bool pending = false;
        for (model &M : S->models) pending = pending || !M.pending.empty();
        if (S->stop && !pending && !any_busy) return;
        if (any_busy || waiting) {
            // poll the running lanes; new requests are picked up on the next pass
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(200));
            lk.lock();
        } else {
            S->cv_work.wait(lk, [&] {
                if (S->stop) return true;
                for (model &M : S->models)
                    if (!M.pending.empty()) return true;
                return false;
            });
        }
    }
}

static int model_init(model &M, const mct_grid *grid, const mct_surf_setup *s, int max_batch)
{
    if (!grid || !s || max_batch < 1 || s->np < 1 || s->nsrc < 1 || s->nrev < 1 || !s->freqs || !s->src ||
        !s->rev || !s->ttime || !s->raystat || s->max_cells < 1)
        return MCT_ERR_ARG;
    const size_t npair = (size_t)s->nsrc * s->nrev;
    M.grid = *grid;
    M.setup = *s;
    M.max_batch = max_batch;
    M.freqs.assign(s->freqs, s->freqs + s->np);
    M.src.assign(s->src, s->src + 2 * (size_t)s->nsrc);
    M.rev.assign(s->rev, s->rev + 2 * (size_t)s->nrev);
    M.ttime.assign(s->ttime, s->ttime + 3 * npair * s->np);
    M.raystat.assign(s->raystat, s->raystat + 2 * npair * s->np);
    M.setup.freqs = M.freqs.data(); M.setup.src = M.src.data(); M.setup.rev = M.rev.data();
    M.setup.ttime = M.ttime.data(); M.setup.raystat = M.raystat.data();
    return MCT_OK;
}

static void server_free(mct_server *S)
{
    for (auto &kv : S->all) delete kv.second;
    for (lane *L : S->lanes) lane_free(L);
    delete S;
}

extern "C" {

int mct_serve_create_mix(int nmodels, const mct_grid *grids, const mct_surf_setup *setups, const int *max_batch,
                         int max_lanes, int64_t hbm_budget, mct_server **out)
{
    if (!out || nmodels < 1 || !grids || !setups || !max_batch || max_lanes < 1 || max_lanes > 16 || hbm_budget < 0)
        return MCT_ERR_ARG;
    *out = nullptr;
    mct_server *S = new (std::nothrow) mct_server();
    if (!S) return MCT_ERR_NOMEM;
    S->max_lanes = max_lanes;
    S->budget = hbm_budget;
    S->models.resize(nmodels);
    int rc = MCT_OK;
    for (int m = 0; m < nmodels && !rc; m++) rc = model_init(S->models[m], &grids[m], &setups[m], max_batch[m]);
    // one lane per model to start with (its footprint is then known)
    for (int m = 0; m < nmodels && !rc; m++) {
        lane *L = nullptr;
        rc = lane_create(S->models[m], &L);
        if (rc) break;
        L->model = m;
        S->models[m].nlanes = 1;
        S->models[m].lane_bytes = L->bytes;
        S->used += L->bytes;
        S->lanes.push_back(L);
        if (S->budget > 0 && S->used > S->budget) rc = MCT_ERR_NOMEM;
    }
    if (rc) { server_free(S); return rc; }
    S->peak = S->used;
    S->dispatcher = std::thread(dispatcher_main, S);
    *out = S;
    return MCT_OK;
}

int mct_serve_create(const mct_grid *grid, const mct_surf_setup *setup, int max_batch, int nlanes,
                     mct_server **out)
{
    if (!grid || !setup || !out || max_batch < 1 || nlanes < 1 || nlanes > 16) return MCT_ERR_ARG;
    *out = nullptr;
    mct_server *S = new (std::nothrow) mct_server();
    if (!S) return MCT_ERR_NOMEM;
    S->max_lanes = nlanes;
    S->models.resize(1);
    int rc = model_init(S->models[0], grid, setup, max_batch);
    for (int i = 0; i < nlanes && !rc; i++) {
        lane *L = nullptr;
        rc = lane_create(S->models[0], &L);
        if (rc) break;
        L->model = 0;
        S->models[0].nlanes++;
        S->models[0].lane_bytes = L->bytes;
        S->used += L->bytes;
        S->lanes.push_back(L);
    }
    if (rc) { server_free(S); return rc; }
    S->peak = S->used;
    S->dispatcher = std::thread(dispatcher_main, S);
    *out = S;
    return MCT_OK;
}

int mct_serve_submit_to(mct_server *S, int model, const double *points, const double *params, int ncells,
                        const double *snoise0, const double *snoise1, int64_t *ticket)
{
    if (!S || !points || !params || !ticket || model < 0 || model >= (int)S->models.size()) return MCT_ERR_ARG;
    const struct model &M = S->models[model];
    const int np = M.setup.np;
    if (ncells < 1 || ncells > M.setup.max_cells) return MCT_ERR_ARG;
    request *r = new (std::nothrow) request();
    if (!r) return MCT_ERR_NOMEM;
    r->model = model;
    r->ncells = ncells;
    r->points.assign(points, points + 3 * (size_t)ncells);
    r->params.assign(params, params + 3 * (size_t)ncells);
    r->noise.assign(2 * (size_t)np, 0.0);
    if (snoise0) memcpy(r->noise.data(), snoise0, sizeof(double) * np);
    if (snoise1) memcpy(r->noise.data() + np, snoise1, sizeof(double) * np);
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (S->stop) { delete r; return MCT_ERR_ARG; }
        r->ticket = S->next_ticket++;
        r->t_submit = clk::now();
        S->all[r->ticket] = r;
        S->models[model].pending.push_back(r);
    }
    S->cv_work.notify_one();
    *ticket = r->ticket;
    return MCT_OK;
}

int mct_serve_submit(mct_server *S, const double *points, const double *params, int ncells,
                     const double *snoise0, const double *snoise1, int64_t *ticket)
{
    return mct_serve_submit_to(S, 0, points, params, ncells, snoise0, snoise1, ticket);
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

int mct_serve_set_fill(mct_server *S, double fill_ms)
{
    if (!S) return MCT_ERR_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    S->fill_ms = fill_ms;
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

int mct_serve_memory(mct_server *S, int64_t *bytes_in_use, int64_t *peak_bytes, int32_t *lanes,
                     int64_t *lane_releases)
{
    if (!S) return MCT_ERR_ARG;
    std::lock_guard<std::mutex> g(S->mu);
    if (bytes_in_use) *bytes_in_use = S->used;
    if (peak_bytes) *peak_bytes = S->peak;
    if (lanes) *lanes = (int32_t)S->lanes.size();
    if (lane_releases) *lane_releases = S->releases;
    return MCT_OK;
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
    server_free(S);
    return MCT_OK;
}

} // extern "C"
