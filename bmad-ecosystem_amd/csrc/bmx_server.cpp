// Tracking server: continuous batching of requests on one device.
//
// The reference's long-term tracking driver runs one beam at a time (lt_tracking_mod.f90:
// 1571-1716, ltt_run_beam_mode); a serving process instead keeps many requests' particles
// resident in a device pool and advances all active requests together, chunk by chunk, with
// the same tile run (enqueue_run) that batches lattice variants. Admission is FIFO by pool
// space; the pool is a first-fit free list over particle slots of one SoA block.
//
// The worker thread keeps one chunk in flight on the compute stream. While it runs, requests
// that arrive are admitted and uploaded on a second stream (same priority), and the results of
// the requests the previous chunk finished are downloaded there; the next chunk starts as soon
// as the running one ends, waiting only for the uploads of the requests it adds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bmx_internal.hpp"

using namespace bmx;

namespace {

// SoA state + the run's two packed particle buffers
constexpr int64_t kBytesPerParticle = 7 * 8 + 3 * 4 + 2 * (int64_t)kPackedBytes;

struct Request {
    int64_t id = 0;
    const bmx_lattice *lat = nullptr;
    int64_t n = 0;
    int32_t n_turns = 0, turn0 = 0, done_turns = 0, chunks = 0;
    bmx_beam_view in{}, out{};
    int64_t off = -1;  // slot offset in the pool
    int64_t t_submit = 0, t_admit = 0, t_done = 0;
    bool done = false;
    int32_t status = BMX_OK;
};

}  // namespace

struct bmx_server {
    bmx_ctx *ctx = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;       // chunks (tracking kernels)
    hipStream_t copy_stream = nullptr;  // request uploads and downloads, overlapping the chunks
    hipEvent_t ev_chunk = nullptr;      // end of the chunk in flight
    hipEvent_t ev_up = nullptr;         // end of the uploads the next chunk waits for
    unsigned long long *d_work = nullptr;
    unsigned long long *h_slots = nullptr;  // pinned: the last chunk's pass counter
    int64_t capacity = 0;   // particle slots
    char *pool = nullptr;
    BeamPtrs base{};
    std::map<int64_t, int64_t> free_list;  // offset -> length
    RunBuffers rb;
    RunConfig cfg;          // run shape of the generic kernel
    bool waves_auto = true;  // waves per SIMD follow the kernel in use
    int32_t chunk_turns = 100;
    std::chrono::steady_clock::time_point t0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::vector<Request *> active;
    std::unordered_map<int64_t, Request *> all;
    int64_t next_id = 1;
    bool stop = false;
    bool arrivals = false;  // a request was submitted since the worker last tried admission
    int64_t total_chunks = 0, busy_ns = 0;
    unsigned long long total_passes = 0;
    std::thread worker;

    int64_t now_ns() const
    {
        return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
    }

    int64_t alloc(int64_t n)
    {
        for (auto it = free_list.begin(); it != free_list.end(); ++it) {
            if (it->second >= n) {
                const int64_t off = it->first, len = it->second;
                free_list.erase(it);
                if (len > n) free_list[off + n] = len - n;
                return off;
            }
        }
        return -1;
    }

    void release(int64_t off, int64_t n)
    {
        auto it = free_list.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_list.end() && off + n == nx->first) {
            it->second += nx->second;
            free_list.erase(nx);
        }
        if (it != free_list.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_list.erase(it);
            }
        }
    }

    BeamPtrs at(int64_t off) const
    {
        return BeamPtrs{base.x + off, base.px + off, base.y + off, base.py + off, base.z + off, base.pz + off,
                        base.beta + off, base.state + off, base.loss_ele + off, base.loss_turn + off};
    }

    hipError_t copy(Request *r, bool up, hipStream_t s)
    {
        const BeamPtrs d = at(r->off);
        const bmx_beam_view &h = up ? r->in : r->out;
        void *hp[10] = {h.x, h.px, h.y, h.py, h.z, h.pz, h.beta, h.state, h.loss_ele, h.loss_turn};
        void *dp[10] = {d.x, d.px, d.y, d.py, d.z, d.pz, d.beta, d.state, d.loss_ele, d.loss_turn};
        hipError_t e = hipSuccess;
        for (int k = 0; k < 10 && e == hipSuccess; k++) {
            const size_t bytes = (size_t)r->n * (k < 7 ? 8 : 4);
            e = up ? hipMemcpyAsync(dp[k], hp[k], bytes, hipMemcpyHostToDevice, s)
                   : hipMemcpyAsync(hp[k], dp[k], bytes, hipMemcpyDeviceToHost, s);
        }
        return e;
    }

    void finish(Request *r, int32_t status)
    {
        r->status = status;
        r->done = true;
        r->t_done = now_ns();
        if (r->off >= 0) release(r->off, r->n);
        r->off = -1;
    }

    void run();
};

void bmx_server::run()
{
    hipSetDevice(device);
    std::vector<Group> groups;
    std::vector<Request *> running;   // requests of the chunk in flight on the compute stream
    std::vector<Request *> ready;     // admitted, uploads queued on the copy stream: next chunk
    std::vector<Request *> copied;    // results queued for download on the copy stream
    bool in_flight = false;           // a chunk is running (ev_chunk not yet reached)
    bool uploads_queued = false;      // ev_up recorded after uploads the next chunk must wait for
    int64_t t_launch = 0;
    hipError_t e = hipSuccess;
    auto fail_all = [&](std::unique_lock<std::mutex> &) {
        for (Request *r : active) finish(r, BMX_ERR_HIP);
        for (Request *r : copied)
            if (!r->done) finish(r, BMX_ERR_HIP);
        active.clear();
        running.clear();
        ready.clear();
        copied.clear();
        in_flight = false;
        uploads_queued = false;
        e = hipSuccess;
        (void)hipGetLastError();
    };
    for (;;) {
        std::vector<Request *> admitted;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (!in_flight && ready.empty() && copied.empty()) {
                cv_work.wait(lk, [&] { return stop || !pending.empty(); });
                if (stop && pending.empty()) break;
            }
            // FIFO admission while the head request fits in the pool
            arrivals = false;
            while (!pending.empty()) {
                Request *r = pending.front();
                const int64_t off = r->n > 0 ? alloc(r->n) : 0;
                if (off < 0) break;
                pending.pop_front();
                r->off = r->n > 0 ? off : -1;
                r->t_admit = now_ns();
                admitted.push_back(r);
                if (r->n > 0 && r->n_turns > 0) active.push_back(r);
            }
        }
        // uploads of the admitted requests on the copy stream, while a chunk may be running; a
        // request of no turns returns its particles unchanged
        for (Request *r : admitted) {
            if (r->n == 0) {
                copied.push_back(r);
                continue;
            }
            if (e == hipSuccess) e = copy(r, true, copy_stream);
            if (r->n_turns == 0) {
                if (e == hipSuccess) e = copy(r, false, copy_stream);
                copied.push_back(r);
            } else {
                ready.push_back(r);
            }
        }
        if (!admitted.empty() && e == hipSuccess) {
            e = hipEventRecord(ev_up, copy_stream);
            uploads_queued = true;
        }
        if (in_flight && e == hipSuccess) {
            const hipError_t q = hipEventQuery(ev_chunk);
            if (q == hipErrorNotReady) {
                // the chunk is still running: admit arrivals meanwhile
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait_for(lk, std::chrono::microseconds(200), [&] { return stop || arrivals; });
                continue;
            }
            e = q;
            if (e == hipSuccess) {
                unsigned long long passes = 0;
                for (int k = 0; k < kWorkSlots; k++) passes += h_slots[k];
                std::vector<Request *> next;
                for (Request *r : running) {
                    r->done_turns += std::min(chunk_turns, r->n_turns - r->done_turns);
                    r->chunks++;
                    if (r->done_turns >= r->n_turns) {
                        // results down on the copy stream, behind the uploads queued above
                        if (e == hipSuccess) e = copy(r, false, copy_stream);
                        copied.push_back(r);
                    } else {
                        next.push_back(r);
                    }
                }
                running.swap(next);
                in_flight = false;
                std::lock_guard<std::mutex> lk(mu);
                total_chunks++;
                total_passes += passes;
                busy_ns += now_ns() - t_launch;
            }
        }
        // the next chunk: the requests still running plus the uploaded ones
        if (!in_flight && e == hipSuccess && (!running.empty() || !ready.empty())) {
            running.insert(running.end(), ready.begin(), ready.end());
            ready.clear();
            if (uploads_queued) {
                e = hipStreamWaitEvent(stream, ev_up, 0);
                uploads_queued = false;
            }
            groups.resize(running.size());
            int32_t max_turns = 0;
            size_t tiles = 0;
            for (size_t i = 0; i < running.size(); i++) {
                Request *r = running[i];
                Group &g = groups[i];
                g.lat = r->lat->desc;
                g.b = at(r->off);
                g.n = (uint32_t)r->n;
                g.list_off = (uint32_t)r->off;
                g.first = 0;
                g.count = g.lat.n_ele;
                g.n_turns = std::min(chunk_turns, r->n_turns - r->done_turns);
                g.turn0 = r->turn0 + r->done_turns;
                max_turns = std::max(max_turns, g.n_turns);
                tiles += (size_t)(r->n + 63) / 64;
            }
            // the cell-specialised kernels when every request of this chunk tracks through
            // lattices that share them
            RunConfig run = cfg;
            if (cfg.fast) {
                const JitKernels *j = running[0]->lat->jit;
                for (Request *r : running)
                    if (r->lat->jit != j) j = nullptr;
                if (j) {
                    run.jit = j;
                    if (waves_auto) run.waves = default_waves(true, j);
                    run.grid = default_grid(device, true, run.waves, j);
                }
            }
            t_launch = now_ns();
            if (e == hipSuccess) e = rb.reserve(groups.size(), (size_t)capacity, tiles + 1);
            if (e == hipSuccess)
                e = enqueue_run(stream, rb, groups.data(), (int32_t)groups.size(), true, max_turns, run, d_work);
            // the chunk's pass count into pinned host memory, the counter cleared for the next
            if (e == hipSuccess)
                e = hipMemcpyAsync(h_slots, d_work, kWorkSlots * sizeof(unsigned long long),
                                   hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipMemsetAsync(d_work, 0, kWorkSlots * sizeof(unsigned long long), stream);
            if (e == hipSuccess) e = hipEventRecord(ev_chunk, stream);
            if (e == hipSuccess) in_flight = true;
        }
        // downloads done (they overlap the chunk just launched): results out, pool space back
        if (!copied.empty() && e == hipSuccess) {
            e = hipStreamSynchronize(copy_stream);
            if (e == hipSuccess) {
                std::lock_guard<std::mutex> lk(mu);
                for (Request *r : copied) finish(r, BMX_OK);
                active.erase(std::remove_if(active.begin(), active.end(), [](Request *r) { return r->done; }),
                             active.end());
                copied.clear();
            }
            cv_done.notify_all();
        }
        if (e != hipSuccess) {
            hipStreamSynchronize(stream);
            hipStreamSynchronize(copy_stream);
            std::unique_lock<std::mutex> lk(mu);
            fail_all(lk);
            cv_done.notify_all();
        }
    }
    hipStreamSynchronize(stream);
    hipStreamSynchronize(copy_stream);
}

extern "C" {

int32_t bmx_server_create(bmx_ctx *ctx, const bmx_server_config *cfg, bmx_server **out)
{
    if (!ctx || !cfg || !out || cfg->hbm_budget < kBytesPerParticle || cfg->chunk_turns < 0 || cfg->compact_turns < 0)
        return set_error(BMX_ERR_ARG, "bmx_server_create: bad argument");
    *out = nullptr;
    bmx_server *S = new (std::nothrow) bmx_server();
    if (!S) return set_error(BMX_ERR_NOMEM, "bmx_server_create: host allocation");
    S->ctx = ctx;
    S->device = ctx->device;
    S->t0 = std::chrono::steady_clock::now();
    S->chunk_turns = cfg->chunk_turns ? cfg->chunk_turns : 100;
    S->capacity = std::min<int64_t>(cfg->hbm_budget / kBytesPerParticle, (int64_t)UINT32_MAX - 1);
    hipSetDevice(S->device);
    const size_t nd = (size_t)S->capacity * 8, ni = (size_t)S->capacity * 4;
    hipError_t e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->copy_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&S->ev_chunk, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&S->ev_up, hipEventDisableTiming);
    if (e == hipSuccess) e = hipHostMalloc((void **)&S->h_slots, kWorkSlots * sizeof(unsigned long long), 0);
    if (e == hipSuccess) e = hipMalloc(&S->d_work, kWorkSlots * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(S->d_work, 0, kWorkSlots * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&S->pool, 7 * nd + 3 * ni);
    if (e == hipSuccess) {
        S->rb.device = S->device;
        e = S->rb.reserve(64, (size_t)S->capacity, (size_t)S->capacity / 64 + 64);
    }
    if (e != hipSuccess) {
        S->rb.release();
        hipFree(S->pool);
        hipFree(S->d_work);
        if (S->h_slots) hipHostFree(S->h_slots);
        if (S->ev_chunk) hipEventDestroy(S->ev_chunk);
        if (S->ev_up) hipEventDestroy(S->ev_up);
        if (S->copy_stream) hipStreamDestroy(S->copy_stream);
        if (S->stream) hipStreamDestroy(S->stream);
        delete S;
        return set_error(BMX_ERR_NOMEM, std::string("bmx_server_create: ") + hipGetErrorString(e));
    }
    char *p = S->pool;
    double **dp[7] = {&S->base.x, &S->base.px, &S->base.y, &S->base.py, &S->base.z, &S->base.pz, &S->base.beta};
    for (int k = 0; k < 7; k++) *dp[k] = (double *)(p + k * nd);
    S->base.state = (int32_t *)(p + 7 * nd);
    S->base.loss_ele = (int32_t *)(p + 7 * nd + ni);
    S->base.loss_turn = (int32_t *)(p + 7 * nd + 2 * ni);
    S->free_list[0] = S->capacity;
    S->cfg.chunk_turns = cfg->compact_turns ? cfg->compact_turns : 5;
    S->cfg.fast = ctx->fast_maps;
    S->waves_auto = ctx->waves == 0;
    S->cfg.waves = ctx->waves ? ctx->waves : default_waves(S->cfg.fast);
    S->cfg.grid = default_grid(S->device, S->cfg.fast, S->cfg.waves);
    S->worker = std::thread([S] { S->run(); });
    *out = S;
    return BMX_OK;
}

int32_t bmx_server_destroy(bmx_server *S)
{
    if (!S) return BMX_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    hipSetDevice(S->device);
    hipStreamSynchronize(S->stream);
    hipStreamSynchronize(S->copy_stream);
    for (auto &kv : S->all) delete kv.second;
    S->rb.release();
    hipFree(S->pool);
    hipFree(S->d_work);
    hipHostFree(S->h_slots);
    hipEventDestroy(S->ev_chunk);
    hipEventDestroy(S->ev_up);
    hipStreamDestroy(S->copy_stream);
    hipStreamDestroy(S->stream);
    delete S;
    return BMX_OK;
}

int32_t bmx_server_submit(bmx_server *S, const bmx_lattice *lat, const bmx_beam_view *in,
                          const bmx_beam_view *out, int64_t n, int32_t n_turns, int32_t turn0,
                          int64_t *id)
{
    if (!S || !lat || !in || !out || !id || n < 0 || n_turns < 0)
        return set_error(BMX_ERR_ARG, "bmx_server_submit: bad argument");
    if (lat->device != S->device) return set_error(BMX_ERR_ARG, "bmx_server_submit: lattice on another device");
    if (n > S->capacity)
        return set_error(BMX_ERR_BUDGET, "bmx_server_submit: " + std::to_string(n) +
                                             " particles exceed the pool of " + std::to_string(S->capacity));
    const bmx_beam_view *v[2] = {in, out};
    for (const bmx_beam_view *h : v)
        if (n > 0 && (!h->x || !h->px || !h->y || !h->py || !h->z || !h->pz || !h->beta || !h->state ||
                      !h->loss_ele || !h->loss_turn))
            return set_error(BMX_ERR_ARG, "bmx_server_submit: null host array");
    Request *r = new (std::nothrow) Request();
    if (!r) return set_error(BMX_ERR_NOMEM, "bmx_server_submit: host allocation");
    r->lat = lat;
    r->n = n;
    r->n_turns = n_turns;
    r->turn0 = turn0;
    r->in = *in;
    r->out = *out;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) {
            delete r;
            return set_error(BMX_ERR_STATE, "bmx_server_submit: server is stopping");
        }
        r->id = S->next_id++;
        r->t_submit = S->now_ns();
        S->all[r->id] = r;
        S->pending.push_back(r);
        S->arrivals = true;
    }
    S->cv_work.notify_all();
    *id = r->id;
    return BMX_OK;
}

int32_t bmx_server_poll(bmx_server *S, int64_t id, int32_t *done)
{
    if (!S || !done) return set_error(BMX_ERR_ARG, "bmx_server_poll: bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return set_error(BMX_ERR_NOT_FOUND, "bmx_server_poll: unknown request");
    *done = it->second->done ? 1 : 0;
    return BMX_OK;
}

int32_t bmx_server_wait(bmx_server *S, int64_t id, bmx_request_stats *stats)
{
    if (!S) return set_error(BMX_ERR_ARG, "bmx_server_wait: null server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return set_error(BMX_ERR_NOT_FOUND, "bmx_server_wait: unknown request");
    Request *r = it->second;
    S->cv_done.wait(lk, [&] { return r->done; });
    if (stats) {
        stats->submit_ns = r->t_submit;
        stats->admit_ns = r->t_admit;
        stats->done_ns = r->t_done;
        stats->status = r->status;
        stats->chunks = r->chunks;
    }
    const int32_t st = r->status;
    S->all.erase(it);
    delete r;
    if (st != BMX_OK) return set_error(st, "bmx_server_wait: request failed on the device");
    return BMX_OK;
}

int32_t bmx_server_totals(bmx_server *S, int64_t *passes, int64_t *chunks, int64_t *busy_ns)
{
    if (!S) return set_error(BMX_ERR_ARG, "bmx_server_totals: null server");
    std::lock_guard<std::mutex> lk(S->mu);
    if (passes) *passes = (int64_t)S->total_passes;
    if (chunks) *chunks = S->total_chunks;
    if (busy_ns) *busy_ns = S->busy_ns;
    return BMX_OK;
}

}  // extern "C"
