// Batches: many (lattice, particle range) groups tracked together, the chunk sequence captured
// once as a hipGraph per turn count, and an autotuner over the run shape.
#include <hip/hip_runtime.h>

#include <map>
#include <new>
#include <string>
#include <vector>

#include "bmx_internal.hpp"

using namespace bmx;

struct bmx_batch {
    bmx_ctx *ctx = nullptr;
    int device = 0;
    std::vector<Group> groups;        // host copy; n_turns / turn0 set per call
    std::vector<const bmx_lattice *> lats;
    size_t n_particles = 0;
    RunBuffers rb;
    RunConfig cfg;
    std::map<int32_t, hipGraphExec_t> graphs;  // by turn count, for the current cfg
    bool grid_auto = true;                      // grid follows the kernel's occupancy
    bool waves_auto = true;                     // waves per SIMD follow the kernel
    void drop_graphs()
    {
        for (auto &kv : graphs) hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
    // The cell-specialised kernels, when every group's lattice carries the same ones and the
    // fast maps are selected; otherwise the generic kernel.
    const JitKernels *shared_jit() const
    {
        if (!cfg.fast || lats.empty()) return nullptr;
        const JitKernels *j = lats[0]->jit;
        for (const bmx_lattice *L : lats)
            if (L->jit != j) return nullptr;
        return j;
    }
    // Follow the context's map choice and the lattices' specialisation (which may be set after
    // the batch was made); a change invalidates the captured graphs.
    void refresh(bool fast)
    {
        const bool fast_changed = cfg.fast != fast;
        cfg.fast = fast;
        const JitKernels *j = shared_jit();
        if (!fast_changed && j == cfg.jit) return;
        cfg.jit = j;
        if (waves_auto) cfg.waves = default_waves(cfg.fast, cfg.jit);
        if (grid_auto) cfg.grid = default_grid(device, cfg.fast, cfg.waves, cfg.jit);
        drop_graphs();
    }
};

namespace {

BeamPtrs offset(const BeamPtrs &b, int64_t first)
{
    return BeamPtrs{b.x + first, b.px + first, b.y + first, b.py + first, b.z + first, b.pz + first,
                    b.beta + first, b.state + first, b.loss_ele + first, b.loss_turn + first};
}

size_t tile_bound(const std::vector<Group> &g)
{
    size_t t = 0;
    for (const Group &x : g) t += (x.n + 63) / 64;
    return t;
}

int fail(hipError_t e, const char *what)
{
    return set_error(BMX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

extern "C" {

int32_t bmx_batch_create(bmx_ctx *ctx, const bmx_group *groups, int32_t n_groups, bmx_batch **out)
{
    if (!ctx || !out || n_groups < 0 || (n_groups > 0 && !groups))
        return set_error(BMX_ERR_ARG, "bmx_batch_create: bad argument");
    *out = nullptr;
    bmx_batch *B = new (std::nothrow) bmx_batch();
    if (!B) return set_error(BMX_ERR_NOMEM, "bmx_batch_create: host allocation");
    B->ctx = ctx;
    B->device = ctx->device;
    size_t off = 0;
    for (int32_t i = 0; i < n_groups; i++) {
        const bmx_group &g = groups[i];
        if (!g.lat || !g.beam || g.first < 0 || g.count < 0 || g.first + g.count > g.beam->n) {
            delete B;
            return set_error(BMX_ERR_ARG, "bmx_batch_create: group " + std::to_string(i) +
                                              ": null handle or range outside its beam");
        }
        if (g.lat->device != ctx->device || g.beam->device != ctx->device) {
            delete B;
            return set_error(BMX_ERR_ARG, "bmx_batch_create: handles on different devices");
        }
        Group G{};
        G.lat = g.lat->desc;
        G.b = offset(g.beam->d, g.first);
        G.n = (uint32_t)g.count;
        G.list_off = (uint32_t)off;
        G.first = 0;
        G.count = G.lat.n_ele;
        off += (size_t)g.count;
        B->groups.push_back(G);
        B->lats.push_back(g.lat);
    }
    if (off >= UINT32_MAX) {
        delete B;
        return set_error(BMX_ERR_ARG, "bmx_batch_create: more than 2^32-1 particles");
    }
    B->n_particles = off;
    hipSetDevice(ctx->device);
    B->rb.device = ctx->device;
    hipError_t e = B->rb.reserve(B->groups.size() ? B->groups.size() : 1, off ? off : 1,
                                 tile_bound(B->groups) + 1);
    if (e != hipSuccess) {
        B->rb.release();
        delete B;
        return set_error(BMX_ERR_NOMEM, std::string("bmx_batch_create: ") + hipGetErrorString(e));
    }
    B->cfg.chunk_turns = ctx->chunk_turns;
    B->cfg.fast = ctx->fast_maps;
    B->cfg.jit = B->shared_jit();
    B->waves_auto = ctx->waves == 0;
    B->cfg.waves = ctx->waves ? ctx->waves : default_waves(B->cfg.fast, B->cfg.jit);
    B->cfg.grid = default_grid(ctx->device, B->cfg.fast, B->cfg.waves, B->cfg.jit);
    *out = B;
    return BMX_OK;
}

int32_t bmx_batch_destroy(bmx_batch *B)
{
    if (!B) return BMX_OK;
    // Independent of the context's lifetime (like lattices and beams): device-wide sync.
    hipSetDevice(B->device);
    hipDeviceSynchronize();
    B->drop_graphs();
    B->rb.release();
    delete B;
    return BMX_OK;
}

int32_t bmx_batch_set_run(bmx_batch *B, int32_t chunk_turns, int32_t grid, int32_t waves)
{
    if (!B || chunk_turns < 0 || grid < 0 || (waves != 0 && (waves < 3 || waves > 5)))
        return set_error(BMX_ERR_ARG, "bmx_batch_set_run: bad argument");
    hipSetDevice(B->ctx->device);
    hipStreamSynchronize(B->ctx->stream);
    B->cfg.chunk_turns = chunk_turns;
    B->cfg.jit = B->shared_jit();
    B->waves_auto = waves == 0;
    B->cfg.waves = waves ? waves : default_waves(B->cfg.fast, B->cfg.jit);
    B->grid_auto = grid <= 0;
    B->cfg.grid = grid > 0 ? grid : default_grid(B->ctx->device, B->cfg.fast, B->cfg.waves, B->cfg.jit);
    B->drop_graphs();
    return BMX_OK;
}

int32_t bmx_batch_track(bmx_batch *B, int32_t n_turns, int32_t turn0)
{
    if (!B || n_turns < 0) return set_error(BMX_ERR_ARG, "bmx_batch_track: bad argument");
    if (n_turns == 0 || B->groups.empty()) return BMX_OK;
    bmx_ctx *ctx = B->ctx;
    hipSetDevice(ctx->device);
    B->refresh(ctx->fast_maps != 0);
    // the previous call's upload of the host group array must be complete before it changes
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(e, "bmx_batch_track");
    for (Group &g : B->groups) {
        g.n_turns = n_turns;
        g.turn0 = turn0;
    }
    e = hipMemcpyAsync(B->rb.d_groups, B->groups.data(), B->groups.size() * sizeof(Group),
                       hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return fail(e, "bmx_batch_track: group upload");
    auto it = B->graphs.find(n_turns);
    if (it == B->graphs.end()) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return fail(e, "hipStreamBeginCapture");
        hipError_t er = enqueue_run(ctx->stream, B->rb, B->groups.data(), (int32_t)B->groups.size(),
                                    false, n_turns, B->cfg, ctx->d_work);
        e = hipStreamEndCapture(ctx->stream, &graph);
        if (er != hipSuccess || e != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            return fail(er != hipSuccess ? er : e, "bmx_batch_track: capture");
        }
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(e, "hipGraphInstantiate");
        it = B->graphs.emplace(n_turns, exec).first;
    }
    e = hipGraphLaunch(it->second, ctx->stream);
    if (e != hipSuccess) return fail(e, "hipGraphLaunch");
    return BMX_OK;
}

int32_t bmx_batch_tune(bmx_batch *B, int32_t sample_turns, int32_t *chunk_out, int32_t *grid_out,
                       int32_t *waves_out, float *best_ms_out)
{
    if (!B || sample_turns <= 0) return set_error(BMX_ERR_ARG, "bmx_batch_tune: bad argument");
    bmx_ctx *ctx = B->ctx;
    hipSetDevice(ctx->device);
    B->refresh(ctx->fast_maps != 0);
    const size_t n = B->n_particles;
    if (n == 0) return BMX_OK;
    // scratch SoA copy of every group's particles, restored before each candidate
    const size_t nd = n * 8, ni = n * 4;
    char *orig = nullptr, *work = nullptr;
    hipError_t e = hipMalloc(&orig, 7 * nd + 3 * ni);
    if (e == hipSuccess) e = hipMalloc(&work, 7 * nd + 3 * ni);
    if (e != hipSuccess) {
        hipFree(orig);
        return set_error(BMX_ERR_NOMEM, "bmx_batch_tune: scratch copy does not fit");
    }
    auto soa = [&](char *blk, size_t first) {
        BeamPtrs p;
        double **d[7] = {&p.x, &p.px, &p.y, &p.py, &p.z, &p.pz, &p.beta};
        for (int k = 0; k < 7; k++) *d[k] = (double *)(blk + k * nd) + first;
        p.state = (int32_t *)(blk + 7 * nd) + first;
        p.loss_ele = (int32_t *)(blk + 7 * nd + ni) + first;
        p.loss_turn = (int32_t *)(blk + 7 * nd + 2 * ni) + first;
        return p;
    };
    std::vector<Group> scratch = B->groups;
    for (size_t i = 0; i < scratch.size(); i++) {
        const Group &g = B->groups[i];
        const BeamPtrs dst = soa(orig, g.list_off);
        void *dp[10] = {dst.x, dst.px, dst.y, dst.py, dst.z, dst.pz, dst.beta, dst.state, dst.loss_ele, dst.loss_turn};
        void *sp[10] = {g.b.x, g.b.px, g.b.y, g.b.py, g.b.z, g.b.pz, g.b.beta, g.b.state, g.b.loss_ele, g.b.loss_turn};
        for (int k = 0; k < 10 && e == hipSuccess && g.n; k++)
            e = hipMemcpyAsync(dp[k], sp[k], (size_t)g.n * (k < 7 ? 8 : 4), hipMemcpyDeviceToDevice, ctx->stream);
        scratch[i].b = soa(work, g.list_off);
        scratch[i].n_turns = sample_turns;
        scratch[i].turn0 = 0;
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    RunConfig best = B->cfg;
    float best_ms = -1;
    const int32_t chunks[] = {5, 10, 25, 50};
    const int32_t waves[] = {3, 4, 5};
    for (int32_t w : waves) {
        const int full = default_grid(ctx->device, B->cfg.fast, w, B->cfg.jit);
        for (int32_t K : chunks) {
            if (K > 2 * sample_turns && K != chunks[0]) continue;
            for (int gsel = 0; gsel < 2 && e == hipSuccess; gsel++) {
                RunConfig c = B->cfg;
                c.chunk_turns = K;
                c.waves = w;
                c.grid = gsel == 0 ? full : (full * 3) / 4;
                if (c.grid <= 0) continue;
                e = hipMemcpyAsync(work, orig, 7 * nd + 3 * ni, hipMemcpyDeviceToDevice, ctx->stream);
                if (e == hipSuccess) e = hipEventRecord(t0, ctx->stream);
                if (e == hipSuccess)
                    e = enqueue_run(ctx->stream, B->rb, scratch.data(), (int32_t)scratch.size(), true,
                                    sample_turns, c, nullptr);
                if (e == hipSuccess) e = hipEventRecord(t1, ctx->stream);
                if (e == hipSuccess) e = hipEventSynchronize(t1);
                float ms = 0;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
                if (e == hipSuccess && (best_ms < 0 || ms < best_ms)) {
                    best_ms = ms;
                    best = c;
                }
            }
        }
    }
    hipStreamSynchronize(ctx->stream);
    if (t0) hipEventDestroy(t0);
    if (t1) hipEventDestroy(t1);
    hipFree(orig);
    hipFree(work);
    if (e != hipSuccess) return fail(e, "bmx_batch_tune");
    B->cfg = best;
    B->grid_auto = false;
    B->waves_auto = false;
    B->drop_graphs();
    if (chunk_out) *chunk_out = best.chunk_turns;
    if (grid_out) *grid_out = best.grid;
    if (waves_out) *waves_out = best.waves;
    if (best_ms_out) *best_ms_out = best_ms;
    return BMX_OK;
}

}  // extern "C"
