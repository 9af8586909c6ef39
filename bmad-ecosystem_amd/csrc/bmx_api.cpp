// C-ABI of the tracking library (declared in include/bmx.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "bmx_internal.hpp"

namespace bmx {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

static int hip_fail(hipError_t e, const char *what)
{
    return set_error(BMX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(call, what)                                  \
    do {                                                     \
        hipError_t _e = (call);                              \
        if (_e != hipSuccess) return bmx::hip_fail(_e, what); \
    } while (0)

// Equal kernel arguments: the same device arrays and values (a captured launch holds them).
static bool same_launch(const LatDesc &a, const LatDesc &b)
{
    return a.hdr == b.hdr && a.prm == b.prm && a.n_ele == b.n_ele && a.sp.mc2 == b.sp.mc2 &&
           a.sp.p0c == b.sp.p0c && a.sp.mc2_sq == b.sp.mc2_sq && a.sp.beta0 == b.sp.beta0 &&
           a.sp.me2 == b.sp.me2 && a.sp.lez_f == b.sp.lez_f && a.sp.lez_thr == b.sp.lez_thr &&
           a.sp.rad.scale == b.sp.rad.scale && a.sp.rad.flags == b.sp.rad.flags &&
           a.sp.rad.key0 == b.sp.rad.key0 && a.sp.rad.key1 == b.sp.rad.key1;
}

static bool same_launch(const BeamPtrs &a, const BeamPtrs &b)
{
    return a.x == b.x && a.px == b.px && a.y == b.y && a.py == b.py && a.z == b.z && a.pz == b.pz &&
           a.beta == b.beta && a.state == b.state && a.loss_ele == b.loss_ele &&
           a.loss_turn == b.loss_turn;
}

}  // namespace bmx

using namespace bmx;

extern "C" {

int32_t bmx_version(void) { return BMX_VERSION; }

const char *bmx_last_error(void) { return g_last_error.c_str(); }

int32_t bmx_ctx_create(int32_t device, bmx_ctx **out)
{
    if (!out) return set_error(BMX_ERR_ARG, "bmx_ctx_create: null out");
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) return set_error(BMX_ERR_ARG, "bmx_ctx_create: no such device");
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    bmx_ctx *c = new (std::nothrow) bmx_ctx();
    if (!c) return set_error(BMX_ERR_NOMEM, "bmx_ctx_create: host allocation");
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreateWithFlags");
    }
    // [0] reduction scratch, [1 .. kWorkSlots] pass counter, then the plain path's turn index
    constexpr size_t kWords = 2 + kWorkSlots;
    e = hipMalloc(&c->d_counter, kWords * sizeof(unsigned long long));
    if (e == hipSuccess) {
        c->d_work = c->d_counter + 1;
        c->d_turn = reinterpret_cast<int32_t *>(c->d_counter + 1 + kWorkSlots);
        e = hipMemset(c->d_counter, 0, kWords * sizeof(unsigned long long));
    }
    // the device address of this device's fixed scratch array (bmx_track_turns_noalloc)
    if (e == hipSuccess) e = noalloc_init(device);
    if (e != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        return hip_fail(e, "hipMalloc(counter)");
    }
    *out = c;
    return BMX_OK;
}

int32_t bmx_ctx_destroy(bmx_ctx *ctx)
{
    if (!ctx) return BMX_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ctx->run.release();
    ctx->timer.release();
    if (ctx->plain.exec) hipGraphExecDestroy(ctx->plain.exec);
    for (auto &ev : ctx->events)
        if (ev) hipEventDestroy(ev);
    hipFree(ctx->d_counter);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return BMX_OK;
}

int32_t bmx_ctx_synchronize(bmx_ctx *ctx)
{
    if (!ctx) return set_error(BMX_ERR_ARG, "bmx_ctx_synchronize: null ctx");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return BMX_OK;
}

void *bmx_ctx_stream(bmx_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int32_t bmx_ctx_set_maps(bmx_ctx *ctx, int32_t form)
{
    if (!ctx || (form != BMX_MAPS_FAST && form != BMX_MAPS_REFERENCE))
        return set_error(BMX_ERR_ARG, "bmx_ctx_set_maps: bad argument");
    ctx->fast_maps = form == BMX_MAPS_FAST;
    return BMX_OK;
}

int32_t bmx_ctx_record(bmx_ctx *ctx, int32_t slot)
{
    if (!ctx || slot < 0 || slot >= BMX_N_EVENTS) return set_error(BMX_ERR_ARG, "bmx_ctx_record: bad slot");
    hipSetDevice(ctx->device);
    if (!ctx->events[slot]) HIP_TRY(hipEventCreate(&ctx->events[slot]), "hipEventCreate");
    HIP_TRY(hipEventRecord(ctx->events[slot], ctx->stream), "hipEventRecord");
    return BMX_OK;
}

int32_t bmx_ctx_elapsed_ms(bmx_ctx *ctx, int32_t slot0, int32_t slot1, float *ms)
{
    if (!ctx || !ms || slot0 < 0 || slot1 < 0 || slot0 >= BMX_N_EVENTS || slot1 >= BMX_N_EVENTS ||
        !ctx->events[slot0] || !ctx->events[slot1])
        return set_error(BMX_ERR_ARG, "bmx_ctx_elapsed_ms: slot not recorded");
    hipSetDevice(ctx->device);
    HIP_TRY(hipEventSynchronize(ctx->events[slot1]), "hipEventSynchronize");
    HIP_TRY(hipEventElapsedTime(ms, ctx->events[slot0], ctx->events[slot1]), "hipEventElapsedTime");
    return BMX_OK;
}

int32_t bmx_ctx_take_work(bmx_ctx *ctx, int64_t *passes)
{
    if (!ctx || !passes) return set_error(BMX_ERR_ARG, "bmx_ctx_take_work: null argument");
    hipSetDevice(ctx->device);
    unsigned long long v[kWorkSlots] = {};
    HIP_TRY(hipMemcpyAsync(v, ctx->d_work, sizeof v, hipMemcpyDeviceToHost, ctx->stream), "copy");
    HIP_TRY(hipMemsetAsync(ctx->d_work, 0, sizeof v, ctx->stream), "memset");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "sync");
    unsigned long long sum = 0;
    for (unsigned long long x : v) sum += x;
    *passes = (int64_t)sum;
    return BMX_OK;
}

int32_t bmx_lattice_create(bmx_ctx *ctx, const bmx_ele *eles, int32_t n_ele, double mc2,
                           int32_t charge, bmx_lattice **out)
{
    if (!ctx || !out) return set_error(BMX_ERR_ARG, "bmx_lattice_create: null argument");
    *out = nullptr;
    HostTable t;
    std::string err;
    int rc = compile_table(eles, n_ele, mc2, charge, t, err);
    if (rc != BMX_OK) return set_error(rc, "bmx_lattice_create: " + err);
    return lattice_upload(ctx, std::move(t), out);
}

int32_t bmx_lattice_create_rad(bmx_ctx *ctx, const bmx_ele *eles, int32_t n_ele, double mc2,
                               int32_t charge, const bmx_rad_map *maps,
                               const bmx_rad_config *config, bmx_lattice **out)
{
    if (!ctx || !out || !maps || !config)
        return set_error(BMX_ERR_ARG, "bmx_lattice_create_rad: null argument");
    *out = nullptr;
    HostTable t;
    std::string err;
    int rc = compile_table(eles, n_ele, mc2, charge, t, err, maps, config);
    if (rc != BMX_OK) return set_error(rc, "bmx_lattice_create_rad: " + err);
    return lattice_upload(ctx, std::move(t), out);
}

int32_t bmx_lattice_set_radiation(bmx_lattice *L, const bmx_rad_config *config)
{
    if (!L || !config || config->reserved != 0 || !std::isfinite(config->synch_rad_scale))
        return set_error(BMX_ERR_ARG, "bmx_lattice_set_radiation: bad argument");
    L->host.sp.rad = rad_settings(*config);
    L->desc.sp.rad = L->host.sp.rad;
    hipSetDevice(L->device);
    // ordered after the context's queued work that reads the descriptor
    HIP_TRY(hipMemcpyAsync(L->d_desc, &L->desc, sizeof(LatDesc), hipMemcpyHostToDevice, L->ctx->stream),
            "bmx_lattice_set_radiation: copy");
    HIP_TRY(hipStreamSynchronize(L->ctx->stream), "bmx_lattice_set_radiation: sync");
    return BMX_OK;
}

}  // extern "C"

namespace bmx {

int32_t lattice_upload(bmx_ctx *ctx, HostTable &&table, bmx_lattice **out)
{
    *out = nullptr;
    bmx_lattice *L = new (std::nothrow) bmx_lattice();
    if (!L) return set_error(BMX_ERR_NOMEM, "bmx_lattice_create: host allocation");
    L->host = std::move(table);
    L->ctx = ctx;
    L->device = ctx->device;
    hipSetDevice(ctx->device);
    const size_t nh = L->host.hdr.size(), np = L->host.prm.size();
    hipError_t e = hipMalloc(&L->d_hdr, nh * sizeof(EleHdr));
    if (e == hipSuccess) e = hipMalloc(&L->d_prm, (np ? np : 1) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&L->d_desc, sizeof(LatDesc));
    if (e == hipSuccess) e = hipMemcpy(L->d_hdr, L->host.hdr.data(), nh * sizeof(EleHdr), hipMemcpyHostToDevice);
    if (e == hipSuccess && np) e = hipMemcpy(L->d_prm, L->host.prm.data(), np * sizeof(double), hipMemcpyHostToDevice);
    L->desc.hdr = L->d_hdr;
    L->desc.prm = L->d_prm;
    L->desc.n_ele = (int32_t)nh;
    L->desc.rare_q = L->host.rare_q ? (L->host.rad ? 2 : 1) : 0;
    L->desc.sp = L->host.sp;
    if (e == hipSuccess) e = hipMemcpy(L->d_desc, &L->desc, sizeof(LatDesc), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(L->d_hdr);
        hipFree(L->d_prm);
        hipFree(L->d_desc);
        delete L;
        return hip_fail(e, "bmx_lattice_create: device table");
    }
    *out = L;
    return BMX_OK;
}

}  // namespace bmx

extern "C" {

int32_t bmx_lattice_destroy(bmx_lattice *L)
{
    if (!L) return BMX_OK;
    // Independent of the context's lifetime: synchronise the device, then free.
    hipSetDevice(L->device);
    hipDeviceSynchronize();
    hipFree(L->d_hdr);
    hipFree(L->d_prm);
    hipFree(L->d_desc);
    delete L;
    return BMX_OK;
}

int32_t bmx_lattice_validate(const bmx_ele *eles, int32_t n_ele, double mc2, int32_t charge,
                             int32_t *n_params)
{
    HostTable t;
    std::string err;
    int rc = compile_table(eles, n_ele, mc2, charge, t, err);
    if (rc != BMX_OK) return set_error(rc, "bmx_lattice_validate: " + err);
    if (n_params) *n_params = (int32_t)t.prm.size();
    return BMX_OK;
}

int32_t bmx_lattice_specialize(bmx_lattice *L)
{
    if (!L) return set_error(BMX_ERR_ARG, "bmx_lattice_specialize: null lattice");
    std::string err;
    const JitKernels *j = nullptr;
    const int rc = jit_specialize(L->device, L->host, &j, err);
    if (rc != BMX_OK) return set_error(rc, "bmx_lattice_specialize: " + err);
    L->jit = j;
    return BMX_OK;
}

int32_t bmx_lattice_info(const bmx_lattice *L, int32_t *n_ele, uint64_t *kind_mask)
{
    if (!L) return set_error(BMX_ERR_ARG, "bmx_lattice_info: null lattice");
    if (n_ele) *n_ele = (int32_t)L->host.hdr.size();
    if (kind_mask) *kind_mask = L->host.kind_mask;
    return BMX_OK;
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

int32_t bmx_beam_create(bmx_ctx *ctx, int64_t n, bmx_beam **out)
{
    if (!ctx || !out || n < 0) return set_error(BMX_ERR_ARG, "bmx_beam_create: bad argument");
    *out = nullptr;
    hipSetDevice(ctx->device);
    const size_t nd = align256((size_t)(n ? n : 1) * 8), ni = align256((size_t)(n ? n : 1) * 4);
    char *blk = nullptr;
    hipError_t e = hipMalloc(&blk, 7 * nd + 3 * ni);
    if (e != hipSuccess) return set_error(BMX_ERR_NOMEM, std::string("bmx_beam_create: ") + hipGetErrorString(e));
    bmx_beam *B = new (std::nothrow) bmx_beam();
    if (!B) {
        hipFree(blk);
        return set_error(BMX_ERR_NOMEM, "bmx_beam_create: host allocation");
    }
    B->ctx = ctx;
    B->device = ctx->device;
    B->n = n;
    B->owned = true;
    B->block = blk;
    double **dp[7] = {&B->d.x, &B->d.px, &B->d.y, &B->d.py, &B->d.z, &B->d.pz, &B->d.beta};
    for (int k = 0; k < 7; k++) *dp[k] = (double *)(blk + k * nd);
    B->d.state = (int32_t *)(blk + 7 * nd);
    B->d.loss_ele = (int32_t *)(blk + 7 * nd + ni);
    B->d.loss_turn = (int32_t *)(blk + 7 * nd + 2 * ni);
    *out = B;
    return BMX_OK;
}

int32_t bmx_beam_wrap(bmx_ctx *ctx, const bmx_beam_view *dev, int64_t n, bmx_beam **out)
{
    if (!ctx || !dev || !out || n < 0) return set_error(BMX_ERR_ARG, "bmx_beam_wrap: bad argument");
    if (n > 0 && (!dev->x || !dev->px || !dev->y || !dev->py || !dev->z || !dev->pz || !dev->beta ||
                  !dev->state || !dev->loss_ele || !dev->loss_turn))
        return set_error(BMX_ERR_ARG, "bmx_beam_wrap: null array");
    bmx_beam *B = new (std::nothrow) bmx_beam();
    if (!B) return set_error(BMX_ERR_NOMEM, "bmx_beam_wrap: host allocation");
    B->ctx = ctx;
    B->device = ctx->device;
    B->n = n;
    B->owned = false;
    B->d = BeamPtrs{dev->x, dev->px, dev->y, dev->py, dev->z, dev->pz, dev->beta,
                    dev->state, dev->loss_ele, dev->loss_turn};
    *out = B;
    return BMX_OK;
}

static int beam_copy(bmx_beam *B, const bmx_beam_view *h, int64_t n, bool up)
{
    if (!B || !h) return set_error(BMX_ERR_ARG, "beam copy: null argument");
    if (n < 0 || n > B->n) return set_error(BMX_ERR_ARG, "beam copy: n exceeds the beam size");
    if (n == 0) return BMX_OK;
    hipSetDevice(B->ctx->device);
    void *hp[10] = {h->x, h->px, h->y, h->py, h->z, h->pz, h->beta, h->state, h->loss_ele, h->loss_turn};
    void *dp[10] = {B->d.x, B->d.px, B->d.y, B->d.py, B->d.z, B->d.pz, B->d.beta, B->d.state,
                    B->d.loss_ele, B->d.loss_turn};
    for (int k = 0; k < 10; k++) {
        if (!hp[k]) return set_error(BMX_ERR_ARG, "beam copy: null host array");
        const size_t bytes = (size_t)n * (k < 7 ? 8 : 4);
        hipError_t e = up ? hipMemcpyAsync(dp[k], hp[k], bytes, hipMemcpyHostToDevice, B->ctx->stream)
                          : hipMemcpyAsync(hp[k], dp[k], bytes, hipMemcpyDeviceToHost, B->ctx->stream);
        if (e != hipSuccess) return hip_fail(e, "beam copy");
    }
    HIP_TRY(hipStreamSynchronize(B->ctx->stream), "beam copy sync");
    return BMX_OK;
}

int32_t bmx_beam_upload(bmx_beam *B, const bmx_beam_view *host, int64_t n) { return beam_copy(B, host, n, true); }
int32_t bmx_beam_download(bmx_beam *B, const bmx_beam_view *host, int64_t n) { return beam_copy(B, host, n, false); }

int32_t bmx_beam_destroy(bmx_beam *B)
{
    if (!B) return BMX_OK;
    if (B->owned) {
        hipSetDevice(B->device);
        hipDeviceSynchronize();
        hipFree(B->block);
    }
    delete B;
    return BMX_OK;
}

int32_t bmx_beam_count_alive(bmx_beam *B, int64_t *n_alive)
{
    if (!B || !n_alive) return set_error(BMX_ERR_ARG, "bmx_beam_count_alive: null argument");
    bmx_ctx *c = B->ctx;
    hipSetDevice(c->device);
    HIP_TRY(hipMemsetAsync(c->d_counter, 0, sizeof(unsigned long long), c->stream), "memset");
    HIP_TRY(launch_count_alive(c->stream, B->d.state, B->n, c->d_counter), "k_count_alive");
    unsigned long long v = 0;
    HIP_TRY(hipMemcpyAsync(&v, c->d_counter, sizeof v, hipMemcpyDeviceToHost, c->stream), "copy");
    HIP_TRY(hipStreamSynchronize(c->stream), "sync");
    *n_alive = (int64_t)v;
    return BMX_OK;
}

}  // extern "C"

namespace bmx {

// Persistent tracking of one group: n_turns turns of `count` elements from element `first`
// (0-based), turn indices turn0 ... (bmx_table.hpp Group). fixed: run buffers of a fixed size
// to use instead of the context's (which grow by allocation); the caller has checked that the
// beam fits them.
int32_t track_group(bmx_ctx *ctx, const bmx_lattice *L, bmx_beam *B, int32_t first, int32_t count,
                    int32_t n_turns, int32_t turn0, const char *who, RunBuffers *fixed)
{
    const std::string w(who);
    if (!ctx || !L || !B) return set_error(BMX_ERR_ARG, w + ": null argument");
    if (n_turns < 0) return set_error(BMX_ERR_ARG, w + ": negative turn count");
    if (L->ctx->device != ctx->device || B->ctx->device != ctx->device)
        return set_error(BMX_ERR_ARG, w + ": handles on different devices");
    if (B->n >= (int64_t)UINT32_MAX) return set_error(BMX_ERR_ARG, w + ": more than 2^32-1 particles");
    if (n_turns == 0 || B->n == 0 || count == 0) return BMX_OK;
    hipSetDevice(ctx->device);
    // the previous call's group upload must have been consumed before the host copy changes
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    RunBuffers *rb = fixed;
    if (!rb) {
        rb = &ctx->run;
        rb->device = ctx->device;
        HIP_TRY(rb->reserve(1, (size_t)B->n, (size_t)(B->n + 63) / 64), "run buffers");
    }
    Group &g = ctx->group;
    g.lat = L->desc;
    g.b = B->d;
    g.n = (uint32_t)B->n;
    g.list_off = 0;
    g.n_turns = n_turns;
    g.turn0 = turn0;
    g.first = first;
    g.count = count;
    RunConfig cfg;
    cfg.chunk_turns = ctx->chunk_turns;
    cfg.fast = ctx->fast_maps;
    cfg.jit = ctx->fast_maps ? L->jit : nullptr;
    cfg.waves = ctx->waves ? ctx->waves : default_waves(ctx->fast_maps, cfg.jit);
    // a start element or a partial turn runs on the generic span kernel; a lattice with stepped
    // quadrupoles or general bends that no cell kernel carries on the generic kernels built
    // with them
    cfg.span = first != 0 || count != L->desc.n_ele;
    const int ext = L->desc.rare_q && (cfg.span || jit_level(cfg.jit) < L->desc.rare_q) ? L->desc.rare_q : 0;
    // Run phase (waves chosen automatically, full turns on the cell kernels): while the live
    // particles fill more than kPhaseRounds rounds of the 4-wave kernel's resident waves, the
    // 5-wave build (more resident waves, a few spilled registers) is faster; below that, with
    // the beam thinned by losses, the 4-wave build is. Measured per 100-turn step of the
    // 901-element ring: 1e6, 5.7e5 and 4.2e5 live particles (3.8, 2.2, 1.6 rounds) run 3-6%
    // faster on 5 waves, 3.4e5 and fewer (1.3 rounds) 1-4% slower.
    if (!ctx->waves && cfg.jit && !cfg.span && !L->host.rare_q) {
        HIP_TRY(hipMemsetAsync(ctx->d_counter, 0, sizeof(unsigned long long), ctx->stream), "memset");
        HIP_TRY(launch_count_alive(ctx->stream, B->d.state, B->n, ctx->d_counter), "k_count_alive");
        unsigned long long alive = 0;
        HIP_TRY(hipMemcpyAsync(&alive, ctx->d_counter, sizeof alive, hipMemcpyDeviceToHost, ctx->stream), "copy");
        HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        const double resident4 = (double)default_grid(ctx->device, true, 4, cfg.jit) * (kTileBlock / 64);
        cfg.waves = (double)((alive + 63) / 64) >= kPhaseRounds * resident4 ? 5 : 4;
    }
    cfg.grid = ctx->grid > 0 ? ctx->grid
                             : default_grid(ctx->device, ctx->fast_maps, cfg.waves, cfg.jit, cfg.span, ext);
    cfg.timer = ctx->profile ? &ctx->timer : nullptr;
    ctx->last = {cfg.chunk_turns, cfg.grid, cfg.waves};
    HIP_TRY(enqueue_run(ctx->stream, *rb, &g, 1, true, n_turns, cfg, ctx->d_work), "k_track_tiles");
    return BMX_OK;
}

}  // namespace bmx

extern "C" {

int32_t bmx_track_turns(bmx_ctx *ctx, const bmx_lattice *L, bmx_beam *B, int32_t n_turns, int32_t turn0)
{
    return track_group(ctx, L, B, 0, L ? L->desc.n_ele : 0, n_turns, turn0, "bmx_track_turns", nullptr);
}

int32_t bmx_track_turns_from(bmx_ctx *ctx, const bmx_lattice *L, bmx_beam *B, int32_t ix_ele_start,
                             int32_t n_turns, int32_t turn0)
{
    if (L && (ix_ele_start < 0 || ix_ele_start > L->desc.n_ele))
        return set_error(BMX_ERR_ARG, "bmx_track_turns_from: ix_ele_start outside 0..n_ele");
    const int32_t n = L ? L->desc.n_ele : 0;
    return track_group(ctx, L, B, n ? ix_ele_start % n : 0, n, n_turns, turn0, "bmx_track_turns_from",
                       nullptr);
}

int32_t bmx_track_bunch(bmx_ctx *ctx, const bmx_lattice *L, bmx_beam *B, int32_t ix_ele1,
                        int32_t ix_ele2, int32_t turn)
{
    if (!L) return set_error(BMX_ERR_ARG, "bmx_track_bunch: null lattice");
    const int32_t n = L->desc.n_ele;
    if (ix_ele1 < 0 || ix_ele1 > n || ix_ele2 < 0 || ix_ele2 > n)
        return set_error(BMX_ERR_ARG, "bmx_track_bunch: element index outside 0..n_ele");
    // beam_mod.f90:148-166: ix_ele1+1 .. ix_ele2, or through the end and round to ix_ele2 when
    // ix_ele2 is at or before ix_ele1
    const int32_t count = ix_ele1 < ix_ele2 ? ix_ele2 - ix_ele1 : n - ix_ele1 + ix_ele2;
    return track_group(ctx, L, B, n ? ix_ele1 % n : 0, count, 1, turn, "bmx_track_bunch", nullptr);
}

int32_t bmx_ctx_profile(bmx_ctx *ctx, int32_t on)
{
    if (!ctx) return set_error(BMX_ERR_ARG, "bmx_ctx_profile: null ctx");
    ctx->profile = on != 0;
    return BMX_OK;
}

int32_t bmx_ctx_kernel_time(bmx_ctx *ctx, double *ms, int64_t *launches)
{
    if (!ctx || !ms || !launches) return set_error(BMX_ERR_ARG, "bmx_ctx_kernel_time: null argument");
    hipSetDevice(ctx->device);
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    double total = 0;
    for (auto &p : ctx->timer.pending) {
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, p.first, p.second), "hipEventElapsedTime");
        total += t;
        ctx->timer.pool.push_back(p.first);
        ctx->timer.pool.push_back(p.second);
    }
    *launches = (int64_t)ctx->timer.pending.size();
    ctx->timer.pending.clear();
    *ms = total;
    return BMX_OK;
}

int32_t bmx_ctx_set_run(bmx_ctx *ctx, int32_t chunk_turns, int32_t grid, int32_t waves)
{
    if (!ctx || chunk_turns < 0 || grid < 0 || (waves != 0 && (waves < 3 || waves > 5)))
        return set_error(BMX_ERR_ARG, "bmx_ctx_set_run: bad argument");
    ctx->chunk_turns = chunk_turns;
    ctx->grid = grid;
    ctx->waves = waves;
    return BMX_OK;
}

int32_t bmx_ctx_last_run(const bmx_ctx *ctx, int32_t *chunk_turns, int32_t *grid, int32_t *waves)
{
    if (!ctx) return set_error(BMX_ERR_ARG, "bmx_ctx_last_run: null context");
    if (chunk_turns) *chunk_turns = ctx->last.chunk_turns;
    if (grid) *grid = ctx->last.grid;
    if (waves) *waves = ctx->last.waves;
    return BMX_OK;
}

int32_t bmx_track_element(bmx_ctx *ctx, const bmx_lattice *L, bmx_beam *B, int32_t ix_ele, int32_t turn)
{
    if (!ctx || !L || !B) return set_error(BMX_ERR_ARG, "bmx_track_element: null argument");
    if (ix_ele < 1 || ix_ele > L->desc.n_ele) return set_error(BMX_ERR_ARG, "bmx_track_element: ix_ele out of range");
    hipSetDevice(ctx->device);
    HIP_TRY(launch_track_element(ctx->stream, L->desc, ix_ele - 1, turn, B->d, B->n, ctx->d_work,
                                 ctx->fast_maps), "k_track_element");
    return BMX_OK;
}

int32_t bmx_track_turns_plain(bmx_ctx *ctx, const bmx_lattice *L, bmx_beam *B, int32_t n_turns, int32_t turn0)
{
    if (!ctx || !L || !B) return set_error(BMX_ERR_ARG, "bmx_track_turns_plain: null argument");
    if (n_turns < 0) return set_error(BMX_ERR_ARG, "bmx_track_turns_plain: negative turn count");
    hipSetDevice(ctx->device);
    if (n_turns == 0 || B->n == 0) return BMX_OK;
    // One turn of element launches is captured in a graph (kept in the context for the same
    // lattice, beam and maps) and replayed once per turn; the launches read the turn index from
    // device memory, which the graph's last node advances.
    bmx_ctx::PlainGraph &pg = ctx->plain;
    if (!pg.exec || !same_launch(pg.lat, L->desc) || !same_launch(pg.beam, B->d) || pg.n != B->n ||
        pg.fast != ctx->fast_maps) {
        if (pg.exec) {
            // replays of the previous graph may still be queued
            HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
            hipGraphExecDestroy(pg.exec);
            pg.exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipError_t e = hipSuccess;
        for (int32_t ie = 0; ie < L->desc.n_ele && e == hipSuccess; ie++)
            e = launch_track_element_at(ctx->stream, L->desc, ie, ctx->d_turn, B->d, B->n, ctx->d_work,
                                        ctx->fast_maps);
        if (e == hipSuccess) e = launch_step_turn(ctx->stream, ctx->d_turn, false, 0);
        const hipError_t e_end = hipStreamEndCapture(ctx->stream, &graph);
        if (e == hipSuccess) e = e_end;
        if (e == hipSuccess) e = hipGraphInstantiate(&pg.exec, graph, nullptr, nullptr, 0);
        if (graph) hipGraphDestroy(graph);
        if (e != hipSuccess) {
            pg.exec = nullptr;
            return hip_fail(e, "bmx_track_turns_plain: turn graph");
        }
        pg.lat = L->desc;
        pg.beam = B->d;
        pg.n = B->n;
        pg.fast = ctx->fast_maps;
    }
    HIP_TRY(launch_step_turn(ctx->stream, ctx->d_turn, true, turn0), "k_step_turn");
    for (int32_t t = 0; t < n_turns; t++) HIP_TRY(hipGraphLaunch(pg.exec, ctx->stream), "hipGraphLaunch");
    return BMX_OK;
}

int32_t bmx_track_orbit(bmx_ctx *ctx, const bmx_lattice *L, bmx_beam *B, int32_t ix_start,
                        int32_t ix_end, int32_t turn, double *orbit)
{
    if (!ctx || !L || !B || !orbit) return set_error(BMX_ERR_ARG, "bmx_track_orbit: null argument");
    if (ix_start < 1 || ix_end > L->desc.n_ele || ix_start > ix_end)
        return set_error(BMX_ERR_ARG, "bmx_track_orbit: element range out of bounds");
    hipSetDevice(ctx->device);
    const size_t bytes = (size_t)B->n * (size_t)(ix_end - ix_start + 1) * 8 * sizeof(double);
    if (bytes == 0) return BMX_OK;
    double *d_rec = nullptr;
    if (hipMalloc(&d_rec, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(BMX_ERR_NOMEM, "bmx_track_orbit: cannot allocate the orbit record");
    }
    int32_t st = BMX_OK;
    hipError_t e = hipMemsetAsync(d_rec, 0, bytes, ctx->stream);
    if (e == hipSuccess)
        e = launch_track_range(ctx->stream, L->desc, ix_start - 1, ix_end - 1, turn, B->d, B->n,
                               d_rec, ctx->d_work, ctx->fast_maps);
    if (e == hipSuccess) e = hipMemcpyAsync(orbit, d_rec, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) st = set_error(BMX_ERR_HIP, std::string("bmx_track_orbit: ") + hipGetErrorString(e));
    hipFree(d_rec);
    return st;
}

}  // extern "C"
