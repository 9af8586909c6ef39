// Chunked tile runs: the host side of k_track_tiles / k_build_tiles.
//
// A run tracks a set of groups (lattice + particle range + turn range) chunk by chunk. Before
// each chunk k_build_tiles turns the alive counts of the previous chunk into wave tiles, so
// lost particles stop taking lanes; the tracking kernel is a persistent grid sized to the
// device. This replaces the reference's per-element fork/join over all particles with
// per-chunk compaction (beam_utils.f90:81-86 skips dead particles one by one).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "bmx_internal.hpp"

namespace bmx {

void RunBuffers::release()
{
    if (device >= 0) hipSetDevice(device);
    hipFree(d_groups);
    for (int k = 0; k < 2; k++) {
        hipFree(d_packed[k]);
        hipFree(d_tiles[k]);
        hipFree(d_cnt[k]);
        d_packed[k] = nullptr;
        d_tiles[k] = nullptr;
        d_cnt[k] = nullptr;
    }
    hipFree(d_small);
    hipFree(d_goff);
    d_goff = nullptr;
    d_groups = nullptr;
    d_small = nullptr;
    cap_groups = cap_particles = cap_tiles = 0;
}

hipError_t RunBuffers::reserve(size_t n_groups, size_t n_particles, size_t n_tiles)
{
    hipError_t e = hipSuccess;
    if (!d_small) {
        e = hipMalloc(&d_small, 4 * sizeof(uint32_t));
        if (e != hipSuccess) return e;
    }
    if (n_groups > cap_groups) {
        const size_t c = std::max(n_groups, 2 * cap_groups);
        hipFree(d_groups);
        hipFree(d_goff);
        for (int k = 0; k < 2; k++) hipFree(d_cnt[k]);
        d_groups = nullptr;
        d_goff = nullptr;
        d_cnt[0] = d_cnt[1] = nullptr;
        cap_groups = 0;
        e = hipMalloc(&d_groups, c * sizeof(Group));
        if (e == hipSuccess) e = hipMalloc(&d_goff, (c + 1) * sizeof(uint32_t));
        for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipMalloc(&d_cnt[k], c * sizeof(uint32_t));
        if (e != hipSuccess) return e;
        cap_groups = c;
    }
    if (n_particles > cap_particles) {
        // slots are indexed by uint32_t
        if (n_particles > UINT32_MAX) return hipErrorInvalidValue;
        const size_t c = std::min<size_t>(std::max(n_particles, cap_particles + cap_particles / 2),
                                          UINT32_MAX);
        for (int k = 0; k < 2; k++) {
            hipFree(d_packed[k]);
            d_packed[k] = nullptr;
        }
        cap_particles = 0;
        for (int k = 0; k < 2 && e == hipSuccess; k++)
            e = hipMalloc(reinterpret_cast<void **>(&d_packed[k]), c * kPackedBytes);
        if (e != hipSuccess) return e;
        cap_particles = c;
    }
    if (n_tiles > cap_tiles) {
        const size_t c = std::max(n_tiles, cap_tiles + cap_tiles / 2);
        for (int k = 0; k < 2; k++) {
            hipFree(d_tiles[k]);
            d_tiles[k] = nullptr;
        }
        cap_tiles = 0;
        for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipMalloc(&d_tiles[k], c * sizeof(WaveTile));
        if (e != hipSuccess) return e;
        cap_tiles = c;
    }
    return e;
}

hipError_t KernelTimer::take(hipEvent_t *ev)
{
    if (pool.empty()) {
        hipEvent_t e = nullptr;
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
        pool.push_back(e);
    }
    *ev = pool.back();
    pool.pop_back();
    return hipSuccess;
}

void KernelTimer::release()
{
    for (auto &p : pending) {
        pool.push_back(p.first);
        pool.push_back(p.second);
    }
    pending.clear();
    for (hipEvent_t e : pool) hipEventDestroy(e);
    pool.clear();
}

int default_grid(int device, bool fast, int waves, const JitKernels *jit, bool span, int ext)
{
    int blocks = 0, cus = 0;
    const hipError_t e = ext   ? track_tiles_occupancy(fast, kExtWaves, &blocks, span, ext)
                         : span ? track_tiles_occupancy(fast, kSpanWaves, &blocks, true)
                         : jit  ? jit_occupancy(jit, waves, &blocks)
                                : track_tiles_occupancy(fast, waves, &blocks);
    if (e != hipSuccess || blocks <= 0) blocks = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        cus = 1;
    return blocks * cus;
}

hipError_t enqueue_run(hipStream_t s, RunBuffers &rb, const Group *h_groups, int32_t n_groups,
                       bool upload, int32_t max_turns, const RunConfig &cfg,
                       unsigned long long *d_work)
{
    if (n_groups <= 0 || max_turns <= 0) return hipSuccess;
    hipError_t e = hipSuccess;
    if (upload) {
        e = hipMemcpyAsync(rb.d_groups, h_groups, (size_t)n_groups * sizeof(Group),
                           hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
    }
    const int32_t K = cfg.chunk_turns > 0 ? cfg.chunk_turns : max_turns;
    uint32_t *n_tiles[2] = {rb.d_small, rb.d_small + 1};
    uint32_t *tile_ctr = rb.d_small + 2;
    // Chunk 0's particles (packed buffer / count 0) from the beams' states: tiles over every
    // particle of every group, a ballot pass over them packing the live ones, then the chunk's
    // tiles from the live counts; chunk c reads packed buffer c % 2 and appends its survivors to
    // the other one.
    const uint32_t stride = (uint32_t)rb.cap_particles;
    size_t all_tiles = 0;
    for (int32_t g = 0; g < n_groups; g++)
        if (h_groups[g].n_turns > 0) all_tiles += (h_groups[g].n + 63) / 64;
    const int collect_grid = (int)std::min<size_t>((all_tiles + 3) / 4, 4096);
    e = launch_build_tiles(s, rb.d_groups, n_groups, nullptr, 0, rb.d_tiles[1], n_tiles[1],
                           rb.d_cnt[0], tile_ctr, rb.d_goff);
    if (e == hipSuccess)
        e = launch_pack_alive(s, collect_grid, rb.d_groups, rb.d_tiles[1], n_tiles[1],
                              rb.d_packed[0], stride, rb.d_cnt[0]);
    if (e == hipSuccess)
        e = launch_build_tiles(s, rb.d_groups, n_groups, rb.d_cnt[0], 0, rb.d_tiles[0], n_tiles[0],
                               rb.d_cnt[1], tile_ctr, rb.d_goff);
    // Lattices with stepped quadrupoles or general bends (K_QMULT, K_BMULT) run on cell kernels
    // generated with them or on the generic kernels built with them (ext); the other kernels
    // do not carry them.
    int rare_q = 0;  // the highest map level among the groups (LatDesc::rare_q)
    for (int32_t g = 0; g < n_groups; g++)
        if (h_groups[g].n_turns > 0 && h_groups[g].lat.rare_q > rare_q) rare_q = h_groups[g].lat.rare_q;
    const bool use_jit = cfg.jit && !cfg.span && jit_level(cfg.jit) >= rare_q;
    const int ext = use_jit ? 0 : rare_q;
    int c = 0;
    for (int32_t off = 0; off < max_turns && e == hipSuccess; off += K, c++) {
        const int cur = c & 1, nxt = cur ^ 1;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        if (cfg.timer) {
            e = cfg.timer->take(&ev0);
            if (e == hipSuccess) e = cfg.timer->take(&ev1);
            if (e == hipSuccess) e = hipEventRecord(ev0, s);
            if (e != hipSuccess) break;
        }
        if (use_jit)
            e = launch_track_tiles_jit(cfg.jit, s, cfg.grid, cfg.waves, rb.d_groups, rb.d_tiles[cur],
                                       n_tiles[cur], rb.d_packed[cur], rb.d_packed[nxt], stride,
                                       rb.d_cnt[nxt], tile_ctr, off, K, d_work);
        else
            e = launch_track_tiles(s, cfg.grid, cfg.waves, rb.d_groups, rb.d_tiles[cur], n_tiles[cur],
                                   rb.d_packed[cur], rb.d_packed[nxt], stride, rb.d_cnt[nxt],
                                   tile_ctr, off, K, d_work, cfg.fast, cfg.span, ext);
        if (cfg.timer && e == hipSuccess) {
            e = hipEventRecord(ev1, s);
            cfg.timer->pending.emplace_back(ev0, ev1);
        }
        if (e == hipSuccess && off + K < max_turns)
            e = launch_build_tiles(s, rb.d_groups, n_groups, rb.d_cnt[nxt], off + K,
                                   rb.d_tiles[nxt], n_tiles[nxt], rb.d_cnt[cur], tile_ctr, rb.d_goff);
    }
    return e;
}

}  // namespace bmx
