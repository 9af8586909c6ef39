// Tracking kernels for gfx950.
//
//  k_track_tiles   persistent tracking of one chunk of turns: one lane per particle, state in
//                  registers for all elements and turns of the chunk; work comes as wave tiles
//                  of one group each, so single-lattice, batched and served runs share it.
//  k_build_tiles   next chunk's tiles from the alive counts (lost particles drop out).
//  k_pack_alive    a run's live particles packed for its first chunk.
//  k_track_element one element for every particle, state round-trips HBM: the GPU form of
//                  track1_bunch_hom (bmad/multiparticle/beam_utils.f90:32-90).
//  k_count_alive   number of live particles (lt_tracking_mod.f90:1672 bunch%n_live).
#include <hip/hip_runtime.h>

#include "bmx_internal.hpp"
#include "bmx_tiles.hpp"

namespace bmx {

// W: waves per SIMD the register allocation targets at least (3: up to 168 VGPRs; 4: up to 128;
// 5: up to 102, spilling a few values of the rf and bend paths to scratch). SPAN: groups may
// track element spans (runs with a start element or a partial turn); the full-turn kernels
// carry no code for them. RARE: the map set of the rare path; MAPS_RARE_Q (with W = 2) for
// lattices that hold a stepped quadrupole or a general bend, MAPS_RARE_QR for those with
// radiation.
template <bool F, int W, bool SPAN = false, int RARE = MAPS_RARE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W))) k_track_tiles(
    const Group *__restrict__ groups, const WaveTile *__restrict__ tiles,
    const uint32_t *__restrict__ n_tiles_p, const double *__restrict__ pk_in,
    double *__restrict__ pk_out, uint32_t stride, uint32_t *__restrict__ cnt_out,
    uint32_t *tile_ctr, int32_t chunk_off, int32_t chunk_turns, unsigned long long *work)
{
    track_tiles<F, GenericTurn<F, RARE>, SPAN>(groups, tiles, n_tiles_p, pk_in, pk_out, stride,
                                               cnt_out, tile_ctr, chunk_off, chunk_turns, work);
}

// Builds the wave tiles of the next chunk from the alive counts per group (cnt_in; nullptr
// for the first chunk: every particle of each group, indexed directly) and resets the counters
// the next chunk appends to. One workgroup of kBuildThreads threads: a scan of the per-group
// tile counts (groups that have no turns left after next_off get none) into goff[0..n_groups],
// then all threads write the tiles, each finding its group by binary search in goff.
constexpr int kBuildThreads = 1024;
__global__ void __launch_bounds__(kBuildThreads) k_build_tiles(
    const Group *__restrict__ groups, int32_t n_groups, const uint32_t *__restrict__ cnt_in,
    int32_t next_off, WaveTile *__restrict__ tiles, uint32_t *__restrict__ n_tiles_p,
    uint32_t *__restrict__ cnt_zero, uint32_t *__restrict__ tile_ctr, uint32_t *__restrict__ goff)
{
    __shared__ uint32_t part[kBuildThreads];
    const int tid = threadIdx.x;
    // contiguous run of groups per thread, so tiles come out in group order
    const int per = (n_groups + kBuildThreads - 1) / kBuildThreads;
    const int g0 = tid * per, g1 = min(n_groups, g0 + per);
    uint32_t mine = 0;
    for (int g = g0; g < g1; g++) {
        const uint32_t c = cnt_in ? cnt_in[g] : groups[g].n;
        if (groups[g].n_turns > next_off) mine += (c + 63) / 64;
    }
    part[tid] = mine;
    __syncthreads();
    // inclusive scan (Hillis-Steele)
    for (int o = 1; o < kBuildThreads; o <<= 1) {
        const uint32_t v = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t pos = part[tid] - mine;
    for (int g = g0; g < g1; g++) {
        goff[g] = pos;
        const uint32_t c = cnt_in ? cnt_in[g] : groups[g].n;
        if (groups[g].n_turns > next_off) pos += (c + 63) / 64;
    }
    const uint32_t total = part[kBuildThreads - 1];
    if (tid == 0) goff[n_groups] = total;
    __syncthreads();
    for (uint32_t t = tid; t < total; t += kBuildThreads) {
        // last g with goff[g] <= t (groups without tiles have goff[g] == goff[g + 1])
        int lo = 0, hi = n_groups - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (goff[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t c = cnt_in ? cnt_in[lo] : groups[lo].n;
        const uint32_t s0 = (t - goff[lo]) * 64;
        WaveTile w;
        w.group = lo;
        w.start = s0;
        w.count = c - s0 < 64 ? c - s0 : 64;
        w.pad = 0;
        tiles[t] = w;
    }
    __syncthreads();
    for (int g = tid; g < n_groups; g += kBuildThreads) cnt_zero[g] = 0;
    if (tid == 0) {
        *n_tiles_p = total;
        *tile_ctr = 0;
    }
}

// The live particles of a run packed for its first chunk: one wave per tile of every group's
// particles (tiles built from the group sizes), a ballot of state == alive and one atomic per
// wave appending the live particles' coordinates and indices to the group's slots of the packed
// buffer (counts in cnt_out, zeroed by the tile build). Dead particles of the caller's beam thus
// take no lane in the run, as track1_bunch_hom skips them (beam_utils.f90:83). Grid-stride over
// the tiles.
__global__ void __launch_bounds__(kBlock) k_pack_alive(
    const Group *__restrict__ groups, const WaveTile *__restrict__ tiles,
    const uint32_t *__restrict__ n_tiles_p, double *__restrict__ pk_out, uint32_t stride,
    uint32_t *__restrict__ cnt_out)
{
    const uint32_t nt = *n_tiles_p;
    const int lane = threadIdx.x & 63;
    const uint32_t n_waves = gridDim.x * (kBlock / 64);
    for (uint32_t t = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); t < nt; t += n_waves) {
        const WaveTile T = tiles[t];
        const Group &G = groups[T.group];
        const uint32_t idx = T.start + lane;
        const bool alive = (uint32_t)lane < T.count && G.b.state[idx] == ST_ALIVE;
        const unsigned long long m = __ballot(alive);
        if (!m) continue;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&cnt_out[T.group], (uint32_t)__popcll(m));
        base = __builtin_amdgcn_readfirstlane(base);
        if (alive) {
            Particle p;
            load_particle(p, G.b, idx);
            store_packed(p, pk_out, stride,
                         G.list_off + base + (uint32_t)__popcll(m & ((1ull << lane) - 1)), idx);
        }
    }
}

// One element, whole bunch: track1 (track1.f90:44-343) per live particle. Q: the kernel carries
// the stepped quadrupole and the general bend (built for lattices that hold one).
template <bool F, int Q>
__device__ __forceinline__ void track_element_pass(const LatDesc &L, int32_t ie, int32_t turn,
                                                   const BeamPtrs &b, int64_t n,
                                                   unsigned long long *work)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = i < n && b.state[i] == ST_ALIVE;
    if (active) {
        Particle p;
        load_particle(p, b, i);
        refresh_energy(p, L.sp);
        if (!orbit_too_large(p))
            track_element<F, Q == 2 ? MAPS_ALL_QR : Q ? MAPS_ALL_Q : MAPS_ALL>(p, load_ele((CHdr)L.hdr + 4 * ie), (CP)L.prm, L.sp,
                                                      false, ie + 1, turn);
        store_particle(p, b, i);
        if (p.state != ST_ALIVE) {
            b.state[i] = p.state;
            b.loss_ele[i] = ie + 1;
            b.loss_turn[i] = turn;
        }
    }
    // one atomic per workgroup: the waves' counts summed in LDS
    if (!work) return;
    unsigned long long s = active ? 1ull : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ unsigned long long part[kBlock / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kBlock / 64; w++) t += part[w];
        if (t) atomicAdd(work + blockIdx.x % kWorkSlots, t);
    }
}

template <bool F, int Q>
__global__ void __launch_bounds__(kBlock) k_track_element(LatDesc L, int32_t ie, int32_t turn,
                                                          BeamPtrs b, int64_t n,
                                                          unsigned long long *work)
{
    track_element_pass<F, Q>(L, ie, turn, b, n, work);
}

// The same with the turn index read from device memory: the launches of one turn captured once
// in a graph serve every turn, the turn advanced by k_step_turn at the end of each replay.
template <bool F, int Q>
__global__ void __launch_bounds__(kBlock) k_track_element_at(LatDesc L, int32_t ie,
                                                             const int32_t *__restrict__ turn_p,
                                                             BeamPtrs b, int64_t n,
                                                             unsigned long long *work)
{
    track_element_pass<F, Q>(L, ie, *turn_p, b, n, work);
}

// turn_p = v (set_v != 0) or turn_p += 1: one thread.
__global__ void k_step_turn(int32_t *turn_p, int32_t set_v, int32_t v)
{
    if (threadIdx.x == 0) *turn_p = set_v ? v : *turn_p + 1;
}

// Elements e0..e1 (0-based, inclusive) once for every live particle, recording the state at
// every element exit: track_all / track_many (track_all.f90:25-145, track_many.f90:50-157)
// with orbit(ix) kept per particle. rec[(i * (e1 - e0 + 1) + (e - e0)) * 8 + k] holds
// x, px, y, py, z, pz, beta, state of particle i after element e. For the element a particle
// is lost in, a loss at the entrance end records zero coordinates with the loss state
// (track_all.f90:84); nothing is written for later elements.
__device__ __forceinline__ void record_orbit(double *r, const Particle &p, bool at_entrance)
{
    r[0] = at_entrance ? 0.0 : p.x;
    r[1] = at_entrance ? 0.0 : p.px;
    r[2] = at_entrance ? 0.0 : p.y;
    r[3] = at_entrance ? 0.0 : p.py;
    r[4] = at_entrance ? 0.0 : p.z;
    r[5] = at_entrance ? 0.0 : p.pz;
    r[6] = p.beta;
    r[7] = (double)p.state;
}

template <bool F, int Q>
__global__ void __launch_bounds__(kBlock) k_track_range(LatDesc L, int32_t e0, int32_t e1,
                                                        int32_t turn, BeamPtrs b, int64_t n,
                                                        double *__restrict__ rec,
                                                        unsigned long long *work)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = i < n && b.state[i] == ST_ALIVE;
    unsigned long long passes = 0;
    if (active) {
        Particle p;
        load_particle(p, b, i);
        refresh_energy(p, L.sp);
        const int64_t ne = (int64_t)(e1 - e0 + 1);
        int32_t lost_e = 0;
        if (orbit_too_large(p)) {
            lost_e = e0 + 1;
            record_orbit(rec + i * ne * 8, p, true);
        } else {
            for (int32_t e = e0; e <= e1; ++e) {
                const bool at_entrance =
                    track_element<F, Q == 2 ? MAPS_ALL_QR : Q ? MAPS_ALL_Q : MAPS_ALL>(p, load_ele((CHdr)L.hdr + 4 * e), (CP)L.prm, L.sp, e > e0, e + 1,
                                     turn);
                passes++;
                record_orbit(rec + (i * ne + (e - e0)) * 8, p, at_entrance);
                if (p.state != ST_ALIVE) {
                    lost_e = e + 1;
                    break;
                }
            }
        }
        store_particle(p, b, i);
        if (p.state != ST_ALIVE) {
            b.state[i] = p.state;
            b.loss_ele[i] = lost_e;
            b.loss_turn[i] = turn;
        }
    }
    add_work(work, passes);
}

__global__ void __launch_bounds__(kBlock) k_count_alive(const int32_t *__restrict__ state, int64_t n,
                                                        unsigned long long *out)
{
    __shared__ unsigned int part[kBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool alive = i < n && state[i] == ST_ALIVE;
    const unsigned long long m = __ballot(alive);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int w = 0; w < kBlock / 64; w++) s += part[w];
        if (s) atomicAdd(out, s);
    }
}

template <bool F, int W, bool SPAN = false, int RARE = MAPS_RARE>
static void launch_tiles(hipStream_t s, int grid, const Group *g, const WaveTile *t, const uint32_t *nt,
                         const double *pi, double *po, uint32_t stride, uint32_t *co, uint32_t *tc,
                         int32_t off, int32_t K, unsigned long long *w)
{
    hipLaunchKernelGGL((k_track_tiles<F, W, SPAN, RARE>), dim3((unsigned)grid), dim3(kBlock), 0, s, g, t, nt,
                       pi, po, stride, co, tc, off, K, w);
}

hipError_t launch_track_tiles(hipStream_t s, int grid, int waves, const Group *d_groups,
                              const WaveTile *d_tiles, const uint32_t *d_n_tiles,
                              const double *d_pk_in, double *d_pk_out, uint32_t stride,
                              uint32_t *d_cnt_out, uint32_t *d_tile_ctr, int32_t chunk_off,
                              int32_t chunk_turns, unsigned long long *d_work, bool fast, bool span,
                              int ext)
{
    if (grid <= 0) return hipSuccess;
#define BMX_TILES(F, W, ...)                                                                      \
    launch_tiles<F, W, ##__VA_ARGS__>(s, grid, d_groups, d_tiles, d_n_tiles, d_pk_in, d_pk_out,   \
                                      stride, d_cnt_out, d_tile_ctr, chunk_off, chunk_turns, d_work)
    if (ext == 2) {
        if (span) {
            if (fast) BMX_TILES(true, kExtWaves, true, MAPS_RARE_QR);
            else BMX_TILES(false, kExtWaves, true, MAPS_RARE_QR);
        } else {
            if (fast) BMX_TILES(true, kExtWaves, false, MAPS_RARE_QR);
            else BMX_TILES(false, kExtWaves, false, MAPS_RARE_QR);
        }
    } else if (ext) {
        if (span) {
            if (fast) BMX_TILES(true, kExtWaves, true, MAPS_RARE_Q);
            else BMX_TILES(false, kExtWaves, true, MAPS_RARE_Q);
        } else {
            if (fast) BMX_TILES(true, kExtWaves, false, MAPS_RARE_Q);
            else BMX_TILES(false, kExtWaves, false, MAPS_RARE_Q);
        }
    } else if (span) {
        if (fast) BMX_TILES(true, kSpanWaves, true);
        else BMX_TILES(false, kSpanWaves, true);
    } else if (fast) {
        if (waves == 5) BMX_TILES(true, 5);
        else if (waves == 4) BMX_TILES(true, 4);
        else BMX_TILES(true, 3);
    } else {
        if (waves == 5) BMX_TILES(false, 5);
        else if (waves == 4) BMX_TILES(false, 4);
        else BMX_TILES(false, 3);
    }
#undef BMX_TILES
    return hipGetLastError();
}

hipError_t launch_build_tiles(hipStream_t s, const Group *d_groups, int32_t n_groups,
                              const uint32_t *d_cnt_in, int32_t next_off, WaveTile *d_tiles,
                              uint32_t *d_n_tiles, uint32_t *d_cnt_zero, uint32_t *d_tile_ctr,
                              uint32_t *d_goff)
{
    hipLaunchKernelGGL(k_build_tiles, dim3(1), dim3(kBuildThreads), 0, s, d_groups, n_groups,
                       d_cnt_in, next_off, d_tiles, d_n_tiles, d_cnt_zero, d_tile_ctr, d_goff);
    return hipGetLastError();
}

hipError_t launch_pack_alive(hipStream_t s, int grid, const Group *d_groups,
                             const WaveTile *d_tiles, const uint32_t *d_n_tiles, double *d_pk_out,
                             uint32_t stride, uint32_t *d_cnt_out)
{
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_alive, dim3((unsigned)grid), dim3(kBlock), 0, s, d_groups, d_tiles,
                       d_n_tiles, d_pk_out, stride, d_cnt_out);
    return hipGetLastError();
}

hipError_t track_tiles_occupancy(bool fast, int waves, int *blocks_per_cu, bool span, int ext)
{
    const void *k;
    if (ext == 2 && span)
        k = fast ? (const void *)k_track_tiles<true, kExtWaves, true, MAPS_RARE_QR>
                 : (const void *)k_track_tiles<false, kExtWaves, true, MAPS_RARE_QR>;
    else if (ext == 2)
        k = fast ? (const void *)k_track_tiles<true, kExtWaves, false, MAPS_RARE_QR>
                 : (const void *)k_track_tiles<false, kExtWaves, false, MAPS_RARE_QR>;
    else if (ext && span)
        k = fast ? (const void *)k_track_tiles<true, kExtWaves, true, MAPS_RARE_Q>
                 : (const void *)k_track_tiles<false, kExtWaves, true, MAPS_RARE_Q>;
    else if (ext)
        k = fast ? (const void *)k_track_tiles<true, kExtWaves, false, MAPS_RARE_Q>
                 : (const void *)k_track_tiles<false, kExtWaves, false, MAPS_RARE_Q>;
    else if (span)
        k = fast ? (const void *)k_track_tiles<true, kSpanWaves, true>
                 : (const void *)k_track_tiles<false, kSpanWaves, true>;
    else if (fast) k = waves == 5 ? (const void *)k_track_tiles<true, 5> : waves == 4 ? (const void *)k_track_tiles<true, 4> : (const void *)k_track_tiles<true, 3>;
    else k = waves == 5 ? (const void *)k_track_tiles<false, 5> : waves == 4 ? (const void *)k_track_tiles<false, 4> : (const void *)k_track_tiles<false, 3>;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, kBlock, 0);
}

hipError_t launch_track_element(hipStream_t s, const LatDesc &lat, int32_t ie, int32_t turn,
                                const BeamPtrs &b, int64_t n, unsigned long long *d_work,
                                bool fast)
{
    int64_t grid = (n + kBlock - 1) / kBlock;
    if (grid <= 0) return hipSuccess;
#define BMX_ELEMENT(F, Q) \
    hipLaunchKernelGGL((k_track_element<F, Q>), dim3((unsigned)grid), dim3(kBlock), 0, s, lat, ie, turn, b, n, d_work)
    if (lat.rare_q == 2) {
        if (fast) BMX_ELEMENT(true, 2);
        else BMX_ELEMENT(false, 2);
    } else if (lat.rare_q) {
        if (fast) BMX_ELEMENT(true, 1);
        else BMX_ELEMENT(false, 1);
    } else {
        if (fast) BMX_ELEMENT(true, 0);
        else BMX_ELEMENT(false, 0);
    }
#undef BMX_ELEMENT
    return hipGetLastError();
}

hipError_t launch_track_element_at(hipStream_t s, const LatDesc &lat, int32_t ie, const int32_t *d_turn,
                                   const BeamPtrs &b, int64_t n, unsigned long long *d_work,
                                   bool fast)
{
    int64_t grid = (n + kBlock - 1) / kBlock;
    if (grid <= 0) return hipSuccess;
#define BMX_ELEMENT(F, Q) \
    hipLaunchKernelGGL((k_track_element_at<F, Q>), dim3((unsigned)grid), dim3(kBlock), 0, s, lat, ie, d_turn, b, n, d_work)
    if (lat.rare_q == 2) {
        if (fast) BMX_ELEMENT(true, 2);
        else BMX_ELEMENT(false, 2);
    } else if (lat.rare_q) {
        if (fast) BMX_ELEMENT(true, 1);
        else BMX_ELEMENT(false, 1);
    } else {
        if (fast) BMX_ELEMENT(true, 0);
        else BMX_ELEMENT(false, 0);
    }
#undef BMX_ELEMENT
    return hipGetLastError();
}

hipError_t launch_step_turn(hipStream_t s, int32_t *d_turn, bool set, int32_t v)
{
    hipLaunchKernelGGL(k_step_turn, dim3(1), dim3(64), 0, s, d_turn, set ? 1 : 0, v);
    return hipGetLastError();
}

hipError_t launch_track_range(hipStream_t s, const LatDesc &lat, int32_t e0, int32_t e1,
                              int32_t turn, const BeamPtrs &b, int64_t n, double *d_rec,
                              unsigned long long *d_work, bool fast)
{
    int64_t grid = (n + kBlock - 1) / kBlock;
    if (grid <= 0) return hipSuccess;
#define BMX_RANGE(F, Q) \
    hipLaunchKernelGGL((k_track_range<F, Q>), dim3((unsigned)grid), dim3(kBlock), 0, s, lat, e0, e1, turn, b, n, d_rec, d_work)
    if (lat.rare_q == 2) {
        if (fast) BMX_RANGE(true, 2);
        else BMX_RANGE(false, 2);
    } else if (lat.rare_q) {
        if (fast) BMX_RANGE(true, 1);
        else BMX_RANGE(false, 1);
    } else {
        if (fast) BMX_RANGE(true, 0);
        else BMX_RANGE(false, 0);
    }
#undef BMX_RANGE
    return hipGetLastError();
}

hipError_t launch_count_alive(hipStream_t s, const int32_t *state, int64_t n,
                              unsigned long long *d_out)
{
    int64_t grid = (n + kBlock - 1) / kBlock;
    if (grid <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_count_alive, dim3((unsigned)grid), dim3(kBlock), 0, s, state, n, d_out);
    return hipGetLastError();
}

}  // namespace bmx
