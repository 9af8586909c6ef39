// Tracking kernels for gfx950.
//
//  k_track_turns   persistent tracking: one lane per particle, state in registers for all
//                  elements and turns of the call; work comes as tiles (one tile per
//                  workgroup) so single-lattice, batched and served launches share the kernel.
//  k_track_element one element for every particle, state round-trips HBM: the GPU form of
//                  track1_bunch_hom (bmad/multiparticle/beam_utils.f90:32-90).
//  k_count_alive   number of live particles (lt_tracking_mod.f90:1672 bunch%n_live).
#include <hip/hip_runtime.h>

#include "bmx_internal.hpp"
#include "bmx_maps.hpp"

namespace bmx {

constexpr int kBlock = 256;

__device__ __forceinline__ void load_particle(Particle &p, const BeamPtrs &b, int64_t i)
{
    p.x = b.x[i];
    p.px = b.px[i];
    p.y = b.y[i];
    p.py = b.py[i];
    p.z = b.z[i];
    p.pz = b.pz[i];
    p.beta = b.beta[i];
    p.state = ST_ALIVE;
}

__device__ __forceinline__ void store_particle(const Particle &p, const BeamPtrs &b, int64_t i)
{
    b.x[i] = p.x;
    b.px[i] = p.px;
    b.y[i] = p.y;
    b.py[i] = p.py;
    b.z[i] = p.z;
    b.pz[i] = p.pz;
    b.beta[i] = p.beta;
}

// Adds the number of particle-element passes tracked by this wave to *work (if non-null):
// element passes of live particles only, as the reference's loop skips dead ones
// (beam_utils.f90:83).
__device__ __forceinline__ void add_work(unsigned long long *work, unsigned long long mine)
{
    if (!work) return;
    unsigned long long s = mine;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(work, s);
}

// Persistent tracking over tiles (tiles == nullptr: implicit tiles of kBlock particles of
// lattice 0 covering [0, n) with the given n_turns / turn0).
template <bool F>
__global__ void __launch_bounds__(kBlock) k_track_turns(const LatDesc *__restrict__ lats,
                                                        const Tile *__restrict__ tiles, BeamPtrs b,
                                                        int64_t n, int32_t n_turns, int32_t turn0,
                                                        unsigned long long *work)
{
    Tile t;
    if (tiles) {
        t = tiles[blockIdx.x];
    } else {
        t.lat = 0;
        t.start = (int64_t)blockIdx.x * kBlock;
        const int64_t rem = n - t.start;
        t.count = rem < kBlock ? (int32_t)rem : kBlock;
        t.n_turns = n_turns;
        t.turn0 = turn0;
    }
    const int li = threadIdx.x;
    const int64_t i = t.start + li;
    // No early return: every lane of the wave takes part in the work reduction at the end.
    const bool active = li < t.count && t.n_turns > 0 && b.state[i] == ST_ALIVE;
    unsigned long long passes = 0;
    if (active) {
        const LatDesc L = lats[t.lat];
        Particle p;
        load_particle(p, b, i);
        refresh_energy(p, L.sp);

        int32_t lost_e = 0, lost_t = 0;
        // Entrance validity check of the first element (track1.f90:196 -> orbit_too_large);
        // later elements need none: a particle leaves element e with the coordinates it enters
        // e+1 with, and element e already ended with the same test (track1.f90:291).
        if (orbit_too_large(p)) {
            lost_e = 1;
            lost_t = t.turn0;
        } else {
            const CHdr hdr = (CHdr)L.hdr;
            const CP prm = (CP)L.prm;
            for (int32_t turn = 0; turn < t.n_turns; ++turn) {
                for (int32_t e = 0; e < L.n_ele; ++e) {
                    track_element<F>(p, hdr + 4 * e, prm, L.sp);
                    if (p.state != ST_ALIVE) {
                        lost_e = e + 1;
                        lost_t = t.turn0 + turn;
                        break;
                    }
                }
                if (p.state != ST_ALIVE) break;
            }
        }
        store_particle(p, b, i);
        if (p.state != ST_ALIVE) {
            b.state[i] = p.state;
            b.loss_ele[i] = lost_e;
            b.loss_turn[i] = lost_t;
            passes = (unsigned long long)(lost_t - t.turn0) * L.n_ele + lost_e;
        } else {
            passes = (unsigned long long)t.n_turns * L.n_ele;
        }
    }
    add_work(work, passes);
}

// One element, whole bunch: track1 (track1.f90:44-343) per live particle.
template <bool F>
__global__ void __launch_bounds__(kBlock) k_track_element(LatDesc L, int32_t ie, int32_t turn,
                                                          BeamPtrs b, int64_t n,
                                                          unsigned long long *work)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool active = i < n && b.state[i] == ST_ALIVE;
    if (active) {
        Particle p;
        load_particle(p, b, i);
        refresh_energy(p, L.sp);
        if (!orbit_too_large(p)) track_element<F>(p, (CHdr)L.hdr + 4 * ie, (CP)L.prm, L.sp);
        store_particle(p, b, i);
        if (p.state != ST_ALIVE) {
            b.state[i] = p.state;
            b.loss_ele[i] = ie + 1;
            b.loss_turn[i] = turn;
        }
    }
    add_work(work, active ? 1ull : 0ull);
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

template <bool F>
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
                const bool at_entrance = track_element<F>(p, (CHdr)L.hdr + 4 * e, (CP)L.prm, L.sp);
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

hipError_t launch_track_turns(hipStream_t s, const LatDesc *d_lats, const Tile *d_tiles,
                              int64_t n_tiles, const BeamPtrs &b, int64_t n, int32_t n_turns,
                              int32_t turn0, unsigned long long *d_work, bool fast)
{
    int64_t grid = d_tiles ? n_tiles : (n + kBlock - 1) / kBlock;
    if (grid <= 0) return hipSuccess;
    if (fast)
        hipLaunchKernelGGL(k_track_turns<true>, dim3((unsigned)grid), dim3(kBlock), 0, s, d_lats,
                           d_tiles, b, n, n_turns, turn0, d_work);
    else
        hipLaunchKernelGGL(k_track_turns<false>, dim3((unsigned)grid), dim3(kBlock), 0, s, d_lats,
                           d_tiles, b, n, n_turns, turn0, d_work);
    return hipGetLastError();
}

hipError_t launch_track_element(hipStream_t s, const LatDesc &lat, int32_t ie, int32_t turn,
                                const BeamPtrs &b, int64_t n, unsigned long long *d_work,
                                bool fast)
{
    int64_t grid = (n + kBlock - 1) / kBlock;
    if (grid <= 0) return hipSuccess;
    if (fast)
        hipLaunchKernelGGL(k_track_element<true>, dim3((unsigned)grid), dim3(kBlock), 0, s, lat, ie,
                           turn, b, n, d_work);
    else
        hipLaunchKernelGGL(k_track_element<false>, dim3((unsigned)grid), dim3(kBlock), 0, s, lat, ie,
                           turn, b, n, d_work);
    return hipGetLastError();
}

hipError_t launch_track_range(hipStream_t s, const LatDesc &lat, int32_t e0, int32_t e1,
                              int32_t turn, const BeamPtrs &b, int64_t n, double *d_rec,
                              unsigned long long *d_work, bool fast)
{
    int64_t grid = (n + kBlock - 1) / kBlock;
    if (grid <= 0) return hipSuccess;
    if (fast)
        hipLaunchKernelGGL(k_track_range<true>, dim3((unsigned)grid), dim3(kBlock), 0, s, lat, e0,
                           e1, turn, b, n, d_rec, d_work);
    else
        hipLaunchKernelGGL(k_track_range<false>, dim3((unsigned)grid), dim3(kBlock), 0, s, lat, e0,
                           e1, turn, b, n, d_rec, d_work);
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
