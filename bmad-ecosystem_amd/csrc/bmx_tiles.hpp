// Device code of the persistent tile kernel, shared by the library's build of it
// (bmx_kernels.hip) and by lattice-specialised builds compiled at run time (bmx_jit.cpp), which
// differ only in the turn policy: how one particle is taken once around the ring.
#pragma once
#include <hip/hip_runtime.h>

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
    p.id = (uint32_t)i;
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

// A lost particle goes back to its beam with its loss record.
__device__ __forceinline__ void store_lost(const Particle &p, const BeamPtrs &b, int64_t i)
{
    store_particle(p, b, i);
    b.state[i] = p.state;
    b.loss_ele[i] = p.loss_ele;
    b.loss_turn[i] = p.loss_turn;
}

// Slot j of a packed particle buffer (bmx_table.hpp, Packed layout).
__device__ __forceinline__ uint32_t load_packed(Particle &p, const double *__restrict__ d,
                                                uint32_t stride, uint32_t j)
{
    p.x = d[j];
    p.px = d[(size_t)stride + j];
    p.y = d[2 * (size_t)stride + j];
    p.py = d[3 * (size_t)stride + j];
    p.z = d[4 * (size_t)stride + j];
    p.pz = d[5 * (size_t)stride + j];
    p.beta = d[6 * (size_t)stride + j];
    p.state = ST_ALIVE;
    p.id = reinterpret_cast<const uint32_t *>(d + kPackedDoubles * (size_t)stride)[j];
    return p.id;
}

__device__ __forceinline__ void store_packed(const Particle &p, double *__restrict__ d,
                                             uint32_t stride, uint32_t j, uint32_t id)
{
    d[j] = p.x;
    d[(size_t)stride + j] = p.px;
    d[2 * (size_t)stride + j] = p.y;
    d[3 * (size_t)stride + j] = p.py;
    d[4 * (size_t)stride + j] = p.z;
    d[5 * (size_t)stride + j] = p.pz;
    d[6 * (size_t)stride + j] = p.beta;
    reinterpret_cast<uint32_t *>(d + kPackedDoubles * (size_t)stride)[j] = id;
}

// Adds the number of particle-element passes tracked by this wave to the pass counter (if
// non-null): element passes of live particles only, as the reference's loop skips dead ones
// (beam_utils.f90:83). The counter is kWorkSlots slots that the reader sums (bmx_table.hpp), so
// the atomics of many waves do not all meet at one address.
__device__ __forceinline__ void add_work(unsigned long long *work, unsigned long long mine)
{
    if (!work) return;
    unsigned long long s = mine;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(work + blockIdx.x % kWorkSlots, s);
}

// The generic turn policy: elements [e0, e1) of the lattice read from its table. The rf
// cavities (and the other rare elements: offsets, momentum windows, quadrupoles with
// multipoles) sit outside the inner element loop: the loop over the other kinds runs until the
// next such element, which is tracked on its own, so its code does not shape the register
// assignment of the common path. The loops are wave-uniform (the element index stays a
// scalar, so the table is read by scalar loads); a lost particle's lane is masked off for the
// rest of the turn. RARE: the map set of the rare path (MAPS_RARE, or MAPS_RARE_Q for a
// lattice with stepped quadrupoles).
template <bool F, int RARE = MAPS_RARE>
__device__ __forceinline__ void track_range(Particle &p, const LatDesc &L, int32_t e0, int32_t e1,
                                            int32_t t_abs)
{
    const CHdr hdr = (CHdr)L.hdr;
    const CP prm = (CP)L.prm;
    int32_t e = e0;
    for (;;) {
        for (; e < e1; ++e) {
            // e is the same on every active lane; read it from the first one so the compiler
            // keeps it (and the table reads) scalar whatever it makes of the loop's control flow
            e = __builtin_amdgcn_readfirstlane(e);
            const CHdr h = hdr + 4 * e;
            if (h[0] == K_RF || (h[1] & (F_OFFSET | F_APE_MOMENTUM | F_RARE_MAP))) break;
            if (p.state == ST_ALIVE)
                track_element<F, MAPS_NO_RF>(p, load_ele(h), prm, L.sp, true, e + 1, t_abs);
        }
        if (e >= e1) break;
        e = __builtin_amdgcn_readfirstlane(e);
        if (p.state == ST_ALIVE)
            track_element<F, RARE>(p, load_ele(hdr + 4 * e), prm, L.sp, true, e + 1, t_abs);
        ++e;
    }
}

template <bool F, int RARE = MAPS_RARE>
struct GenericTurn {
    static constexpr int kRare = RARE;
    __device__ __forceinline__ static void turn(Particle &p, const LatDesc &L, int32_t t_abs)
    {
        track_range<F, RARE>(p, L, 0, L.n_ele, t_abs);
    }
};

// `count` elements from element `first` (0-based), wrapping past the last element to element 0
// (beam_mod.f90:159-166, track_bunch's forward loops): the turn of a group that does not start
// at element 0 or does not pass every element.
template <bool F, int RARE>
__device__ __forceinline__ void track_span(Particle &p, const LatDesc &L, int32_t first,
                                           int32_t count, int32_t t_abs)
{
    // at most two segments (first .. end, then 0 .. rest), one call site of the element loop
    int32_t e0 = first, left = count;
    while (left > 0) {
        const int32_t e1 = left < L.n_ele - e0 ? e0 + left : L.n_ele;
        track_range<F, RARE>(p, L, e0, e1, t_abs);
        left -= e1 - e0;
        e0 = 0;
    }
}

// The entrance aperture test of element e for a particle that does not arrive from element
// e - 1 in this launch (it resumes at e): made here when the table marks it as a repeat of the
// previous element's exit test (F_APE_ENT_SAME), which the element loop skips. A particle that
// did pass that exit test passes this one too, so making it again changes no result.
__device__ __forceinline__ void entrance_retest(Particle &p, const LatDesc &L, int32_t e,
                                                int32_t turn)
{
    const EleRef h = load_ele((CHdr)L.hdr + 4 * e);
    if (!(h.flags & F_APE_ENT_SAME)) return;
    CP ap = (CP)L.prm + h.off + ((h.flags & F_OFFSET) ? kOffsetLen : 0);
    const bool out = aperture_fails(p, h.flags, ap);
    if (any_lane(out)) {
        if (out) {
            aperture(p, h.flags, ap);
            p.loss_ele = uniform_here(e + 1);
            p.loss_turn = uniform_here(turn);
        }
    }
}

// Tracks one particle through up to n_turns turns starting at turn index turn0 and records the
// loss element (1-based) and turn. SPAN false: full turns from element 0 (first = 0, count =
// n_ele) by the turn policy Turn; SPAN true: turns of the group's element span (first, count)
// through the generic element loop. Returns the number of element passes made.
template <bool F, class Turn, bool SPAN>
__device__ __forceinline__ unsigned long long track_particle(Particle &p, const LatDesc &L,
                                                             int32_t n_turns, int32_t turn0,
                                                             int32_t first, int32_t count)
{
    refresh_energy(p, L.sp);
    p.loss_ele = 0;
    p.loss_turn = 0;
    // Entrance validity check of the first element (track1.f90:196 -> orbit_too_large); later
    // elements need none: a particle leaves element e with the coordinates it enters e+1 with,
    // and element e already ended with the same test (track1.f90:291).
    if (orbit_too_large(p)) {
        p.loss_ele = first + 1;
        p.loss_turn = turn0;
    } else {
        if constexpr (SPAN) {
            if (first > 0) entrance_retest(p, L, first, turn0);
        }
        // a wave stops once all its particles are lost
        for (int32_t turn = 0; turn < n_turns; ++turn) {
            if (!any_lane(p.state == ST_ALIVE)) break;
            if constexpr (SPAN) track_span<F, Turn::kRare>(p, L, first, count, turn0 + turn);
            else Turn::turn(p, L, turn0 + turn);
        }
    }
    if (p.state != ST_ALIVE) {
        // passes up to and including the loss element, counted from `first`
        const int32_t pos = p.loss_ele - 1 - first;
        return (unsigned long long)(p.loss_turn - turn0) * count + (pos < 0 ? pos + L.n_ele : pos) + 1;
    }
    return (unsigned long long)n_turns * count;
}

// Persistent tracking of one chunk of turns over a list of wave tiles.
//
// Each wave repeatedly takes the next tile (atomic counter), reads its particles from the packed
// buffer pk_in and tracks them for the chunk's turns with the state in registers. A particle
// lost in the chunk goes back to its beam with its loss record at once; one that survives is
// appended to its group's slots of pk_out (one atomic per wave), so the next chunk's tiles hold
// live particles only, waves stay full as the beam is lost, and every tile reads whole cache
// lines. In its group's last chunk the survivors go back to the beam instead. The order of the
// packed slots depends on scheduling; each particle's result does not.
//
// Every resident wave takes tiles (as many waves per SIMD as the register budget allows hide
// each other's latencies better than fewer waves with evenly divided tiles).
// chunk_off: turns already done by earlier chunks of this call; chunk_turns: turns per chunk.
// SPAN: groups may track element spans (Group first/count); false: every group tracks full
// turns from element 0, and the kernel carries no code for spans.
template <bool F, class Turn, bool SPAN = false>
__device__ __forceinline__ void track_tiles(
    const Group *__restrict__ groups, const WaveTile *__restrict__ tiles,
    const uint32_t *__restrict__ n_tiles_p, const double *__restrict__ pk_in,
    double *__restrict__ pk_out, uint32_t stride, uint32_t *__restrict__ cnt_out,
    uint32_t *tile_ctr, int32_t chunk_off, int32_t chunk_turns, unsigned long long *work)
{
    const uint32_t nt = *n_tiles_p;
    const int lane = threadIdx.x & 63;
    const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (wave >= nt) return;  // whole wave
    unsigned long long passes = 0;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(tile_ctr, 1u);
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= nt) break;
        const WaveTile T = tiles[t];
        const Group &G = groups[T.group];
        const int32_t rem = G.n_turns - chunk_off;
        const bool last = rem <= chunk_turns;  // the group's last chunk of this run
        const int32_t n_turns = last ? rem : chunk_turns;
        const int32_t turn0 = G.turn0 + chunk_off;
        bool keep = false;
        uint32_t id = 0;
        Particle p;
        if ((uint32_t)lane < T.count) {
            id = load_packed(p, pk_in, stride, G.list_off + T.start + lane);
            passes += track_particle<F, Turn, SPAN>(p, G.lat, n_turns, turn0,
                                                    SPAN ? G.first : 0,
                                                    SPAN ? G.count : G.lat.n_ele);
            if (p.state != ST_ALIVE) store_lost(p, G.b, id);
            else if (last) store_particle(p, G.b, id);
            else keep = true;
        }
        const unsigned long long m = __ballot(keep);
        if (m) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(&cnt_out[T.group], (uint32_t)__popcll(m));
            base = __builtin_amdgcn_readfirstlane(base);
            if (keep) {
                const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1));
                store_packed(p, pk_out, stride, G.list_off + base + rank, id);
            }
        }
    }
    add_work(work, passes);
}

}  // namespace bmx
