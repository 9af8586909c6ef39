// Internal declarations shared by the kernels, the table compiler and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/bmx.h"
#include "bmx_table.hpp"

namespace bmx {

// ---- kernel launchers (bmx_kernels.hip) ----
hipError_t launch_track_tiles(hipStream_t s, int grid, int waves, const Group *d_groups,
                              const WaveTile *d_tiles, const uint32_t *d_n_tiles,
                              const double *d_pk_in, double *d_pk_out, uint32_t stride,
                              uint32_t *d_cnt_out, uint32_t *d_tile_ctr, int32_t chunk_off,
                              int32_t chunk_turns, unsigned long long *d_work, bool fast, bool span,
                              int ext = 0);
hipError_t launch_build_tiles(hipStream_t s, const Group *d_groups, int32_t n_groups,
                              const uint32_t *d_cnt_in, int32_t next_off, WaveTile *d_tiles,
                              uint32_t *d_n_tiles, uint32_t *d_cnt_zero, uint32_t *d_tile_ctr,
                              uint32_t *d_goff);
hipError_t launch_pack_alive(hipStream_t s, int grid, const Group *d_groups,
                             const WaveTile *d_tiles, const uint32_t *d_n_tiles, double *d_pk_out,
                             uint32_t stride, uint32_t *d_cnt_out);
// ext: 0 the plain map sets, 1 with the stepped quadrupole and general bend (LatDesc::rare_q),
// 2 with radiation as well.
hipError_t track_tiles_occupancy(bool fast, int waves, int *blocks_per_cu, bool span = false,
                                 int ext = 0);
// Waves per SIMD of the kernel for runs with element spans (Group first/count other than the
// full turn from element 0).
constexpr int kSpanWaves = 3;
// Waves per SIMD of the generic kernels for lattices with stepped quadrupoles or general bends
// (the register budget of 2 waves, so that their long maps do not spill).
constexpr int kExtWaves = 2;
hipError_t launch_track_element(hipStream_t s, const LatDesc &lat, int32_t ie, int32_t turn,
                                const BeamPtrs &b, int64_t n, unsigned long long *d_work,
                                bool fast);
hipError_t launch_track_element_at(hipStream_t s, const LatDesc &lat, int32_t ie, const int32_t *d_turn,
                                   const BeamPtrs &b, int64_t n, unsigned long long *d_work,
                                   bool fast);
hipError_t launch_step_turn(hipStream_t s, int32_t *d_turn, bool set, int32_t v);
hipError_t launch_track_range(hipStream_t s, const LatDesc &lat, int32_t e0, int32_t e1,
                              int32_t turn, const BeamPtrs &b, int64_t n, double *d_rec,
                              unsigned long long *d_work, bool fast);
hipError_t launch_count_alive(hipStream_t s, const int32_t *state, int64_t n,
                              unsigned long long *d_out);

// ---- table compiler (bmx_compile.cpp) ----
struct HostTable {
    std::vector<EleHdr> hdr;
    std::vector<double> prm;
    Species sp{};
    uint64_t kind_mask = 0;  // bit = bmx_ele key present
    bool rare_q = false;     // some element compiled to K_QMULT or K_BMULT
    bool rad = false;        // some element carries radiation kicks (rare_q is then set too)
};
// Returns BMX_OK or an error code with a message in err.
// rad: null, or rm0 and rm1 of every element (2 n_ele maps) for radiation kicks; rad_cfg the
// lattice's radiation settings (with rad).
int compile_table(const bmx_ele *eles, int32_t n_ele, double mc2, int32_t charge, HostTable &out,
                  std::string &err, const bmx_rad_map *rad = nullptr,
                  const bmx_rad_config *rad_cfg = nullptr);
// The device form of radiation settings.
RadCfg rad_settings(const bmx_rad_config &c);

// ---- error reporting (bmx_api.cpp) ----
int set_error(int code, const std::string &msg);

// ---- chunked tile runs (bmx_run.cpp) ----
// Device buffers of a tile run over a set of groups: the groups, two packed particle buffers
// (cap_particles slots each, the Packed layout of bmx_table.hpp), two tile lists with their
// counts, per-group alive counters for the two packed buffers and the tile counter.
struct RunBuffers {
    int device = -1;
    size_t cap_groups = 0, cap_particles = 0, cap_tiles = 0;
    Group *d_groups = nullptr;
    double *d_packed[2] = {nullptr, nullptr};
    WaveTile *d_tiles[2] = {nullptr, nullptr};
    uint32_t *d_cnt[2] = {nullptr, nullptr};
    uint32_t *d_goff = nullptr;   // first tile of each group (n_groups + 1), tile builder scratch
    uint32_t *d_small = nullptr;  // [0..1] tile counts of the two tile lists, [2] tile counter
    void release();
    // Grows the buffers to hold the given sizes (allocates; not for use under graph capture).
    hipError_t reserve(size_t n_groups, size_t n_particles, size_t n_tiles);
};

// Event pairs around tracking-kernel launches (profiling of bmx_track_turns).
struct KernelTimer {
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    hipError_t take(hipEvent_t *ev);
    void release();
};

// Threads per workgroup of the tile kernels.
constexpr int kTileBlock = 256;

// Lattice-specialised tile kernels (bmx_jit.cpp): compiled at run time for a lattice that is
// mostly one repeated cell; shared by every lattice with the same cell structure.
struct JitKernels;
int jit_specialize(int device, const HostTable &t, const JitKernels **out, std::string &err);
hipError_t jit_occupancy(const JitKernels *j, int waves, int *blocks_per_cu);
// The cell kernels carry the stepped quadrupole (their lattice holds one).
bool jit_rare_q(const JitKernels *j);
// The map level the cell kernels carry: 0 plain, 1 stepped elements, 2 with radiation.
int jit_level(const JitKernels *j);
hipError_t launch_track_tiles_jit(const JitKernels *j, hipStream_t s, int grid, int waves,
                                  const Group *d_groups, const WaveTile *d_tiles,
                                  const uint32_t *d_n_tiles, const double *d_pk_in,
                                  double *d_pk_out, uint32_t stride, uint32_t *d_cnt_out,
                                  uint32_t *d_tile_ctr, int32_t chunk_off, int32_t chunk_turns,
                                  unsigned long long *d_work);

// Parameters of one run.
struct RunConfig {
    int32_t chunk_turns = 5;   // turns per chunk (alive lists are rebuilt between chunks)
    int32_t grid = 0;          // workgroups of the tracking kernel (0: occupancy limit x CUs)
    int32_t waves = 5;         // waves per SIMD the tracking kernel is compiled for (3, 4, 5)
    bool fast = true;
    KernelTimer *timer = nullptr;  // if set: an event pair around every tracking launch
    const JitKernels *jit = nullptr;  // cell-specialised kernels shared by every group's lattice
    bool span = false;  // some group tracks an element span: the generic span kernel runs
};

// Uploads the groups (asynchronously, from host memory that must stay valid until the stream
// reaches the copy) and enqueues the chunks that track every group for its n_turns: the tile
// build of the first chunk, then per chunk the tracking kernel and the next chunk's tile
// build. Enqueues nothing that allocates or synchronises, so a run can be captured in a graph
// once upload is false and the groups are already on the device.
hipError_t enqueue_run(hipStream_t s, RunBuffers &rb, const Group *h_groups, int32_t n_groups,
                       bool upload, int32_t max_turns, const RunConfig &cfg,
                       unsigned long long *d_work);
// Default waves per SIMD of the tracking kernel: the variant measured fastest on the 901-element
// ring for each kernel (cell-specialised 4: at most 128 VGPRs, so 4 waves per SIMD are resident,
// without spilling; generic fast maps 5; reference maps 3, whose 4- and 5-wave builds spill).
// Cell kernels that carry the stepped quadrupole or the general bend are all built for 2 waves per
// SIMD (their longer maps would spill otherwise), whichever variant is asked for.
// The live-particle rounds of the 4-wave cell kernel above which the 5-wave build runs.
constexpr double kPhaseRounds = 1.45;
inline int32_t default_waves(bool fast, const JitKernels *jit = nullptr)
{
    return jit ? (jit_rare_q(jit) ? 3 : 4) : (fast ? 5 : 3);
}
// Workgroups of the tracking kernel that fit on the device at once (of the cell-specialised
// kernels when jit is set).
int default_grid(int device, bool fast, int waves, const JitKernels *jit = nullptr,
                 bool span = false, int ext = 0);

}  // namespace bmx

struct bmx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    unsigned long long *d_counter = nullptr;  // scratch for reductions
    unsigned long long *d_work = nullptr;     // particle-element passes since the last take
    hipEvent_t events[BMX_N_EVENTS] = {};
    bool fast_maps = true;  // BMX_MAPS_FAST (default) or BMX_MAPS_REFERENCE
    int32_t chunk_turns = 5;
    int32_t grid = 0;       // 0: default_grid
    int32_t waves = 0;      // waves per SIMD of the tracking kernel variant (0: default_waves)
    struct LastRun {
        int32_t chunk_turns = 0, grid = 0, waves = 0;
    } last;                 // the run shape the last bmx_track_turns* call used
    bmx::RunBuffers run;    // buffers of bmx_track_turns
    bmx::Group group{};     // host copy of the single group uploaded by bmx_track_turns
    bool profile = false;   // time every tracking launch of bmx_track_turns
    bmx::KernelTimer timer;
    int32_t *d_turn = nullptr;  // turn index of the plain path's captured turn (in d_counter)
    // the plain path's turn of element launches, captured once per (lattice, beam, maps) and
    // replayed for every turn
    struct PlainGraph {
        bmx::LatDesc lat{};    // the kernel arguments the graph was captured with
        bmx::BeamPtrs beam{};
        int64_t n = -1;
        bool fast = true;
        hipGraphExec_t exec = nullptr;
    } plain;
};

struct bmx_lattice {
    bmx_ctx *ctx = nullptr;
    int device = 0;
    bmx::HostTable host;
    bmx::EleHdr *d_hdr = nullptr;
    double *d_prm = nullptr;
    bmx::LatDesc desc{};              // device pointers inside
    bmx::LatDesc *d_desc = nullptr;   // one-element array for the tile kernel
    const bmx::JitKernels *jit = nullptr;  // set by bmx_lattice_specialize
};

struct bmx_beam {
    bmx_ctx *ctx = nullptr;
    int device = 0;
    int64_t n = 0;
    bool owned = false;
    void *block = nullptr;  // single allocation when owned
    bmx::BeamPtrs d{};
};

namespace bmx {
// Uploads a compiled element table as a new lattice on ctx's device (bmx_lattice_create's
// device half).
int32_t lattice_upload(bmx_ctx *ctx, HostTable &&table, bmx_lattice **out);
// Persistent tracking of one group (bmx_api.cpp); fixed: run buffers to use instead of the
// context's growing ones (nullptr: the context's).
int32_t track_group(bmx_ctx *ctx, const bmx_lattice *L, bmx_beam *B, int32_t first, int32_t count,
                    int32_t n_turns, int32_t turn0, const char *who, RunBuffers *fixed);
// Records the device address of the fixed scratch array of the allocation-free entry point on
// `device`, which must be current (bmx_noalloc.hip); once per device.
hipError_t noalloc_init(int device);
}  // namespace bmx

