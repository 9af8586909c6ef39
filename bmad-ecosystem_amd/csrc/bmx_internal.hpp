// Internal declarations shared by the kernels, the table compiler and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/bmx.h"
#include "bmx_table.hpp"

namespace bmx {

// ---- kernel launchers (bmx_kernels.hip) ----
hipError_t launch_track_turns(hipStream_t s, const LatDesc *d_lats, const Tile *d_tiles,
                              int64_t n_tiles, const BeamPtrs &b, int64_t n, int32_t n_turns,
                              int32_t turn0, unsigned long long *d_work, bool fast);
hipError_t launch_track_element(hipStream_t s, const LatDesc &lat, int32_t ie, int32_t turn,
                                const BeamPtrs &b, int64_t n, unsigned long long *d_work,
                                bool fast);
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
};
// Returns BMX_OK or an error code with a message in err.
int compile_table(const bmx_ele *eles, int32_t n_ele, double mc2, int32_t charge, HostTable &out,
                  std::string &err);

// ---- error reporting (bmx_api.cpp) ----
int set_error(int code, const std::string &msg);

}  // namespace bmx

struct bmx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    unsigned long long *d_counter = nullptr;  // scratch for reductions
    unsigned long long *d_work = nullptr;     // particle-element passes since the last take
    hipEvent_t events[BMX_N_EVENTS] = {};
    bool fast_maps = true;  // BMX_MAPS_FAST (default) or BMX_MAPS_REFERENCE
};

struct bmx_lattice {
    bmx_ctx *ctx = nullptr;
    int device = 0;
    bmx::HostTable host;
    bmx::EleHdr *d_hdr = nullptr;
    double *d_prm = nullptr;
    bmx::LatDesc desc{};              // device pointers inside
    bmx::LatDesc *d_desc = nullptr;   // one-element array for the tile kernel
};

struct bmx_beam {
    bmx_ctx *ctx = nullptr;
    int device = 0;
    int64_t n = 0;
    bool owned = false;
    void *block = nullptr;  // single allocation when owned
    bmx::BeamPtrs d{};
};
