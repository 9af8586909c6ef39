// The allocation-free tracking entry point (bmx_track_turns_noalloc).
//
// bmx_track_turns keeps its run buffers (alive lists, wave tiles, counters, the uploaded group)
// in the context and grows them with hipMalloc when a beam is larger than any before. Callers
// that must not allocate -- code under HIP graph capture, tight serving loops, harnesses that
// forbid it -- use this variant instead: its run buffers live in one file-scope device array of
// fixed size, `scratch`, which every GPU the library runs on holds until the process ends.
// Host code reaches each device's copy through the address hipGetSymbolAddress returns for it,
// recorded once per device when a context is created on it (noalloc_init); the array's name is
// never used as a device address on the host.
//
// Every call on one GPU shares that GPU's array and leaves its contents behind; the kernels
// reset every counter they read (the tile builds zero the alive counts and the tile counter)
// before use, so nothing needs clearing between calls. Calls that overlap on one GPU (two
// contexts, two host threads) race on the array and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "bmx_internal.hpp"

namespace bmx {
// Beams of at most 2^18 particles: two packed particle buffers (60 B per particle each: seven
// coordinates and the particle's index), two tile lists (one 16-byte tile per 64 particles
// each), the group and its counters -- about 30 MiB.
constexpr int64_t kNoAllocMaxParticles = int64_t(1) << 18;

namespace {

constexpr size_t align256(size_t v) { return (v + 255) / 256 * 256; }

constexpr size_t kMaxTiles = (size_t)(kNoAllocMaxParticles + 63) / 64;
constexpr size_t kOffGroups = 0;
constexpr size_t kOffGoff = kOffGroups + align256(sizeof(Group));
constexpr size_t kOffCnt0 = kOffGoff + 256;   // goff: n_groups + 1 = 2 uint32
constexpr size_t kOffCnt1 = kOffCnt0 + 256;
constexpr size_t kOffSmall = kOffCnt1 + 256;  // 4 uint32
constexpr size_t kOffPacked0 = kOffSmall + 256;
constexpr size_t kOffPacked1 = kOffPacked0 + align256(kNoAllocMaxParticles * kPackedBytes);
constexpr size_t kOffTiles0 = kOffPacked1 + align256(kNoAllocMaxParticles * kPackedBytes);
constexpr size_t kOffTiles1 = kOffTiles0 + align256(kMaxTiles * sizeof(WaveTile));
constexpr size_t kScratchBytes = kOffTiles1 + align256(kMaxTiles * sizeof(WaveTile));

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char *g_scratch[kMaxDevices] = {};

// The run buffers of one call: views into the device's scratch array (never released).
RunBuffers scratch_buffers(int device, unsigned char *base)
{
    RunBuffers rb;
    rb.device = device;
    rb.d_groups = reinterpret_cast<Group *>(base + kOffGroups);
    rb.d_goff = reinterpret_cast<uint32_t *>(base + kOffGoff);
    rb.d_cnt[0] = reinterpret_cast<uint32_t *>(base + kOffCnt0);
    rb.d_cnt[1] = reinterpret_cast<uint32_t *>(base + kOffCnt1);
    rb.d_small = reinterpret_cast<uint32_t *>(base + kOffSmall);
    rb.d_packed[0] = reinterpret_cast<double *>(base + kOffPacked0);
    rb.d_packed[1] = reinterpret_cast<double *>(base + kOffPacked1);
    rb.d_tiles[0] = reinterpret_cast<WaveTile *>(base + kOffTiles0);
    rb.d_tiles[1] = reinterpret_cast<WaveTile *>(base + kOffTiles1);
    rb.cap_groups = 1;
    rb.cap_particles = (size_t)kNoAllocMaxParticles;
    rb.cap_tiles = kMaxTiles;
    return rb;
}

}  // namespace
}  // namespace bmx

// The fixed scratch array (one copy per device). Host code uses only the per-device address
// noalloc_init records, never this name.
static __device__ __align__(256) unsigned char scratch[bmx::kScratchBytes];

namespace bmx {

hipError_t noalloc_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<unsigned char *>(p);
    return hipSuccess;
}

}  // namespace bmx

using namespace bmx;

extern "C" int32_t bmx_track_turns_noalloc(bmx_ctx *ctx, const bmx_lattice *lat, bmx_beam *beam,
                                           int32_t n_turns, int32_t turn0)
{
    if (!ctx || !lat || !beam) return set_error(BMX_ERR_ARG, "bmx_track_turns_noalloc: null argument");
    if (beam->n > kNoAllocMaxParticles)
        return set_error(BMX_ERR_ARG, "bmx_track_turns_noalloc: " + std::to_string(beam->n) +
                                          " particles exceed the fixed scratch of " +
                                          std::to_string(kNoAllocMaxParticles));
    unsigned char *base = nullptr;
    if (ctx->device >= 0 && ctx->device < kMaxDevices) {
        std::lock_guard<std::mutex> lock(g_mutex);
        base = g_scratch[ctx->device];
    }
    if (!base) return set_error(BMX_ERR_HIP, "bmx_track_turns_noalloc: no scratch address for the device");
    RunBuffers rb = scratch_buffers(ctx->device, base);
    return track_group(ctx, lat, beam, 0, lat->desc.n_ele, n_turns, turn0, "bmx_track_turns_noalloc",
                       &rb);
}
