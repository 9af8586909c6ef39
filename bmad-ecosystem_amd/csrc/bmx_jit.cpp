// Lattice-specialised tile kernels, compiled at run time.
//
// A ring is mostly one cell repeated: the same element kinds and flags in the same order, with
// parameter blocks of the same size. For such a lattice the element loop of the tile kernel
// can be written out once per cell position with the kind and flags as compile-time constants:
// every kind dispatch and flag test folds away, and each map's results flow straight into the
// next element's code instead of through the joins of a dispatch. The source of that turn
// policy is generated here, compiled with hiprtc for the device together with the library's
// own device headers (embedded at build time, bmx_jit_sources.cpp), and loaded as a module.
// Elements outside the repeated run (a head cell that differs, the tail with the rf cavity)
// keep the generic table-driven loop. Kernels are cached per device and cell signature, so
// lattices that differ only in element strengths (error variants) share one.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "bmx_internal.hpp"

extern const char bmx_jit_src_table[];
extern const char bmx_jit_src_maps[];
extern const char bmx_jit_src_tiles[];

namespace bmx {

struct JitKernels {
    hipModule_t module = nullptr;
    hipFunction_t fn[3] = {};  // waves per SIMD 3, 4, 5
    bool rare_q = false;       // carries K_QMULT / K_BMULT on its rare path (all built for 2 waves)
    int level = 0;             // 0 plain, 1 stepped elements, 2 with radiation kicks
};

namespace {

// head elements [0, H) generic, then N cells of P elements, cell c's parameters at
// hdr[H].off + c * S, then the tail generic.
struct Period {
    int32_t H = 0, P = 0, N = 0, S = 0;
};

bool cell_repeats(const HostTable &t, int32_t H, int32_t P, int32_t N, int32_t S)
{
    for (int32_t c = 1; c < N; c++)
        for (int32_t j = 0; j < P; j++) {
            const EleHdr &a = t.hdr[H + j], &b = t.hdr[H + c * P + j];
            if (a.kind != b.kind || a.flags != b.flags || a.n != b.n || b.off != a.off + c * S)
                return false;
        }
    return true;
}

// The shortest cell (up to 64 elements) that repeats over at least three quarters of the
// lattice, starting at element 0 or after one differing head cell.
bool find_period(const HostTable &t, Period &out)
{
    const int32_t n = (int32_t)t.hdr.size();
    for (int32_t P = 1; P <= 64 && 4 * P <= n; P++) {
        for (int32_t H : {0, P}) {
            if (H + 2 * P > n) continue;
            const int32_t S = t.hdr[H + P].off - t.hdr[H].off;
            int32_t N = 1;
            while (H + (N + 1) * P <= n && cell_repeats(t, H, P, N + 1, S)) N++;
            if (N >= 4 && 4 * N * P >= 3 * n) {
                out = {H, P, N, S};
                return true;
            }
        }
    }
    return false;
}

std::string generate(const HostTable &t, const Period &pd, std::string &sig)
{
    bool q_kinds = false;  // the lattice holds stepped quadrupoles or general bends
    for (const EleHdr &h : t.hdr) q_kinds |= h.kind == K_QMULT || h.kind == K_BMULT;
    std::ostringstream body, key;
    key << pd.H << ':' << pd.P << ':' << pd.N << ':' << pd.S << ':' << t.hdr.size() << ':'
        << t.hdr[pd.H].off;
    for (int32_t j = 0; j < pd.P; j++) {
        const EleHdr &h = t.hdr[pd.H + j];
        const int32_t rel = h.off - t.hdr[pd.H].off;
        key << '/' << h.kind << ',' << h.flags << ',' << h.n << ',' << rel;
        // the rare set only where the element is a cavity, a quadrupole with multipoles or
        // carries offsets or momentum windows: elsewhere pz cannot change and the exit test for
        // rel_p < 0 drops out, as in the generic loop
        const char *maps = (h.kind == K_RF || (h.flags & (F_OFFSET | F_APE_MOMENTUM | F_RARE_MAP))) ? "kRare" : "MAPS_NO_RF";
        body << "            if (p.state == ST_ALIVE)\n"
             << "                track_element<true, " << maps << ">(p, EleRef{" << h.kind << ", " << h.flags
             << ", " << h.n << ", " << rel << "}, base, L.sp, true, e1 + " << j << ", t_abs);\n";
    }
    if (t.rare_q) key << "/q";
    if (t.rad) key << "/r";
    sig = key.str();
    std::ostringstream s;
    s << "#include <stdint.h>\n"
      << bmx_jit_src_table << '\n'
      << bmx_jit_src_maps << '\n'
      << bmx_jit_src_tiles << '\n'
      << "namespace bmx {\n"
      << "struct CellTurn {\n"
      << "    static constexpr int kRare = "
      << (t.rad ? (q_kinds ? "MAPS_RARE_QR" : "MAPS_RARE_R") : t.rare_q ? "MAPS_RARE_Q" : "MAPS_RARE") << ";\n"
      << "    __device__ __forceinline__ static void turn(Particle &p, const LatDesc &L, int32_t t_abs)\n"
      << "    {\n"
      << "        const CP prm = (CP)L.prm;\n";
    if (pd.H > 0) s << "        track_range<true, kRare>(p, L, 0, " << pd.H << ", t_abs);\n";
    s << "        for (int32_t c = 0; c < " << pd.N << "; ++c) {\n"
      << "            c = __builtin_amdgcn_readfirstlane(c);\n"
      << "            const CP base = prm + (" << t.hdr[pd.H].off << " + c * " << pd.S << ");\n"
      << "            const int32_t e1 = " << pd.H + 1 << " + c * " << pd.P << ";\n"
      << body.str()
      << "        }\n"
      << "        track_range<true, kRare>(p, L, " << pd.H + pd.N * pd.P << ", L.n_ele, t_abs);\n"
      << "    }\n"
      << "};\n"
      << "}  // namespace bmx\n";
    const char *args =
        "(const bmx::Group *__restrict__ g, const bmx::WaveTile *__restrict__ t, "
        "const uint32_t *__restrict__ nt, const double *__restrict__ pi, double *__restrict__ po, "
        "uint32_t ps, uint32_t *__restrict__ co, uint32_t *tc, int32_t off, int32_t K, "
        "unsigned long long *w)";
    // the stepped quadrupole and the general bend need the 2-wave budget; a lattice that is on
    // the stepped-element map sets only for its radiation kicks gets the variants asked for
    for (int w = 3; w <= 5; w++)
        s << "extern \"C\" __global__ void __launch_bounds__(bmx::kBlock) "
          << "__attribute__((amdgpu_waves_per_eu(" << (q_kinds ? 2 : w) << "))) bmx_cell_tiles_" << w << args
          << "\n{\n    bmx::track_tiles<true, bmx::CellTurn>(g, t, nt, pi, po, ps, co, tc, off, K, w);\n}\n";
    return s.str();
}

std::mutex g_mutex;
std::map<std::pair<int, std::string>, std::unique_ptr<JitKernels>> g_cache;

}  // namespace

int jit_specialize(int device, const HostTable &t, const JitKernels **out, std::string &err)
{
    *out = nullptr;
    Period pd;
    if (!find_period(t, pd)) {
        err = "no repeated cell covering three quarters of the lattice";
        return BMX_ERR_UNSUPPORTED;
    }
    std::string sig;
    const std::string src = generate(t, pd, sig);
    std::lock_guard<std::mutex> lock(g_mutex);
    auto &slot = g_cache[{device, sig}];
    if (slot) {
        *out = slot.get();
        return BMX_OK;
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "bmx_cell.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        err = "hiprtcCreateProgram failed";
        return BMX_ERR_HIP;
    }
    const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=fast",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        err = std::string("cell kernel compilation failed: ") + log.substr(0, 2000);
        hiprtcDestroyProgram(&prog);
        return BMX_ERR_HIP;
    }
    size_t size = 0;
    hiprtcGetCodeSize(prog, &size);
    std::vector<char> code(size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    auto k = std::make_unique<JitKernels>();
    k->rare_q = t.rare_q;
    k->level = t.rare_q ? (t.rad ? 2 : 1) : 0;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipModuleLoadData(&k->module, code.data());
    for (int w = 3; w <= 5 && e == hipSuccess; w++) {
        const std::string name = "bmx_cell_tiles_" + std::to_string(w);
        e = hipModuleGetFunction(&k->fn[w - 3], k->module, name.c_str());
    }
    if (e != hipSuccess) {
        err = std::string("cell kernel load: ") + hipGetErrorString(e);
        return BMX_ERR_HIP;
    }
    slot = std::move(k);
    *out = slot.get();
    return BMX_OK;
}

bool jit_rare_q(const JitKernels *j) { return j && j->rare_q; }
int jit_level(const JitKernels *j) { return j ? j->level : 0; }

hipError_t jit_occupancy(const JitKernels *j, int waves, int *blocks_per_cu)
{
    const int w = waves < 3 ? 3 : (waves > 5 ? 5 : waves);
    return hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, j->fn[w - 3], kTileBlock, 0);
}

hipError_t launch_track_tiles_jit(const JitKernels *j, hipStream_t s, int grid, int waves,
                                  const Group *d_groups, const WaveTile *d_tiles,
                                  const uint32_t *d_n_tiles, const double *d_pk_in,
                                  double *d_pk_out, uint32_t stride, uint32_t *d_cnt_out,
                                  uint32_t *d_tile_ctr, int32_t chunk_off, int32_t chunk_turns,
                                  unsigned long long *d_work)
{
    if (grid <= 0) return hipSuccess;
    const int w = waves < 3 ? 3 : (waves > 5 ? 5 : waves);
    void *args[] = {&d_groups, &d_tiles,   &d_n_tiles,  &d_pk_in,     &d_pk_out, &stride,
                    &d_cnt_out, &d_tile_ctr, &chunk_off, &chunk_turns, &d_work};
    return hipModuleLaunchKernel(j->fn[w - 3], (unsigned)grid, 1, 1, kTileBlock, 1, 1, 0, s, args,
                                 nullptr);
}

}  // namespace bmx
