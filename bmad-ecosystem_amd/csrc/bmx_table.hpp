// Device element table shared by the host table compiler and the kernels.
//
// A lattice is compiled once from raw bmx_ele records into a compact table: one 16-byte header
// per element plus a flat array of FP64 parameters holding only what the per-particle map of
// that element kind needs (element constants that Bmad re-derives per particle per element,
// e.g. multipole_ele_to_ab in track_a_quadrupole.f90:46-47 or the bend's cos/sin/sinc of the
// bend angle in track_a_bend.f90:165-175, are computed here once).
#pragma once
#include <stdint.h>

namespace bmx {

enum Kind : int32_t {
    K_DRIFT = 0,
    K_QUAD = 1,
    K_MULT = 2,   // thick multipole: kick - (drift - kick) * n_step (track_a_thick_multipole.f90)
    K_BEND = 3,
    K_RF = 4,
    K_MARKER = 5,
    // quadrupole with multipoles or kicks beside k1 and/or electric multipoles: n_step steps of
    // kick - quadrupole step - kick, electric hard-edge energy kicks at both ends
    // (track_a_quadrupole.f90:17-148)
    K_QMULT = 6,
    // sector bend with what K_BEND leaves out: offsets, pitches and ref_tilt (offset_particle's
    // bend branch), multipoles and kicks (dg from their b0 part), k1 (the k1 body map), electric
    // multipoles, the full edge (exact_bend_edge_kick); n_step steps of kick - body - kick
    // (track_a_bend.f90:18-333)
    K_BMULT = 7,
    // thin multipole (multipole$, ab_multipole$): kicks of any order 0..21 at zero length,
    // offset_particle without the tilt around them (track1_bmad.f90:142-168)
    K_THIN = 8,
    K_NKIND = 9,
};

enum Flags : int32_t {
    F_APE_ENT = 1 << 0,     // aperture checked at the entrance end
    F_APE_EXIT = 1 << 1,    // aperture checked at the exit end
    F_APE_ELLIP = 1 << 2,   // elliptical (else rectangular)
    F_EDGE_ENT = 1 << 3,    // Hwang edge kick at the bend entrance
    F_EDGE_EXIT = 1 << 4,   // Hwang edge kick at the bend exit
    F_EDGE_SIMPLE = 1 << 5, // both edges have e = 0, fint*hgap = 0, k1 = 0 (only the g y^2 terms)
    F_BEND_DRIFT = 1 << 6,  // g = dg = 0: the bend body is a drift (track_a_bend.f90:117-118)
    F_BEND_SERIES = 1 << 7, // g_tot = g and |angle| <= 0.1: the momentum form of the exact body
    F_ORDER_SHIFT = 8,      // K_MULT: bits 8..11 = mask of multipole orders 0..3 present
    // The entrance test repeats the exit test of the element before (same type and values,
    // same coordinates): a particle arriving from that element has passed it already. Set with
    // F_APE_ENT cleared; kernels that start at this element make the test.
    F_APE_ENT_SAME = 1 << 12,
    // The exit aperture fails for every |x| or |y| above orbit_too_large's limit (its limits all
    // lie well inside it), so that part of the exit validity test need not be gated separately.
    F_APE_EXIT_COVERS = 1 << 13,
    F_MULT_NORMAL = 1 << 14,  // K_MULT: no skew strengths (every a_n = 0)
    // Elliptical aperture centred on the axis (x0 = y0 = 0). Used where the flags are
    // compile-time constants (cell-specialised kernels); the table-driven loop masks it out.
    F_APE_CENTERED = 1 << 15,
    // offset_particle's non-bend branch around the element map (offset_particle.f90:183-492):
    // a block of kOffsetLen values precedes the aperture values. F_MISALIGN: offsets or pitches
    // (the frame change with the drift to the element edge); F_TILT: a tilt rotation.
    F_OFFSET = 1 << 16,
    F_MISALIGN = 1 << 17,
    F_TILT = 1 << 18,
    // collimator momentum windows: 8 values after the aperture values (px, py, z, pz half
    // width and centre); F_APE_MOM_ELLIP: the ecollimator's elliptical pairs
    F_APE_MOMENTUM = 1 << 19,
    F_APE_MOM_ELLIP = 1 << 20,
    // K_QMULT: bits 21..24 = mask of electric multipole orders 0..3 present (the magnetic
    // orders use bits 8..11 as for K_MULT)
    F_ELEC_SHIFT = 21,
    // An element whose map only the rare element path carries (K_QMULT): the loops send it
    // there as they do cavities and elements with offsets or momentum windows.
    F_RARE_MAP = 1 << 25,
    // K_QMULT edges (apply_element_edge_kick.f90:107-110,165-178,270-271): F_QEDGE_ENT /
    // F_QEDGE_EXIT: the element has edge fields (init_fringe_info.f90:44-54) and fringe_at puts
    // them at that end; F_QHARD: the hard multipole edge (fringe_type hard_edge_only or full);
    // F_QSOFT: the soft quadrupole edge (soft_edge_only or full). The electric edge comes with
    // F_QEDGE_* whenever electric orders are present.
    F_QEDGE_ENT = 1 << 26,
    F_QEDGE_EXIT = 1 << 27,
    F_QHARD = 1 << 28,
    F_QSOFT = 1 << 29,
    // K_BMULT (with F_QEDGE_* for the ends fringe_at names, F_EDGE_ENT / F_EDGE_EXIT where
    // bend_edge_kick acts, F_EDGE_SIMPLE, F_BEND_DRIFT, F_BEND_SERIES as for K_BEND, the order
    // masks as for K_QMULT, F_MISALIGN: the bend frame change, F_TILT: ref_tilt alone):
    F_BSEMI = F_MULT_NORMAL,  // dg = 0: the semi-linear body near the zero orbit applies
    F_BK1 = F_QSOFT,          // b1 /= 0: the k1 body map
    F_BEXACT = F_QHARD,       // fringe_type full: hard multipole edge and exact_bend_edge_kick
    // K_THIN: the dipole order of a multipole$ in the reference-orbit form (multipole_kick's
    // ref_orb_offset, multipole_mod.f90:250-254)
    F_THIN_REF0 = F_QHARD,
    // radiation damping and/or fluctuation kicks at both ends of the element (track1_radiation,
    // radiation_mod.f90:61-147): the element's radiation block follows its parameters, and the
    // element takes the rare path of the radiation map sets (F_RARE_MAP is set with it)
    F_RAD = 1 << 30,
};
constexpr int kOffsetLen = 24;

struct EleHdr {
    int32_t kind;
    int32_t flags;
    int32_t n;    // K_MULT: n_step; K_RF: n_slice
    int32_t off;  // offset of the element's parameters in the parameter array (doubles)
};

// Parameter layouts (index into prm + off):
//  K_DRIFT : [0] L
//  K_QUAD  : [0] L, [1] b1 = k1 L, [2] b1 L, [3] b1 / L, [4] t_lim (fast form's series
//            domain; hdr.n = number of series terms used)
//  K_MULT  : [0] step_len, [1..4] a_n * r_step/2 (n = 0..3), [5..8] b_n * r_step/2
//  K_BEND  : [0] L, [1] g, [2] g_tot, [3] angle = g L, [4] cos_a, [5] sin_a, [6] sinc_a,
//            [7] cosc_a, [8] beta_ref, [9] sincc_a,
//            [10..15] entrance gt, gt2, gs2, k1_tane, fg_factor, tan_e^2, [16..21] exit,
//            [22] L^2 g cosc_a, [23] L / beta_ref, [24] L sinc_a, [25] gs2/2 entrance,
//            [26] gs2/2 exit, [27] 1 - cos_a, [28] 1 / g
//  K_RF    : [0] L/n_slice, [1] voltage * charge_dir, [2] phi0, [3] rf_frequency,
//            [4] 0.5/rf_frequency, [5] twopi * rf_frequency, [6] twopi * phi0_multipass
//  K_MARKER: (none)
//  K_QMULT : [0..4] one step as a K_QUAD block: step_len s, b1 s/L, b1 s^2/L, b1/L, t_lim;
//            [5..8] a_n * r_step/2, [9..12] b_n * r_step/2 (magnetic, n = 0..3, b1 removed);
//            [13..16] a_n * charge/p0c * s/2, [17..20] -b_n * charge/p0c * s/2 (electric
//            kick strengths without the particle's 1/beta); [21..24] charge a_n/(n+1),
//            [25..28] charge b_n/(n+1) (electric edge potential); [29] k1/12 (hard edge,
//            n = 1), [30] k1 fq1, [31] k1 fq2 (soft edge); hdr.n = n_step
//  K_BMULT : [0..28] one body step as a K_BEND block (L = step_len, g_tot = g + dg with dg
//            from the multipoles' b0, edge constants of the element); [29] g + the element's dg
//            (the edges' g_tot); [30] k1 = b1/L; [31..34] a_n r_step/2, [35..38] b_n r_step/2
//            (magnetic); [39..42] a_n charge/p0c step_len/2, [43..46] -b_n charge/p0c
//            step_len/2 (electric field strengths); [47..50] charge a_n/(n+1), [51..54]
//            charge b_n/(n+1) (electric edge potential); [55] k1/12 (hard multipole edge);
//            entrance [56] cos e1, [57] sin e1, [58] sin 2 e1, [59] fint hgap, exit [60..63]
//            likewise with e2, fintx hgapx; [64] 1/beta_ref^2; [65] e1, [66] e2; frame change
//            [67..75] W (entering, row-major), [76..78] d (r = W ((x, y, 0) - d)), [79..87] W
//            (leaving), [88..90] c (r = W (x, y, 0) + c); [91] cos ref_tilt, [92] sin
//            ref_tilt, [93] L; hdr.n = n_step
//  K_THIN  : [0] k0l cos t0, [1] k0l sin t0 (the reference-orbit dipole, F_THIN_REF0),
//            [2 .. 2+N) a_n, [2+N .. 2+2N) b_n (orders 0 .. N-1 as a/b strengths, the order
//            tilts folded in); hdr.n = N (0..22), so the block holds 2 + 2N values
// preceded, when F_OFFSET, by the offset block (hdr.off points at it; then the aperture values,
// then the parameters):
//  [0] x_offset, [1] y_offset, [2] z_offset, [3] L/2, [4..12] W^-1 (entrance frame change,
//  row-major), [13..21] W (exit), [22] cos tilt, [23] sin tilt
// and, when F_APE_ENT|F_APE_EXIT, by 4 aperture values (hdr.off points at them otherwise), then
// when F_APE_MOMENTUM by 8 momentum-window values:
//  elliptical : x_center, 1/x_width2, y_center, 1/y_width2
//  rectangular: x1_lim (= -x1_limit), x2_lim, y1_lim, y2_lim (0 = no limit)
// Parameter count per kind (K_THIN: 2 + 2 hdr.n, here its minimum)
constexpr int kParams[K_NKIND] = {1, 5, 9, 29, 7, 0, 32, 94, 2};
// With F_RAD, the element's parameters are followed by its radiation block: the maps of the
// upstream and the downstream half (rad_map_ele_struct rm0, rm1, bmad_struct.f90:988-1000), each
// kRadEdge values: [0..5] ref_orb, [6..41] damp_dmat (row-major), [42..47] xfer_damp_vec,
// [48..68] stoc_mat's lower triangle row by row (row i holds columns 0..i).
constexpr int kRadEdge = 69;
constexpr int kRadLen = 2 * kRadEdge;

// Radiation settings of a lattice (bmad_com%radiation_damping_on, %radiation_fluctuations_on,
// %radiation_zero_average, %synch_rad_scale), read by the kicks of F_RAD elements.
enum : uint32_t { RAD_DAMP = 1, RAD_FLUCT = 2, RAD_ZERO_AVERAGE = 4 };
struct RadCfg {
    double scale;       // synch_rad_scale
    uint32_t flags;     // RAD_*
    uint32_t key0, key1;  // the fluctuation stream's key (the seed)
    uint32_t pad;
};

struct Species {
    double mc2;      // [eV]
    double p0c;      // reference momentum [eV]
    double mc2_sq;   // mc2^2
    // low_energy_z_correction.f90:22-58 constants of the (single) reference energy
    double beta0;    // p0c / E_tot
    double me2;      // (mc2 / E_tot)^2
    double lez_f;    // beta0^2 (2 beta0^2 - (mc2/E_tot)^2 / 2)
    double lez_thr;  // 3e-7 E_tot
    RadCfg rad;      // radiation settings of the lattice (only F_RAD elements read them)
};

// Per-lattice device descriptor.
struct LatDesc {
    const EleHdr *hdr;
    const double *prm;
    int32_t n_ele;
    int32_t rare_q;  // 1 when the table holds K_QMULT or K_BMULT elements, 2 when it also has
                     // radiation kicks, else 0 (read on the host only)
    Species sp;
};

// Device pointers of one beam (SoA).
struct BeamPtrs {
    double *x, *px, *y, *py, *z, *pz, *beta;
    int32_t *state, *loss_ele, *loss_turn;
};

// A group: particles [0, n) of one beam view tracked through one lattice for turns
// turn0 .. turn0 + n_turns - 1. A single-lattice call is one group; batched variants and served
// requests are one group each. list_off locates the group's segment in the packed particle
// buffers of the run (Packed below).
// Each turn passes through `count` elements starting at element `first` (0-based) and wrapping
// past the last element to element 0: first = 0, count = n_ele is the full turn from the start
// of the lattice; a turn that starts after element s (LTT's track_bunch(ele_start, ele_start),
// lt_tracking_mod.f90:3223) is first = s, count = n_ele; track_bunch's ele1 -> ele2 span
// (beam_mod.f90:148-166) is one turn of first = ix_ele1, count < n_ele possibly.
struct Group {
    LatDesc lat;
    BeamPtrs b;
    uint32_t n;
    uint32_t list_off;
    int32_t n_turns;
    int32_t turn0;
    int32_t first;
    int32_t count;
};

// Slots of a particle-element pass counter: kernels add to slot blockIdx.x % kWorkSlots and the
// host reads their sum.
constexpr int kWorkSlots = 16;

// The live particles of a run between two chunks, packed: slot j of the run holds coordinate k
// (0..6: x, px, y, py, z, pz, beta) at d[k * stride + j] and the particle's index in its beam
// at ((uint32_t *)(d + kPackedDoubles * stride))[j]. Each group owns slots list_off ..
// list_off + n - 1; a chunk reads the live particles from one such buffer and appends the ones
// that survive it to the other, so every chunk reads and writes whole cache lines however many
// of the beam's particles are lost.
constexpr int kPackedDoubles = 7;
constexpr uint64_t kPackedBytes = kPackedDoubles * sizeof(double) + sizeof(uint32_t);

// One wave's unit of work: up to 64 particles of one group for one chunk of turns. start
// indexes the group's slots in the packed buffer (or the particles of its beam when the run
// packs its live particles before the first chunk).
struct WaveTile {
    int32_t group;
    uint32_t start;
    uint32_t count;
    uint32_t pad;
};

}  // namespace bmx
