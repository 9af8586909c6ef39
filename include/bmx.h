/*
 * bmx.h -- C-ABI of the MI355X (gfx950) bmad_standard multi-particle tracking library.
 *
 * The library replaces the inner loops of Bmad's beam tracking:
 *   bmad/multiparticle/beam_mod.f90:40   track_beam   (lattice loop over elements, per bunch)
 *   bmad/multiparticle/beam_mod.f90:103  track_bunch
 *   bmad/multiparticle/beam_mod.f90:238  track1_bunch (one element, whole bunch)
 *   bmad/multiparticle/beam_utils.f90:32 track1_bunch_hom (OpenMP loop over particles, :81-86)
 *   bmad/code/track1.f90:44              track1       (one particle, one element)
 *   bmad/code/track_all.f90:25           track_all    (one particle, all elements, stop on loss)
 *   bsim/modules/lt_tracking_mod.f90:1571 ltt_run_beam_mode (turn loop)
 * for the element kinds drift / quadrupole / sextupole / sbend (Hwang edge) / rfcavity / marker
 * and elliptical or rectangular apertures, all arithmetic in FP64.
 *
 * Conventions (mirroring the reference):
 *  - Phase-space coordinates are Bmad's (x, px, y, py, z, pz) with px, py, pz normalised by the
 *    reference momentum p0c (bmad/modules/bmad_struct.f90:551-582, coord_struct%vec).
 *  - Particle loss is data, not an error (bmad/code/track1.f90:39-41): it is recorded in state[]
 *    with Bmad's codes (bmad_struct.f90:527-532), the element index in loss_ele[] (ix_ele,
 *    track1.f90:88-94, 1-based like lat%branch%ele(1:n)) and the 0-based turn in loss_turn[]
 *    (lt_tracking_mod.f90:1395).
 *  - Every entry point returns BMX_OK (0) or a positive error code; bmx_last_error() gives the
 *    message of the last failure on the calling thread. Nothing aborts or prints.
 *  - All arrays passed to the upload, download and submit calls are caller-owned host memory; arrays
 *    passed to bmx_beam_wrap are caller-owned device memory that must outlive the handle.
 *  - A handle is not reentrant: one context per host thread (the server owns its own thread).
 */
#ifndef BMX_H
#define BMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BMX_VERSION 1

/* ---- status codes ---------------------------------------------------------------------- */
#define BMX_OK              0
#define BMX_ERR_ARG         1  /* bad argument (null handle, size mismatch, ...) */
#define BMX_ERR_UNSUPPORTED 2  /* element kind or attribute outside this path (track1_bmad.f90:263-270) */
#define BMX_ERR_HIP         3  /* HIP runtime failure */
#define BMX_ERR_P0C         4  /* reference momentum mismatch (track1.f90:151-169) */
#define BMX_ERR_NOMEM       5  /* device allocation failed */
#define BMX_ERR_BUDGET      6  /* request larger than the server's HBM budget */
#define BMX_ERR_STATE       7  /* call not valid in the handle's current state */
#define BMX_ERR_NOT_FOUND   8  /* unknown request id */
#define BMX_ERR_UNSTABLE    9  /* one-turn matrix with an eigenvalue off the unit circle: no tune */

/* ---- coord_struct%state codes (bmad_struct.f90:527-532) -------------------------------- */
#define BMX_STATE_ALIVE      1
#define BMX_STATE_LOST       2
#define BMX_STATE_LOST_NEG_X 3
#define BMX_STATE_LOST_POS_X 4
#define BMX_STATE_LOST_NEG_Y 5
#define BMX_STATE_LOST_POS_Y 6
#define BMX_STATE_LOST_Z     7
#define BMX_STATE_LOST_PZ    8

/* ---- element keys (bmad_struct.f90:1696-1710) ------------------------------------------ */
#define BMX_DRIFT            1
#define BMX_SBEND            2
#define BMX_QUADRUPOLE       3
#define BMX_SEXTUPOLE        5
#define BMX_RFCAVITY         9
#define BMX_MARKER          14
#define BMX_KICKER          15
#define BMX_OCTUPOLE        17
#define BMX_MULTIPOLE       19  /* thin: a_pole = knl, b_pole = tn, a_pole_elec = ksl */
#define BMX_AB_MULTIPOLE    22  /* thin: a_pole, b_pole */
#define BMX_MONITOR         31
#define BMX_INSTRUMENT      32
#define BMX_HKICKER         33
#define BMX_VKICKER         34
#define BMX_RCOLLIMATOR     35
#define BMX_ECOLLIMATOR     36
#define BMX_PIPE            44
#define BMX_THICK_MULTIPOLE 67

/* fringe_type (bmad_struct.f90:27,168-169) */
#define BMX_FRINGE_NONE           1
#define BMX_FRINGE_SOFT_EDGE_ONLY 2
#define BMX_FRINGE_HARD_EDGE_ONLY 3
#define BMX_FRINGE_FULL           4
#define BMX_FRINGE_BASIC_BEND     7
/* fringe_at / aperture_at (bmad_struct.f90:373-374) */
#define BMX_ENTRANCE_END 1
#define BMX_EXIT_END     2
#define BMX_BOTH_ENDS    3
#define BMX_NO_END       4
#define BMX_CONTINUOUS   5
/* aperture_type (bmad_struct.f90:158) */
#define BMX_APERTURE_RECTANGULAR 2
#define BMX_APERTURE_ELLIPTICAL  3

#define BMX_N_POLE 22  /* n_pole_maxx + 1: multipole orders 0..21 */

/*
 * bmx_ele -- one lattice element, as the raw attribute values of ele_struct
 * (bmad_struct.f90:1422-1530, ele%value(:) at :1469). A Fortran host fills it straight from
 * lat%branch(ix)%ele(i)%value(...) after Bmad's own bookkeeping; ds_step is the bookkept value
 * (attribute_bookkeeper.f90:292-398) and p0c / e_tot the element's reference momentum / energy.
 * Zero means "not set" for every optional attribute, as in ele%value.
 */
typedef struct bmx_ele {
    int32_t key;                   /* BMX_DRIFT ... */
    int32_t is_on;                 /* ele%is_on */
    int32_t fringe_type;           /* value(fringe_type$) */
    int32_t fringe_at;             /* value(fringe_at$) */
    int32_t aperture_type;         /* ele%aperture_type */
    int32_t aperture_at;           /* ele%aperture_at */
    int32_t offset_moves_aperture; /* ele%offset_moves_aperture */
    int32_t n_multipole;           /* 0: a_pole/b_pole not associated; else number of orders used */
    double l;
    double k1, k2, k3;
    double g, dg, e1, e2, fint, fintx, hgap, hgapx;
    double hkick, vkick, kick;
    double tilt, ref_tilt;
    double x_offset, y_offset, z_offset, x_pitch, y_pitch;
    double voltage, rf_frequency, phi0, phi0_multipass;
    double x1_limit, x2_limit, y1_limit, y2_limit;
    double px_aperture_width2, px_aperture_center, py_aperture_width2, py_aperture_center;
    double z_aperture_width2, z_aperture_center, pz_aperture_width2, pz_aperture_center;
    double ds_step;
    double p0c, e_tot;
    double r0_mag;
    double a_pole[BMX_N_POLE];
    double b_pole[BMX_N_POLE];
    /* Electric multipoles: ele%a_pole_elec / %b_pole_elec (bmad_struct.f90:1483-1484) and
     * value(r0_elec$); n_multipole_elec = 0 when they are not associated. */
    double r0_elec;
    double a_pole_elec[BMX_N_POLE];
    double b_pole_elec[BMX_N_POLE];
    int32_t n_multipole_elec;
    /* value(exact_multipoles$) of a bend: 1 off (0 is taken as off), 2 horizontally_pure,
     * 3 vertically_pure (bmad_struct.f90:256-261) */
    int32_t exact_multipoles;
    /* value(fq1$), value(fq2$): soft quadrupole edge coefficients (fringe_mod.f90:556-661) */
    double fq1, fq2;
} bmx_ele;

/*
 * bmx_beam_view -- structure-of-arrays particle state: coord_struct%vec(1:6), %beta, %state
 * (bmad_struct.f90:551-582) plus the loss bookkeeping of track1.f90:88-94 / lt_tracking_mod.f90:1395.
 * Depending on the call the pointers are host or device addresses.
 */
typedef struct bmx_beam_view {
    double *x, *px, *y, *py, *z, *pz, *beta;
    int32_t *state, *loss_ele, *loss_turn;
} bmx_beam_view;

typedef struct bmx_ctx bmx_ctx;
typedef struct bmx_lattice bmx_lattice;
typedef struct bmx_beam bmx_beam;
typedef struct bmx_batch bmx_batch;
typedef struct bmx_server bmx_server;

/* ---- library ----------------------------------------------------------------------------- */
/* Library version (BMX_VERSION). */
int32_t bmx_version(void);
/* Message of the last failed call on this thread (empty if none). Replaces out_io error
 * reporting (sim_utils/io/output_mod.f90:198). */
const char *bmx_last_error(void);

/* ---- context: one device + one HIP stream ------------------------------------------------ */
int32_t bmx_ctx_create(int32_t device, bmx_ctx **out);
int32_t bmx_ctx_destroy(bmx_ctx *ctx);
int32_t bmx_ctx_synchronize(bmx_ctx *ctx);
/* The context's HIP stream (hipStream_t), for callers that order their own work against it. */
void *bmx_ctx_stream(bmx_ctx *ctx);
/* Particle-element passes tracked by this context's tracking calls since the last take
 * (element passes of particles alive at the element entrance, the work track1_bunch_hom
 * does, beam_utils.f90:81-86); synchronises the stream and resets the count. */
int32_t bmx_ctx_take_work(bmx_ctx *ctx, int64_t *passes);
/* Formulation of the per-particle maps used by this context's tracking calls:
 * BMX_MAPS_FAST (default) re-expresses the drift, quadrupole and bend bodies without
 * per-particle transcendental calls (convergent series, angle-addition identities, cached
 * reciprocals); BMX_MAPS_REFERENCE evaluates the reference's formulas as written
 * (track_a_drift.f90, quad_mat2_calc.f90, track_a_bend.f90:163-209). Both agree to rounding. */
#define BMX_MAPS_FAST 1
#define BMX_MAPS_REFERENCE 2
int32_t bmx_ctx_set_maps(bmx_ctx *ctx, int32_t form);
/* Run shape of this context's bmx_track_turns calls: turns per chunk (alive particles are
 * compacted between chunks; 0 = the whole call is one chunk; default 5), workgroups of the
 * persistent tracking grid (0 = as many as fit on the device at once, the default) and the
 * waves per SIMD the tracking kernel variant is compiled for at least (3, 4 or 5; 0 = chosen
 * per call: for a cell-specialised lattice 5 while the live particles fill more than 1.45
 * rounds of the 4-wave kernel's resident waves and 4 below (the run phase, from a count of the
 * live particles at the start of the call), else 5 with the fast maps, 3 with the reference
 * maps). */
int32_t bmx_ctx_set_run(bmx_ctx *ctx, int32_t chunk_turns, int32_t grid, int32_t waves);
/* The run shape (turns per chunk, workgroups, waves per SIMD) the context's last
 * bmx_track_turns / bmx_track_turns_from / bmx_track_bunch call used. No counterpart in the
 * reference (its tracking has no run shape); for reports and autotuning. */
int32_t bmx_ctx_last_run(const bmx_ctx *ctx, int32_t *chunk_turns, int32_t *grid, int32_t *waves);
/* Kernel timing of bmx_track_turns: while on, every launch of the tracking kernel is bracketed
 * by HIP events on the context's stream; bmx_ctx_kernel_time synchronises and returns the sum
 * of those launch durations and their number since the last call. */
int32_t bmx_ctx_profile(bmx_ctx *ctx, int32_t on);
int32_t bmx_ctx_kernel_time(bmx_ctx *ctx, double *ms, int64_t *launches);
/* Timing on the context's stream: record HIP event `slot` (0..BMX_N_EVENTS-1) now, and the
 * elapsed milliseconds between two recorded slots (synchronises on the later one). Replaces
 * the wall-clock run_timer of the reference drivers (sim_utils/misc/run_timer.f90:32). */
#define BMX_N_EVENTS 64
int32_t bmx_ctx_record(bmx_ctx *ctx, int32_t slot);
int32_t bmx_ctx_elapsed_ms(bmx_ctx *ctx, int32_t slot0, int32_t slot1, float *ms);

/* ---- lattice: flattened element table (lat%branch(ix)%ele(1:n_ele_track)) ----------------- *
 * Replaces the per-particle re-derivation of element constants in track_a_* (e.g.
 * multipole_ele_to_ab calls in track_a_quadrupole.f90:46-47, track_a_bend.f90:58-77) by one
 * compilation of the table, uploaded once. mc2 [eV] and charge are the tracked species
 * (particle_species_mod.f90:1498,1550); p0c of every element must equal the beam's p0c.      */
int32_t bmx_lattice_create(bmx_ctx *ctx, const bmx_ele *eles, int32_t n_ele, double mc2,
                           int32_t charge, bmx_lattice **out);
int32_t bmx_lattice_destroy(bmx_lattice *lat);
/* Compile the element table on the host only (no device needed) and report whether every
 * element is inside this path: BMX_OK, or BMX_ERR_UNSUPPORTED / BMX_ERR_P0C / BMX_ERR_ARG with
 * the offending element in bmx_last_error(). n_params receives the table size in doubles. */
int32_t bmx_lattice_validate(const bmx_ele *eles, int32_t n_ele, double mc2, int32_t charge,
                             int32_t *n_params);
/* Compile, at run time for this device, tile kernels specialised to the lattice's repeated
 * cell: the shortest run of elements whose kinds, flags and parameter-block sizes repeat over at
 * least three quarters of the table, written out once with every kind dispatch and flag test
 * resolved. Strengths stay in the table, so lattices differing only in strengths (error
 * variants) share one compiled kernel. Used by the fast-map runs of bmx_track_turns and
 * bmx_batch_track; the maps are the generic kernel's, so results agree with it to rounding
 * (losses identical in the parity tests). BMX_ERR_UNSUPPORTED when no
 * such cell exists (the generic kernel is then used). No counterpart in the reference, whose
 * element loop is track_all (bmad/code/track_all.f90:1) over the generic element dispatch. */
int32_t bmx_lattice_specialize(bmx_lattice *lat);
/* Number of elements and bit set of element kinds present (bit = key). */
int32_t bmx_lattice_info(const bmx_lattice *lat, int32_t *n_ele, uint64_t *kind_mask);

/* ---- radiation: damping and fluctuation kicks at the element ends -------------------------- *
 * bmx_rad_map -- one half of an element's radiation map, rad_map_struct (bmad_struct.f90:988-994)
 * as radiation_map_setup leaves it in ele%rad_map%rm0 (upstream half) and %rm1 (downstream
 * half), row-major here: damp_dmat[i][j] = damp_dmat(i+1, j+1). stoc_mat is the lower
 * triangular kick matrix (tracking_rad_map_setup.f90:71-82); its upper triangle is not read.
 * xfer_damp_mat is not used by the kick and has no field. */
typedef struct bmx_rad_map {
    double ref_orb[6];
    double damp_dmat[6][6];
    double xfer_damp_vec[6];
    double stoc_mat[6][6];
} bmx_rad_map;

/* bmad_com%radiation_damping_on, %radiation_fluctuations_on, %radiation_zero_average and
 * %synch_rad_scale (bmad_struct.f90), and the seed of the fluctuation stream. The fluctuation
 * kicks of a particle are a function of (seed, the particle's index, turn, element, edge), the
 * index counted from the first particle of what one call tracks (the beam; a bmx_batch group's
 * range; a bmx_server request): a call repeats the kicks of an earlier call over the same turns,
 * so a run split over several calls continues the turn count (bmx_track_turns_from) or changes
 * the seed. The deviates are Philox4x32-10 blocks under the seed with the counter (index, turn,
 * element 1-based, 4 edge + block), turned into normal deviates by Box-Muller. The reference
 * draws them from one sequential generator (ran_gauss), so its kicks are not reproduced number
 * for number; their distribution is the same. */
typedef struct bmx_rad_config {
    int32_t damping_on;
    int32_t fluctuations_on;
    int32_t zero_average;
    int32_t reserved;        /* 0 */
    double synch_rad_scale;  /* 1 in the reference's default */
    uint64_t seed;
} bmx_rad_config;

/* bmx_lattice_create with radiation: maps[2 i] and maps[2 i + 1] are rm0 and rm1 of element i.
 * track1_radiation (radiation_mod.f90:61-147) is applied at both ends of every element that
 * radiates in the reference -- is_on, l /= 0 and not a drift, pipe, marker, multipole or
 * ab_multipole (radiation_mod.f90:88-93) -- and whose maps are not all zero, in track1's order
 * (track1.f90:197-202, 291-326): the upstream half's kick after the entrance aperture test, the
 * downstream half's after the body's orbit test and before the exit aperture test. A particle
 * whose pz falls below -1 is lost (lost_pz). Elements that radiate take the generic kernel's
 * rare path (a cell-specialised kernel writes them out like any other element). */
int32_t bmx_lattice_create_rad(bmx_ctx *ctx, const bmx_ele *eles, int32_t n_ele, double mc2,
                               int32_t charge, const bmx_rad_map *maps,
                               const bmx_rad_config *config, bmx_lattice **out);
/* Change the radiation settings (switches, scale, seed) of a lattice without recompiling it:
 * tracking calls made after it use them (a bmx_batch or bmx_server request keeps the settings
 * its lattice had when it was created or submitted). */
int32_t bmx_lattice_set_radiation(bmx_lattice *lat, const bmx_rad_config *config);

/* ---- beam: particle state resident in HBM --------------------------------------------------- */
/* Library-owned device storage for n particles. */
int32_t bmx_beam_create(bmx_ctx *ctx, int64_t n, bmx_beam **out);
/* Caller-owned device storage (e.g. torch tensors); no copy is made. */
int32_t bmx_beam_wrap(bmx_ctx *ctx, const bmx_beam_view *device, int64_t n, bmx_beam **out);
/* Host -> device and device -> host copies of the first n particles (blocking). */
int32_t bmx_beam_upload(bmx_beam *beam, const bmx_beam_view *host, int64_t n);
int32_t bmx_beam_download(bmx_beam *beam, const bmx_beam_view *host, int64_t n);
int32_t bmx_beam_destroy(bmx_beam *beam);
/* ---- beam files (bmad/multiparticle/beam_file_io.f90), host memory, no device needed ------ *
 * Read Bmad's ASCII beam files: the old "!ASCII::3" layout (read_beam_file, :380-804) or the
 * column layout (read_beam_ascii, :834-1040). bmx_beam_file_info gives the bunch count and the
 * first bunch's particle count. bmx_beam_file_read fills bunch ix_bunch (1-based) into host
 * arrays of capacity cap: x..pz and state from the file, loss_ele = the file's ix_ele for lost
 * particles (0 for live ones), loss_turn = 0, beta from a beta column or else from pz, p0c and
 * mc2 as init_coord sets it; n_out and charge_tot (sum of particle charges or the file's
 * charge_tot) are optional. */
int32_t bmx_beam_file_info(const char *path, int32_t *n_bunch, int64_t *n_particle);
int32_t bmx_beam_file_read(const char *path, int32_t ix_bunch, const bmx_beam_view *host, int64_t cap,
                           double p0c, double mc2, int64_t *n_out, double *charge_tot);
/* Write n particles as one bunch in the column layout of write_ascii_beam_file (:123-350):
 * constant columns as '#' parameters (charge per particle, species, location Downstream_End,
 * ix_branch 0, direction and time_dir 1), then index, x..pz, p0c, beta, state name and ix_ele (the
 * given ix_ele for live particles, loss_ele for lost ones). Coordinates carry 17 significant
 * digits (the reference writes 16), so reading a written file back is bit-identical: a
 * tracking checkpoint. append != 0 adds the bunch to an existing file. */
int32_t bmx_beam_file_write(const char *path, const bmx_beam_view *host, int64_t n, const char *species,
                            double charge, double p0c, int32_t ix_ele, int32_t append);
/* openPMD HDF5 beam files of charged particles (bmad/hdf5/hdf5_read_beam.f90,
 * hdf5_write_beam.f90), host memory, no device needed. libhdf5 (1.10) is loaded at the first
 * call: BMX_HDF5_LIB if set, else libhdf5.so.103 / libhdf5.so, else /opt/conda/lib;
 * BMX_ERR_UNSUPPORTED when none loads. bmx_beam_h5_info: bunch count, the first bunch's particle
 * count and species name (speciesType; species may be null). bmx_beam_h5_read: bunch ix_bunch
 * (1-based), the newer layout (a species group under particlesPath) or the older one
 * (particlesPath holds the records); positions and positionOffset, momenta in eV/c and their
 * offsets, time, totalMomentum and totalMomentumOffset (p0c, which must be one value), state,
 * elementIndex (loss_ele of lost particles), as hdf5_read_beam converts them without a lattice
 * element: px = Px / p0c, pz = totalMomentum / p0c, beta from (1 + pz) p0c and mc2,
 * z = -beta c (t - t_ref); a file not written by Bmad (or this library) has its non-alive
 * states read as lost. charge_tot = totalCharge * chargeUnitSI. Photon bunches are refused
 * (BMX_ERR_UNSUPPORTED). bmx_beam_h5_write: n particles as one bunch /data/NNNNN/particles/
 * <species>/ in hdf5_write_beam's layout (equal values as constant records), the absolute time
 * (not carried here) as t_ref = -(t - t_ref); append != 0 adds the bunch to an existing file.
 * Reading a written file back gives x and y exactly and the other coordinates to within a few
 * ulp (px, py and pz pass through px p0c, py p0c and pz p0c, z through -z / (beta c)); the ASCII
 * writer is the bit-identical checkpoint. */
int32_t bmx_beam_h5_info(const char *path, int32_t *n_bunch, int64_t *n_particle, char *species,
                         int32_t species_len);
int32_t bmx_beam_h5_read(const char *path, int32_t ix_bunch, const bmx_beam_view *host, int64_t cap,
                         double mc2, int64_t *n_out, double *p0c, double *charge_tot);
int32_t bmx_beam_h5_write(const char *path, const bmx_beam_view *host, int64_t n, const char *species,
                          double charge, double p0c, int32_t ix_ele, int32_t append);
/* Count of particles with state == alive (lt_tracking_mod.f90:1672, bunch%n_live). */
int32_t bmx_beam_count_alive(bmx_beam *beam, int64_t *n_alive);

/* ---- tracking ---------------------------------------------------------------------------- */
/* Persistent tracking: n_turns full turns starting at element 1, turn index turn0 .. turn0+n_turns-1.
 * Equivalent of ltt_run_beam_mode's turn loop (lt_tracking_mod.f90:1631-1716) calling
 * track_bunch (beam_mod.f90:103) once per turn. State stays in registers across elements and
 * the turns of a chunk; between chunks lost particles are dropped from the work list.
 * At most 2^32-1 particles. Asynchronous on the context's stream. */
int32_t bmx_track_turns(bmx_ctx *ctx, const bmx_lattice *lat, bmx_beam *beam, int32_t n_turns,
                        int32_t turn0);
/* The same turn loop with every turn starting after element ix_ele_start (0..n_ele; 0 and
 * n_ele are bmx_track_turns): elements ix_ele_start+1 .. n_ele, then 1 .. ix_ele_start, all
 * with the turn's index. This is LTT's turn, track_bunch(lat, bunch, ele_start, ele_start)
 * (bsim/modules/lt_tracking_mod.f90:3223, called from the turn loop at :1652). A particle that
 * resumes at element ix_ele_start+1 has its entrance aperture tested there. Asynchronous. */
int32_t bmx_track_turns_from(bmx_ctx *ctx, const bmx_lattice *lat, bmx_beam *beam,
                             int32_t ix_ele_start, int32_t n_turns, int32_t turn0);
/* track_bunch (bmad/multiparticle/beam_mod.f90:103, forward direction): the bunch at the end of
 * element ix_ele1 (0..n_ele; not tracked through) to the end of element ix_ele2 (0..n_ele):
 * elements ix_ele1+1 .. ix_ele2, or, when ix_ele2 <= ix_ele1, ix_ele1+1 .. n_ele then
 * 1 .. ix_ele2 (beam_mod.f90:148-166). Losses record loss_turn = turn. Persistent kernel;
 * asynchronous. */
int32_t bmx_track_bunch(bmx_ctx *ctx, const bmx_lattice *lat, bmx_beam *beam, int32_t ix_ele1,
                        int32_t ix_ele2, int32_t turn);
/* bmx_track_turns without any device allocation. No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Its run buffers (packed particle buffers, wave tiles,
 * counters) are a fixed device array of about 30 MiB per GPU, held until the process ends;
 * beams of more than 2^18 particles are refused with BMX_ERR_ARG. All calls on one GPU share
 * that array, so calls that overlap on one GPU (other contexts, other host threads) race: the
 * caller serialises them. Same results as bmx_track_turns. Asynchronous. */
int32_t bmx_track_turns_noalloc(bmx_ctx *ctx, const bmx_lattice *lat, bmx_beam *beam,
                                int32_t n_turns, int32_t turn0);
/* One element for the whole bunch: track1_bunch_hom (beam_utils.f90:32-90) on the GPU, state
 * read from and written back to HBM. ix_ele is 1-based. Asynchronous. */
int32_t bmx_track_element(bmx_ctx *ctx, const bmx_lattice *lat, bmx_beam *beam, int32_t ix_ele,
                          int32_t turn);
/* Plain path: n_turns x n_ele launches of the per-element kernel (the reference's loop nest:
 * turn -> element -> particles, beam_mod.f90:159-166 + beam_utils.f90:81-86), state through HBM
 * at every element. One turn of launches is captured as a HIP graph (kept in the context for
 * the same lattice, beam and maps) and replayed per turn. Asynchronous. */
int32_t bmx_track_turns_plain(bmx_ctx *ctx, const bmx_lattice *lat, bmx_beam *beam,
                              int32_t n_turns, int32_t turn0);

/* ---- batches: many lattices and particle sets in one launch sequence ------------------------ *
 * A group is a range of one beam tracked through one lattice. A batch tracks all of its groups
 * together: their particles share the persistent grid wave by wave (every wave works on one
 * group at a time), lost particles are compacted out between chunks, and the whole sequence
 * of chunks is captured once as a hipGraph and replayed by every bmx_batch_track call with the
 * same turn count. Groups may share beams (disjoint ranges) and lattices. The batch keeps the
 * handles; they must outlive it. */
typedef struct bmx_group {
    const bmx_lattice *lat;
    bmx_beam *beam;
    int64_t first;  /* first particle of the range */
    int64_t count;  /* particles in the range */
} bmx_group;
int32_t bmx_batch_create(bmx_ctx *ctx, const bmx_group *groups, int32_t n_groups, bmx_batch **out);
int32_t bmx_batch_destroy(bmx_batch *batch);
/* Every group n_turns turns, turn indices turn0 .. turn0+n_turns-1 (asynchronous). */
int32_t bmx_batch_track(bmx_batch *batch, int32_t n_turns, int32_t turn0);
/* Run shape of the batch: turns per chunk, persistent grid (0 = device limit) and the waves
 * per SIMD the tracking kernel is compiled for at least (3, 4 or 5; 0 = the default: 3 for
 * cell-specialised lattices, else 5 with the fast maps, 3 with the reference maps). */
int32_t bmx_batch_set_run(bmx_batch *batch, int32_t chunk_turns, int32_t grid, int32_t waves);
/* Autotuner: times sample_turns turns of every candidate run shape (chunk length x kernel
 * occupancy x grid size) on a scratch copy of the batch's particles (the batch's beams are not
 * modified), keeps the fastest for the batch and returns it. Blocking. */
int32_t bmx_batch_tune(bmx_batch *batch, int32_t sample_turns, int32_t *chunk_turns,
                       int32_t *grid, int32_t *waves, float *best_ms);

/* ---- closed orbit and one-turn matrix (bmad/code/closed_orbit_calc.f90:192) ----------------- *
 * For each of n_lat lattices: the fixed point of the one-turn map, found by Newton steps whose
 * Jacobian is the one-turn matrix from five-point central differences of tracked probes (25
 * particles per lattice per iteration, all lattices in one batch). n_dim = 4 closes x, px, y, py at the pz of
 * the guess with the rf cavities off (as closed_orbit_calc's i_dim = 4); n_dim = 6 closes all
 * six coordinates with the rf on. orbit[6 i .. 6 i + 5] holds lattice i's guess on entry and its
 * closed orbit on return; mat (optional, 36 per lattice, row-major) receives the one-turn matrix
 * d out_r / d in_c at the last iterate (finite differences, about 1e-9 relative). Converged when
 * every Newton step is below |orbit| rel_tol + abs_tol (closed_orbit_calc.f90:534). status[i]:
 * 0 converged, 1 not within max_iter, 2 a probe particle was lost, 3 singular (M - 1). */
int32_t bmx_closed_orbit(bmx_ctx *ctx, bmx_lattice *const *lats, int32_t n_lat, int32_t n_dim,
                         double *orbit, double *mat, int32_t max_iter, double rel_tol,
                         double abs_tol, int32_t *status);

/* Normal-mode tunes of a one-turn matrix, as make_N computes them (bmad/modules/mode3_mod.f90:
 * eigen_decomp_6mat and order_evecs_by_plane_dominance): for each conjugate pair of eigenvalues
 * the member whose eigenvector v has Im(v^T S conj(v)) >= 0 is conjugated to lambda, and the
 * tune is the angle of conj(lambda) in [0, 2 pi); the modes are ordered by plane dominance (the
 * largest |Re x| of the eigenvectors, normalised to unit length with their largest component
 * real, first; then the largest |Re y|). mat: 6x6 row-major. n_dim = 6: the three modes of mat;
 * n_dim = 4: the two transverse modes of its x-y block (the one-turn matrix with the rf off).
 * tunes: n_dim / 2 values in radians. Host only (no device). BMX_ERR_UNSTABLE when an eigenvalue
 * is real or off the unit circle by more than 1e-6. */
int32_t bmx_normal_mode_tunes(const double *mat, int32_t n_dim, double *tunes);

/* Chromaticities of n_lat lattices at once, as chrom_calc does for a closed ring
 * (bmad/code/chrom_calc.f90:60-190): closed orbits with the rf off (bmx_closed_orbit, n_dim = 4)
 * at pz - dE_low and pz + delta_e, the normal-mode tunes of their one-turn matrices
 * (bmx_normal_mode_tunes, n_dim = 4), and chrom = (tune_high - tune_low) / (delta_e + dE_low) per
 * mode, tunes in units of 2 pi; dE_low = delta_e unless pz - delta_e < -0.99 (then pz + 0.99).
 * delta_e <= 0 means 1e-4. chrom: 2 per lattice (a, b modes); tune (optional, 2 per lattice):
 * the fractional tunes at pz itself, in [0, 1). status[i]: 0 ok, 1..3 the failing closed
 * orbit's bmx_closed_orbit status, 4 no tune (unstable). All 3 n_lat closed orbits are found
 * in one batch. */
int32_t bmx_chromaticity(bmx_ctx *ctx, bmx_lattice *const *lats, int32_t n_lat, double delta_e,
                         double pz, double *chrom, double *tune, int32_t *status);

/* ---- server: continuous batching of tracking requests ---------------------------------------- *
 * A server owns a worker thread, a HIP stream and a device pool of hbm_budget bytes for
 * particle state. Submitted requests (a lattice, n particles in caller-owned host arrays, a
 * turn count) wait in FIFO order until their particles fit in the pool; admitted requests stay
 * resident and advance together in scheduling chunks of chunk_turns turns (one tile run over
 * all active requests per chunk, lost particles compacted every compact_turns turns); requests
 * arriving while a chunk runs join at the next chunk boundary. A finished request's particles
 * are copied to its host_out arrays and its pool space is reused. Each particle's result is
 * the same as tracking its request alone with bmx_track_turns. */
typedef struct bmx_server_config {
    int64_t hbm_budget;     /* bytes of device memory for particle state (188 B per particle) */
    int32_t chunk_turns;    /* turns per scheduling chunk (0 = 100) */
    int32_t compact_turns;  /* turns between alive-list compactions inside a chunk (0 = 5) */
} bmx_server_config;
typedef struct bmx_request_stats {
    int64_t submit_ns, admit_ns, done_ns;  /* steady clock, ns since the server was created */
    int32_t status;                        /* BMX_OK, or the error that ended the request */
    int32_t chunks;                        /* scheduling chunks the request took part in */
} bmx_request_stats;
int32_t bmx_server_create(bmx_ctx *ctx, const bmx_server_config *cfg, bmx_server **out);
/* Waits for every submitted request, then stops the worker and frees the pool. */
int32_t bmx_server_destroy(bmx_server *server);
/* Queue a request. host_in / host_out (may be the same arrays) must stay valid until the
 * request is waited for; BMX_ERR_BUDGET if n particles can never fit in the pool. */
int32_t bmx_server_submit(bmx_server *server, const bmx_lattice *lat, const bmx_beam_view *host_in,
                          const bmx_beam_view *host_out, int64_t n, int32_t n_turns, int32_t turn0,
                          int64_t *request_id);
/* done = 1 once the request's results are in host_out. */
int32_t bmx_server_poll(bmx_server *server, int64_t request_id, int32_t *done);
/* Block until the request is done, return its timing and forget it. */
int32_t bmx_server_wait(bmx_server *server, int64_t request_id, bmx_request_stats *stats);
/* Totals since creation: particle-element passes tracked, scheduling chunks run, and the
 * time the worker spent with at least one request active. */
int32_t bmx_server_totals(bmx_server *server, int64_t *passes, int64_t *chunks, int64_t *busy_ns);

/* Elements ix_start..ix_end (1-based, inclusive) once for every live particle, with the state
 * at every element exit written to orbit (caller-owned host memory of
 * n * (ix_end - ix_start + 1) * 8 doubles; orbit[(i * n_range + j) * 8 + k] = x, px, y, py, z,
 * pz, beta, state of particle i after element ix_start + j; zero after the element a
 * particle is lost in). track_all / track_many with their orbit(0:n) arrays
 * (track_all.f90:25-145, track_many.f90:50-157); loss_turn of a particle lost here = turn.
 * Blocking. */
int32_t bmx_track_orbit(bmx_ctx *ctx, const bmx_lattice *lat, bmx_beam *beam, int32_t ix_start,
                        int32_t ix_end, int32_t turn, double *orbit);

#ifdef __cplusplus
}
#endif

#endif /* BMX_H */
