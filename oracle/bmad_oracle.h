/*
 * bmad_oracle.h -- CPU restatement of Bmad's bmad_standard charged-particle tracking path.
 *
 * TEST INFRASTRUCTURE. This is the checker that the GPU library is compared against. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product
 * library never links or calls it.
 *
 * Element input is the public raw-attribute record bmx_ele (include/bmx.h); all bookkeeping
 * from those attributes (multipole normalisation, step counts, edge constants) is redone here
 * independently of the library, following the reference files cited in bmad_oracle.c.
 */
#ifndef BMAD_ORACLE_H
#define BMAD_ORACLE_H

#include <stdint.h>
#include "../include/bmx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Special functions (sim_utils/special_functions). */
double bmo_sqrt_one(double x);
double bmo_sinc(double x);
double bmo_cosc(double x);
double bmo_sincc(double x);
double bmo_modulo2(double x, double amp);
double bmo_c_multi(int32_t n, int32_t m);
/* multipole_kt_to_ab (bmad/code/multipole_kt_to_ab.f90) over orders 0 .. 21. */
void bmo_multipole_kt_to_ab(const double *knl, const double *knsl, const double *tn, double *an,
                            double *bn);

/* One particle through one element: track1 (bmad/code/track1.f90:44-343) with
 * tracking_method = bmad_standard. vec[6], beta and state are updated in place.
 * Returns 0, or BMX_ERR_UNSUPPORTED for elements/attributes the restatement does not cover. */
int32_t bmo_track1(const bmx_ele *ele, double mc2, int32_t charge, double p0c, double *vec,
                   double *beta, int32_t *state);

/* Low-level pieces exported for known-answer tests. */
void bmo_hwang_bend_edge_kick(const bmx_ele *ele, int32_t charge, int32_t first_edge,
                              double *vec);
void bmo_check_aperture_limit(const bmx_ele *ele, int32_t second_edge, double *vec,
                              int32_t *state, double *unstable_factor);

/* Bunch tracking over n_turns full turns of n_ele elements, SoA in place.
 * Mirrors ltt_run_beam_mode -> track_bunch -> track1_bunch_hom (beam_utils.f90:81-86): per
 * element an OpenMP parallel loop over particles, skipping dead ones. Loss bookkeeping:
 * loss_ele = 1-based element index (track1.f90:88-94), loss_turn = turn0-based index of the turn
 * (lt_tracking_mod.f90:1395). n_threads <= 0 uses the OpenMP default.
 * Returns 0 or an error code. *n_tracked receives the number of particle-element passes
 * tracked (alive at the element entrance). */
int32_t bmo_track_bunch(const bmx_ele *eles, int32_t n_ele, double mc2, int32_t charge,
                        double p0c, int64_t n, double *x, double *px, double *y, double *py,
                        double *z, double *pz, double *beta, int32_t *state, int32_t *loss_ele,
                        int32_t *loss_turn, int32_t n_turns, int32_t turn0, int32_t n_threads,
                        int64_t *n_tracked);

/* The same with radiation kicks (track1_radiation, radiation_mod.f90:61-147) at both ends of
 * the elements that radiate, with maps[2 i], maps[2 i + 1] = rm0, rm1 of element i and the
 * settings cfg; the fluctuation stream of particle ip is keyed by (cfg->seed, ip, turn,
 * element, edge). maps = NULL is bmo_track_bunch. */
int32_t bmo_track_bunch_rad(const bmx_ele *eles, int32_t n_ele, double mc2, int32_t charge,
                            double p0c, int64_t n, double *x, double *px, double *y, double *py,
                            double *z, double *pz, double *beta, int32_t *state, int32_t *loss_ele,
                            int32_t *loss_turn, int32_t n_turns, int32_t turn0, int32_t n_threads,
                            int64_t *n_tracked, const bmx_rad_map *maps, const bmx_rad_config *cfg);
/* Philox4x32-10 of one counter block under a key, and the six deviates of one kick. */
void bmo_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out);
void bmo_rad_gauss6(uint64_t seed, uint32_t id, int32_t turn, int32_t ix1, int32_t edge, double *g);

/* Single-particle trace: coordinates after each element for n_turns (out has
 * 6 * n_ele * n_turns doubles; entries after loss are the frozen lost coordinates). */
int32_t bmo_track_trace(const bmx_ele *eles, int32_t n_ele, double mc2, int32_t charge,
                        double p0c, double *vec, double *beta, int32_t *state, int32_t n_turns,
                        double *out);

#ifdef __cplusplus
}
#endif

#endif
