/*
 * bmad_oracle.c -- CPU restatement of Bmad's bmad_standard charged-particle tracking path,
 * plain C99 + GNU __float128, FP64, scalar, compiled with -ffp-contract=off.
 *
 * TEST INFRASTRUCTURE: this is the checker the MI355X library is compared against. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. It follows the
 * reference's routines statement by statement (same branches, same small-argument thresholds,
 * same left-to-right evaluation order of each Fortran expression); each function cites the
 * reference file:line it restates (paths relative to the bmad-ecosystem tree).
 *
 * Coverage: drift, quadrupole, sextupole / octupole / thick_multipole / kickers / collimators /
 * instrument / monitor / pipe (thick multipole kick-drift-kick), sbend without body multipoles
 * (Hwang edge kicks, exact and semi-linear body), rfcavity (standing wave, no coupler kick),
 * marker; offsets, pitches and tilt of non-bend elements (offset_particle non-bend branch);
 * elliptical and rectangular apertures incl. ecollimator/rcollimator momentum windows.
 * Electric multipoles, bend misalignments, bend body multipoles and k1 in bends return
 * BMX_ERR_UNSUPPORTED. Time (coord_struct%t, qp) is not carried: no result on this path
 * depends on it under relative time tracking (bmad_struct.f90:2350).
 */
#include <math.h>
#include <quadmath.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "bmad_oracle.h"

/* sim_utils/interfaces/physical_constants.f90:11-33 */
static const double PI = 3.141592653589793238462643383279;
#define TWOPI (2 * PI)
static const double C_LIGHT = 2.99792458e8;
/* bmad_struct.f90:2314 bmad_com%max_aperture_limit */
static const double MAX_APERTURE_LIMIT = 1e3;

/* A validity test that repeats one made just before on the same coordinates (the entrance and
 * exit aperture checks call orbit_too_large again, check_aperture_limit.f90:61; a drift or a
 * marker changes none of px, py, pz). The FLOP-counting build (count_flops.cpp) defines this to
 * leave the expression's operations uncounted, as the GPU kernels do not repeat these tests;
 * here it is the expression itself. */
#ifndef BMO_REPEATED_TEST
#define BMO_REPEATED_TEST(expr) (expr)
#endif

#define FIRST_TRACK_EDGE 11  /* bmad_struct.f90:379 */
#define SECOND_TRACK_EDGE 12

typedef struct {
    double vec[6];
    double beta, p0c, mc2;
    int charge, state;
} coord_t;

static double sq(double x) { return x * x; }

/* ------------------------------------------------------------------------------------------
 * special functions
 * ------------------------------------------------------------------------------------------ */

/* sim_utils/special_functions/sqrt_one.f90:18-41 (nd = 0): sqrt(1+x) - 1 without cancellation. */
double bmo_sqrt_one(double x)
{
    double s = sqrt(1 + x);
    double rad = s + 1;
    return x / rad;
}

/* sim_utils/special_functions/sinc.f90:15-40 (nd = 0). */
double bmo_sinc(double x)
{
    if (fabs(x) < 1e-8) return 1.0;
    return sin(x) / x;
}

/* sim_utils/special_functions/cosc.f90:15-40 (nd = 0): (1 - cos x) / x^2. */
double bmo_cosc(double x)
{
    if (fabs(x) < 1e-8) return 0.5;
    double s = bmo_sinc(x / 2);
    return 0.5 * (s * s);
}

/* sim_utils/special_functions/sincc.f90:15-45 (nd = 0): (x - sin x) / x^3. */
double bmo_sincc(double x)
{
    const double c0 = 1.0 / 6.0, c1 = -1.0 / 120.0, c2 = 1.0 / 5040.0, c3 = -1.0 / 362880.0;
    double x2 = x * x;
    if (fabs(x) < 0.1) return c0 + x2 * (c1 + x2 * (c2 + x2 * c3));
    return (x - sin(x)) / (x * x2);
}

/* Fortran MODULO for reals: fmod, then shifted to the sign of p. */
static double f_modulo(double a, double p)
{
    double r = fmod(a, p);
    if (r != 0 && ((r < 0) != (p < 0))) r += p;
    return r;
}

static __float128 f_moduloq(__float128 a, __float128 p)
{
    __float128 r = fmodq(a, p);
    if (r != 0 && ((r < 0) != (p < 0))) r += p;
    return r;
}

/* sim_utils/special_functions/modulo2_mod.f90:85-99 (modulo2_dp). */
double bmo_modulo2(double x, double amp)
{
    double m = f_modulo(x, 2 * amp);
    if (m >= amp) m = m - 2 * amp;
    return m;
}

/* modulo2_mod.f90:121-135 (modulo2_qp). */
static __float128 modulo2_q(__float128 x, __float128 amp)
{
    __float128 m = f_moduloq(x, 2 * amp);
    if (m >= amp) m = m - 2 * amp;
    return m;
}

/* bmad/low_level/c_multi.f90:19-63 with no_n_fact = .true.: the same table build (binomials,
 * division by n!, sign pattern on m mod 4) and the same re-multiplication by n!.
 * The reference builds the table lazily behind a SAVEd flag on first call, which races when
 * the first call happens inside an OpenMP region; here it is built once when the library is
 * loaded. */
static double c_multi_tab[BMX_N_POLE][BMX_N_POLE];
static double c_multi_nfact[BMX_N_POLE];

__attribute__((constructor)) static void c_multi_init(void)
{
    double (*c)[BMX_N_POLE] = c_multi_tab;
    memset(c_multi_tab, 0, sizeof c_multi_tab);
    c[0][0] = 1;
    for (int in = 1; in < BMX_N_POLE; in++) {
        c[in][0] = 1;
        c[in][in] = 1;
        for (int im = 1; im < in; im++) c[in][im] = c[in - 1][im - 1] + c[in - 1][im];
    }
    c_multi_nfact[0] = 1;
    for (int in = 0; in < BMX_N_POLE; in++) {
        if (in > 0) c_multi_nfact[in] = in * c_multi_nfact[in - 1];
        for (int im = 0; im <= in; im++) {
            c[in][im] = c[in][im] / c_multi_nfact[in];
            if (im % 4 == 0) c[in][im] = -c[in][im];
            if (im % 4 == 3) c[in][im] = -c[in][im];
        }
    }
}

double bmo_c_multi(int32_t n, int32_t m) { return c_multi_tab[n][m] * c_multi_nfact[n]; }

/* Fortran x**m for an integer variable m (libgcc __powidf2 repeated squaring). */
static double powi(double x, int m)
{
    unsigned n = m < 0 ? -(unsigned)m : (unsigned)m;
    double y = (n % 2) ? x : 1;
    while (n >>= 1) {
        x = x * x;
        if (n % 2) y = y * x;
    }
    return m < 0 ? 1 / y : y;
}

/* bmad/low_level/mexp.f90:14: x^m with 0^0 = 1. */
static double mexp(double x, int m)
{
    if (m < 0) return 0;
    if (m == 0) return 1;
    return powi(x, m);
}

/* Fortran NINT: round half away from zero. */
static long f_nint(double x) { return lround(x); }

/* ------------------------------------------------------------------------------------------
 * element attribute helpers
 * ------------------------------------------------------------------------------------------ */

static int is_thick_multipole_key(int key)
{
    /* track1_bmad.f90:87-88, 172-173, 215-216 */
    switch (key) {
    case BMX_RCOLLIMATOR: case BMX_ECOLLIMATOR: case BMX_MONITOR: case BMX_INSTRUMENT:
    case BMX_PIPE: case BMX_KICKER: case BMX_HKICKER: case BMX_VKICKER: case BMX_OCTUPOLE:
    case BMX_THICK_MULTIPOLE: case BMX_SEXTUPOLE:
        return 1;
    }
    return 0;
}

static int has_orientation_attributes(int key) { return key != BMX_MARKER; }

static int has_hkick_attributes(int key)
{
    switch (key) {
    case BMX_DRIFT: case BMX_MARKER: case BMX_KICKER: case BMX_HKICKER: case BMX_VKICKER:
        return 0;
    }
    return 1;
}

static int is_misaligned(const bmx_ele *e)
{
    return e->x_offset != 0 || e->y_offset != 0 || e->z_offset != 0 || e->x_pitch != 0 ||
           e->y_pitch != 0;
}

static int max_nonzero(const double *a, const double *b, int ubound)
{
    for (int i = ubound; i >= 0; i--) {
        if (a[i] != 0) return i;
        if (b[i] != 0) return i;
    }
    return -1;
}

/* bmad/code/multipole_ele_to_ab.f90:472-600 convert_this_ab, magnetic, use_ele_tilt = F
 * (so the tilt difference is zero), field_master = F, scale_multipoles = T. */
static int convert_this_ab(const bmx_ele *e, double *a, double *b)
{
    for (int n = 0; n < BMX_N_POLE; n++) {
        a[n] = e->a_pole[n];
        b[n] = e->b_pole[n];
    }
    if (max_nonzero(a, b, BMX_N_POLE - 1) < 0) return 0;
    double cnst;
    int ref_exp;
    switch (e->key) {
    case BMX_SBEND:
        cnst = e->l * (e->g + e->dg); ref_exp = 0; break;
    case BMX_KICKER:
        if (e->hkick == 0) cnst = e->vkick;
        else if (e->vkick == 0) cnst = e->hkick;
        else cnst = sqrt(sq(e->hkick) + sq(e->vkick));
        ref_exp = 0;
        break;
    case BMX_HKICKER: case BMX_VKICKER:
        cnst = e->kick; ref_exp = 0; break;
    case BMX_QUADRUPOLE:
        cnst = e->k1 * e->l; ref_exp = 1; break;
    case BMX_SEXTUPOLE:
        cnst = e->k2 * e->l; ref_exp = 2; break;
    case BMX_OCTUPOLE:
        cnst = e->k3 * e->l; ref_exp = 3; break;
    case BMX_THICK_MULTIPOLE: case BMX_RFCAVITY:
        return 0; /* multipoles do not scale */
    default:
        return BMX_ERR_UNSUPPORTED;
    }
    double radius = e->r0_mag;
    if (radius == 0) {
        for (int n = 0; n < BMX_N_POLE; n++) {
            a[n] = cnst * a[n];
            b[n] = cnst * b[n];
        }
    } else {
        double factor = cnst * powi(radius, ref_exp + 1);
        for (int n = 0; n < BMX_N_POLE; n++) {
            factor = factor / radius;
            a[n] = factor * a[n];
            b[n] = factor * b[n];
        }
    }
    return 0;
}

/* multipole_ele_to_ab.f90:326-405 this_ele_non_multipoles, magnetic. */
static void this_ele_non_multipoles(const bmx_ele *e, double *a_kick, double *b_kick)
{
    for (int i = 0; i < 4; i++) a_kick[i] = b_kick[i] = 0;
    switch (e->key) {
    case BMX_HKICKER: b_kick[0] = -e->kick; break;
    case BMX_VKICKER: a_kick[0] = e->kick; break;
    case BMX_KICKER:
        b_kick[0] = -e->hkick;
        a_kick[0] = e->vkick;
        break;
    default: {
        double hk = e->hkick, vk = e->vkick;
        if (has_hkick_attributes(e->key) && (hk != 0 || vk != 0)) {
            double tilt = (e->key == BMX_SBEND) ? e->ref_tilt : e->tilt;
            if (tilt != 0) {
                double cos_t = cos(tilt), sin_t = sin(tilt);
                b_kick[0] = -hk * cos_t - vk * sin_t;
                a_kick[0] = -hk * sin_t + vk * cos_t;
            } else {
                b_kick[0] = -hk;
                a_kick[0] = vk;
            }
        }
    }
    }
    switch (e->key) {
    case BMX_QUADRUPOLE: b_kick[1] = e->k1 * e->l; break;
    case BMX_SBEND:
        b_kick[0] = b_kick[0] + e->dg * e->l;
        b_kick[1] = e->k1 * e->l;
        b_kick[2] = e->k2 * e->l / 2.0;
        break;
    case BMX_SEXTUPOLE: b_kick[2] = e->k2 * e->l / 2.0; break;
    case BMX_OCTUPOLE: b_kick[3] = e->k3 * e->l / 6.0; break;
    }
}

/* multipole_ele_to_ab.f90:31-236 (non-slave element, magnetic, use_ele_tilt = F), with the
 * optional b1 pull-out of pull_this_b1 (:664-676). Returns ix_pole_max; *err set on failure. */
static int multipole_ele_to_ab(const bmx_ele *e, int include_kicks, double *a, double *b,
                               double *b1, int *err)
{
    *err = 0;
    for (int n = 0; n < BMX_N_POLE; n++) a[n] = b[n] = 0;
    if (b1) *b1 = 0;
    if (!e->is_on || e->key == BMX_PIPE) return -1;
    if (e->n_multipole > 0) {
        *err = convert_this_ab(e, a, b);
        if (*err) return -1;
    }
    double ak[4], bk[4];
    this_ele_non_multipoles(e, ak, bk);
    if (include_kicks) {
        for (int i = 0; i < 4; i++) {
            a[i] = a[i] + ak[i];
            b[i] = b[i] + bk[i];
        }
    }
    int ix = max_nonzero(a, b, BMX_N_POLE - 1);
    if (b1) {
        *b1 = b[1];
        b[1] = 0;
        if (ix <= 1) ix = max_nonzero(a, b, 1);
    }
    return ix;
}

/* multipole_ele_to_ab.f90:31-236 for pole_type = electric$ (non-slave element, use_ele_tilt = F):
 * ele%a_pole_elec / %b_pole_elec scaled by convert_this_ab's electric branch (:519-530: with
 * r0_elec = r not 0 or 1, order n is multiplied by r^-n); this_ele_non_multipoles adds nothing
 * for the elements here (only an elseparator has electric kicks, :389-398). Returns
 * ix_pole_max. */
static int multipole_ele_to_ab_elec(const bmx_ele *e, double *a, double *b)
{
    for (int n = 0; n < BMX_N_POLE; n++) a[n] = b[n] = 0;
    if (!e->is_on || e->key == BMX_PIPE || e->n_multipole_elec <= 0) return -1;
    for (int n = 0; n < BMX_N_POLE; n++) {
        a[n] = e->a_pole_elec[n];
        b[n] = e->b_pole_elec[n];
    }
    if (max_nonzero(a, b, BMX_N_POLE - 1) < 0) return -1;
    double radius = e->r0_elec;
    if (radius != 0 && radius != 1) {
        double factor = radius;
        for (int n = 0; n < BMX_N_POLE; n++) {
            factor = factor / radius;
            a[n] = factor * a[n];
            b[n] = factor * b[n];
        }
    }
    return max_nonzero(a, b, BMX_N_POLE - 1);
}

/* ------------------------------------------------------------------------------------------
 * coordinate transforms
 * ------------------------------------------------------------------------------------------ */

/* bmad/code/tilt_coords.f90:18-52 */
static void tilt_coords(double tilt_val, double *v)
{
    if (tilt_val == 0) return;
    double s = sin(tilt_val), c = cos(tilt_val);
    double x = v[0], y = v[2], px = v[1], py = v[3];
    v[0] = c * x + s * y;
    v[2] = -s * x + c * y;
    v[1] = c * px + s * py;
    v[3] = -s * px + c * py;
}

/* bmad/geometry/floor_angles_to_w_mat.f90:17-57 with psi = 0 kept symbolic. */
static void floor_angles_to_w_mat(double theta, double phi, double psi, double w[3][3], int inv)
{
    double s_the = sin(theta), c_the = cos(theta);
    double s_phi = sin(phi), c_phi = cos(phi);
    double s_psi = sin(psi), c_psi = cos(psi);
    if (!inv) {
        w[0][0] = c_the * c_psi - s_the * s_phi * s_psi;
        w[0][1] = -c_the * s_psi - s_the * s_phi * c_psi;
        w[0][2] = s_the * c_phi;
        w[1][0] = c_phi * s_psi;
        w[1][1] = c_phi * c_psi;
        w[1][2] = s_phi;
        w[2][0] = -s_the * c_psi - c_the * s_phi * s_psi;
        w[2][1] = s_the * s_psi - c_the * s_phi * c_psi;
        w[2][2] = c_the * c_phi;
    } else {
        w[0][0] = c_the * c_psi - s_the * s_phi * s_psi;
        w[0][1] = c_phi * s_psi;
        w[0][2] = -s_the * c_psi - c_the * s_phi * s_psi;
        w[1][0] = -c_the * s_psi - s_the * s_phi * c_psi;
        w[1][1] = c_phi * c_psi;
        w[1][2] = s_the * s_psi - c_the * s_phi * c_psi;
        w[2][0] = s_the * c_phi;
        w[2][1] = s_phi;
        w[2][2] = c_the * c_phi;
    }
}

static void matvec3(double w[3][3], const double *r, double *out)
{
    double t[3];
    for (int i = 0; i < 3; i++) t[i] = w[i][0] * r[0] + w[i][1] * r[1] + w[i][2] * r[2];
    for (int i = 0; i < 3; i++) out[i] = t[i];
}

/* sim_utils/math/rotation_3d_mod.f90:737-835 rotate_vec and rotate_mat (left multiplication by
 * the rotation about axis 0 = x, 1 = y, 2 = z); nothing for angle 0. */
static void rotate_vec(double *v, int axis, double angle)
{
    if (angle == 0) return;
    double ca = cos(angle), sa = sin(angle), a0, a1;
    switch (axis) {
    case 0: a0 = ca * v[1] - sa * v[2]; a1 = sa * v[1] + ca * v[2]; v[1] = a0; v[2] = a1; break;
    case 1: a0 = sa * v[2] + ca * v[0]; a1 = ca * v[2] - sa * v[0]; v[0] = a0; v[2] = a1; break;
    default: a0 = ca * v[0] - sa * v[1]; a1 = sa * v[0] + ca * v[1]; v[0] = a0; v[1] = a1; break;
    }
}

static void rotate_mat(double m[3][3], int axis, double angle)
{
    if (angle == 0) return;
    for (int j = 0; j < 3; j++) {
        double col[3] = {m[0][j], m[1][j], m[2][j]};
        rotate_vec(col, axis, angle);
        m[0][j] = col[0];
        m[1][j] = col[1];
        m[2][j] = col[2];
    }
}

/* A point and frame (floor_position_struct's r and w). */
typedef struct {
    double r[3];
    double w[3][3];
} bmo_pos_t;

static void mat3_mul(double a[3][3], double b[3][3], double out[3][3])
{
    double t[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
    memcpy(out, t, sizeof t);
}

/* bmad/geometry/bend_shift.f90:19-61: the position and frame a distance delta_s further along a
 * bend of curvature g (with the bend's ref_tilt when has_tilt). */
static void bend_shift(bmo_pos_t *p, double g, double delta_s, int has_tilt, double ref_tilt)
{
    double angle = delta_s * g;
    if (angle == 0) {
        p->r[2] = p->r[2] - delta_s;
        return;
    }
    double tlt = has_tilt ? ref_tilt : 0.0;
    double S[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (tlt != 0) {
        rotate_mat(S, 2, -ref_tilt);
        rotate_mat(S, 1, angle);
        rotate_mat(S, 2, ref_tilt);
    } else {
        rotate_mat(S, 1, angle);
    }
    double cos_one = -2 * sq(sin(angle / 2));
    double L[3] = {cos_one / g, 0.0 / g, -sin(angle) / g};
    if (has_tilt) rotate_vec(L, 2, ref_tilt);
    double r[3];
    matvec3(S, p->r, r);
    for (int i = 0; i < 3; i++) p->r[i] = r[i] + L[i];
    mat3_mul(S, p->w, p->w);
}

/* bmad/geometry/w_mat_for_tilt.f90:14-38 (not inverse): rotation about z by tilt. */
static void w_mat_for_tilt(double tilt, double w[3][3])
{
    double c = cos(tilt), s = sin(tilt);
    double m[3][3] = {{c, -s, 0.0}, {s, c, 0.0}, {0.0, 0.0, 1.0}};
    memcpy(w, m, sizeof m);
}

static void pos_apply(bmo_pos_t *p, double w[3][3])
{
    matvec3(w, p->r, p->r);
    mat3_mul(w, p->w, p->w);
}

static void track_a_drift_ex(coord_t *o, double length, int include_ref_motion);

/* bmad/code/offset_particle.f90:66-492, non-bend branch (lines 183-335 / 339-492), for an
 * element with orientation = +1 and a particle with direction = time_dir = +1.
 * has_s_pos/s_pos: optional s_pos argument (then drift_to_edge defaults to no$). */
static void offset_particle(const bmx_ele *e, int set, coord_t *o, int set_tilt,
                            int set_hvkicks, int has_s_pos, double s_pos)
{
    const int sign_z_vel = 1;
    int set_hv = set_hvkicks && e->is_on && has_hkick_attributes(e->key);
    if (e->key == BMX_KICKER || e->key == BMX_HKICKER || e->key == BMX_VKICKER)
        set_hv = set_hvkicks && e->is_on;
    int set_t = set_tilt && has_orientation_attributes(e->key);
    int drift_to; /* 0 = no, 1 = upstream, 2 = downstream */
    if (!has_orientation_attributes(e->key)) drift_to = 0;
    else if (has_s_pos) drift_to = 0;
    else if (set) drift_to = 1;
    else drift_to = 2;

    double length = e->l;
    double L_half = 0.5 * length;
    double s_pos0;
    if (has_s_pos) s_pos0 = s_pos;
    else if (set) s_pos0 = 0;
    else s_pos0 = length;
    double ds_center = L_half - s_pos0;
    double s_target = (drift_to == 1) ? (1 - sign_z_vel) * L_half : (1 + 1) * L_half;
    double rel_p = 1 + o->vec[5];
    const double charge_dir = 1.0 * sign_z_vel;
    int set_hv1 = 0, set_hv2 = 0;
    if (set_hv) {
        if (e->key == BMX_KICKER || e->key == BMX_HKICKER || e->key == BMX_VKICKER) set_hv2 = 1;
        else set_hv1 = 1;
    }
    int misaligned = has_orientation_attributes(e->key) && is_misaligned(e);
    double w[3][3], r[3], ds;

    const int bend_frame = e->key == BMX_SBEND && (e->g != 0 || e->ref_tilt != 0);
    if (set) {
        double s_body;
        if (misaligned && bend_frame) {
            /* offset_particle.f90:206-232: along the bend to its centre, the misalignment
             * (ele_misalignment_L_S_calc.f90:16-50 with roll = 0: L = offsets, S =
             * floor_angles_to_w_mat(x_pitch, y_pitch, 0)), the ref_tilt rotation about the
             * bend's entrance, back to the entrance face */
            bmo_pos_t ps = {{o->vec[0], o->vec[2], 0.0}, {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
            const double g = e->g, rt = e->ref_tilt;
            bend_shift(&ps, g, 1 * ds_center, 1, rt);
            double Sm[3][3], St[3][3];
            floor_angles_to_w_mat(e->x_pitch, e->y_pitch, 0.0, Sm, 0);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) St[i][j] = Sm[j][i];
            double d[3] = {ps.r[0] - e->x_offset, ps.r[1] - e->y_offset, ps.r[2] - e->z_offset};
            matvec3(St, d, ps.r);
            mat3_mul(St, ps.w, ps.w);
            if (rt != 0) {
                bend_shift(&ps, g, -L_half, 1, rt);
                double wt[3][3];
                w_mat_for_tilt(-rt, wt);
                pos_apply(&ps, wt);
                bend_shift(&ps, g, L_half, 0, 0.0);
            }
            if (drift_to != 0) {
                bend_shift(&ps, g, -sign_z_vel * L_half, 0, 0.0);
                ps.r[2] = ps.r[2] + (1 - sign_z_vel) * L_half;
            }
            s_body = ps.r[2];
            double pz = sq(rel_p) - sq(o->vec[1]) - sq(o->vec[3]);
            if (pz <= 0) {
                o->state = BMX_STATE_LOST_PZ;
            } else {
                double p0[3] = {o->vec[1], o->vec[3], sign_z_vel * sqrt(pz)}, pv[3];
                matvec3(ps.w, p0, pv);
                o->vec[1] = pv[0];
                o->vec[3] = pv[1];
                o->vec[0] = ps.r[0];
                o->vec[2] = ps.r[1];
            }
        } else if (misaligned) {
            r[0] = o->vec[0] - e->x_offset;
            r[1] = o->vec[2] - e->y_offset;
            r[2] = 0.0 - (e->z_offset + 1 * ds_center);
            floor_angles_to_w_mat(e->x_pitch, e->y_pitch, 0.0, w, 1);
            matvec3(w, r, r);
            r[2] = r[2] + L_half;
            s_body = r[2];
            double pz = sq(rel_p) - sq(o->vec[1]) - sq(o->vec[3]);
            if (pz <= 0) {
                o->state = BMX_STATE_LOST_PZ;
            } else {
                double p0[3] = {o->vec[1], o->vec[3], sign_z_vel * sqrt(pz)}, p[3];
                matvec3(w, p0, p);
                o->vec[1] = p[0];
                o->vec[3] = p[1];
                o->vec[0] = r[0];
                o->vec[2] = r[1];
            }
        } else {
            if (e->key == BMX_SBEND && e->ref_tilt != 0) tilt_coords(e->ref_tilt, o->vec);
            s_body = s_pos0;
        }
        ds = s_target - s_body;
        if (fabs(ds) > 1e-14 && drift_to != 0 && o->state == BMX_STATE_ALIVE)
            track_a_drift_ex(o, ds, 0);
        if (set_hv1) {
            o->vec[1] = o->vec[1] + 0.5 * charge_dir * e->hkick;
            o->vec[3] = o->vec[3] + 0.5 * charge_dir * e->vkick;
        }
        if (set_t && e->key != BMX_SBEND && e->tilt != 0) tilt_coords(e->tilt, o->vec);
        if (set_hv2) {
            if (e->key == BMX_HKICKER) o->vec[1] = o->vec[1] + 0.5 * charge_dir * e->kick;
            else if (e->key == BMX_VKICKER) o->vec[3] = o->vec[3] + 0.5 * charge_dir * e->kick;
            else {
                o->vec[1] = o->vec[1] + 0.5 * charge_dir * e->hkick;
                o->vec[3] = o->vec[3] + 0.5 * charge_dir * e->vkick;
            }
        }
    } else {
        if (set_hv2) {
            if (e->key == BMX_HKICKER) o->vec[1] = o->vec[1] + 0.5 * charge_dir * e->kick;
            else if (e->key == BMX_VKICKER) o->vec[3] = o->vec[3] + 0.5 * charge_dir * e->kick;
            else {
                o->vec[1] = o->vec[1] + 0.5 * charge_dir * e->hkick;
                o->vec[3] = o->vec[3] + 0.5 * charge_dir * e->vkick;
            }
        }
        if (set_t && e->key != BMX_SBEND) tilt_coords(-e->tilt, o->vec);
        if (set_hv1) {
            o->vec[1] = o->vec[1] + 0.5 * charge_dir * e->hkick;
            o->vec[3] = o->vec[3] + 0.5 * charge_dir * e->vkick;
        }
        double s_lab;
        if (misaligned && bend_frame) {
            /* offset_particle.f90:404-430, the reverse of the set branch */
            bmo_pos_t ps = {{o->vec[0], o->vec[2], 0.0}, {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
            const double g = e->g, rt = e->ref_tilt;
            bend_shift(&ps, g, ds_center, 0, 0.0);
            if (rt != 0) {
                bend_shift(&ps, g, -L_half, 0, 0.0);
                double wt[3][3];
                w_mat_for_tilt(rt, wt);
                pos_apply(&ps, wt);
                bend_shift(&ps, g, L_half, 1, rt);
            }
            double Sm[3][3];
            floor_angles_to_w_mat(e->x_pitch, e->y_pitch, 0.0, Sm, 0);
            double rr[3];
            matvec3(Sm, ps.r, rr);
            ps.r[0] = rr[0] + e->x_offset;
            ps.r[1] = rr[1] + e->y_offset;
            ps.r[2] = rr[2] + e->z_offset;
            mat3_mul(Sm, ps.w, ps.w);
            if (drift_to != 0) {
                bend_shift(&ps, g, sign_z_vel * L_half, 1, rt);
                ps.r[2] = ps.r[2] + (1 + sign_z_vel) * L_half;
            }
            s_lab = ps.r[2];
            double pz = sq(rel_p) - sq(o->vec[1]) - sq(o->vec[3]);
            if (pz <= 0) {
                o->state = BMX_STATE_LOST_PZ;
            } else {
                double p0[3] = {o->vec[1], o->vec[3], sign_z_vel * sqrt(pz)}, pv[3];
                matvec3(ps.w, p0, pv);
                o->vec[1] = pv[0];
                o->vec[3] = pv[1];
                o->vec[0] = ps.r[0];
                o->vec[2] = ps.r[1];
            }
        } else if (misaligned) {
            r[0] = o->vec[0];
            r[1] = o->vec[2];
            r[2] = 0.0 - ds_center;
            floor_angles_to_w_mat(e->x_pitch, e->y_pitch, 0.0, w, 0);
            matvec3(w, r, r);
            r[0] = r[0] + e->x_offset;
            r[1] = r[1] + e->y_offset;
            r[2] = r[2] + (e->z_offset + L_half);
            s_lab = r[2];
            double pz = sq(rel_p) - sq(o->vec[1]) - sq(o->vec[3]);
            if (pz <= 0) {
                o->state = BMX_STATE_LOST_PZ;
            } else {
                double p0[3] = {o->vec[1], o->vec[3], sign_z_vel * sqrt(pz)}, p[3];
                matvec3(w, p0, p);
                o->vec[1] = p[0];
                o->vec[3] = p[1];
                o->vec[0] = r[0];
                o->vec[2] = r[1];
            }
        } else {
            if (e->key == BMX_SBEND && e->ref_tilt != 0) tilt_coords(-e->ref_tilt, o->vec);
            s_lab = s_pos0;
        }
        ds = s_target - s_lab;
        if (fabs(ds) > 1e-14 && drift_to != 0 && o->state == BMX_STATE_ALIVE)
            track_a_drift_ex(o, ds, 0);
    }
}

/* ------------------------------------------------------------------------------------------
 * drift
 * ------------------------------------------------------------------------------------------ */

/* bmad/low_level/track_a_drift.f90:25-116 for orb%direction = ele_orientation = +1, beta > 0. */
static void track_a_drift_ex(coord_t *o, double length, int include_ref_motion)
{
    if (length == 0) return;
    double delta = o->vec[5];
    double rel_pc = 1 + delta;
    double px_rel = 1 * o->vec[1] / rel_pc;
    double py_rel = 1 * o->vec[3] / rel_pc;
    double pxy2 = sq(px_rel) + sq(py_rel);
    if (pxy2 >= 1) {
        o->state = BMX_STATE_LOST_PZ;
        return;
    }
    double ps_rel = sqrt(1 - pxy2);
    o->vec[0] = o->vec[0] + length * px_rel / ps_rel;
    o->vec[2] = o->vec[2] + length * py_rel / ps_rel;
    double dz;
    if (include_ref_motion) {
        double mc2 = o->mc2;
        dz = length * (bmo_sqrt_one((sq(mc2) * (2 * delta + sq(delta))) /
                                    (sq(o->p0c * rel_pc) + sq(mc2))) +
                       bmo_sqrt_one(-pxy2) / ps_rel);
    } else {
        dz = -length / ps_rel;
    }
    o->vec[4] = o->vec[4] + 1 * dz;
}

static void track_a_drift(coord_t *o, double length) { track_a_drift_ex(o, length, 1); }

/* ------------------------------------------------------------------------------------------
 * multipole kicks
 * ------------------------------------------------------------------------------------------ */

/* bmad/modules/multipole_mod.f90:314-401 ab_multipole_kick: magnetic (ele_orientation = 1,
 * same species as the reference, f = 1) or electric (f = charge / (beta p0c), a2 = a f,
 * b2 = -b f, :343-346). */
static void ab_multipole_kick_t(double a, double b, int n, const coord_t *o, double scale,
                                int electric, double *kx, double *ky)
{
    *kx = 0;
    *ky = 0;
    if (a == 0 && b == 0) return;
    double a2, b2;
    if (electric) {
        double f = o->charge / (o->beta * o->p0c);
        a2 = a * f;
        b2 = -b * f;
    } else {
        a2 = a * 1.0;
        b2 = b * 1.0;
    }
    a2 = scale * a2;
    b2 = scale * b2;
    double x = o->vec[0], y = o->vec[2];
    for (int m = 0; m <= n; m += 2) {
        double f = bmo_c_multi(n, m) * mexp(x, n - m) * mexp(y, m);
        *kx = *kx + b2 * f;
        *ky = *ky - a2 * f;
    }
    for (int m = 1; m <= n; m += 2) {
        double f = bmo_c_multi(n, m) * mexp(x, n - m) * mexp(y, m);
        *kx = *kx + a2 * f;
        *ky = *ky + b2 * f;
    }
}

static void ab_multipole_kick(double a, double b, int n, const coord_t *o, double scale,
                              double *kx, double *ky)
{
    ab_multipole_kick_t(a, b, n, o, scale, 0, kx, ky);
}

/* multipole_mod.f90:82-174 ab_multipole_kicks, magnetic branch. */
static void ab_multipole_kicks(const double *an, const double *bn, int ix_pole_max,
                               const bmx_ele *e, coord_t *o, double scale)
{
    double kx_tot = 0, ky_tot = 0, kx, ky;
    double r = e->p0c / o->p0c;
    for (int n = 0; n <= ix_pole_max; n++) {
        ab_multipole_kick(r * an[n], r * bn[n], n, o, scale, &kx, &ky);
        kx_tot = kx_tot + kx;
        ky_tot = ky_tot + ky;
    }
    o->vec[1] = o->vec[1] + kx_tot;
    o->vec[3] = o->vec[3] + ky_tot;
}

/* multipole_mod.f90:82-174 ab_multipole_kicks, electric branch (:128-143): the kicks change the
 * particle's energy, pz += (1+pz) sqrt_one(alpha) with alpha = (kx (2 px + kx) + ky (2 py + ky))
 * / (1+pz)^2 (lost_pz if alpha < -1), beta from the new momentum, z scaled by beta_new/beta_old.
 * scale is the longitudinal length of the field region. */
static void ab_multipole_kicks_elec(const double *an, const double *bn, int ix_pole_max,
                                    const bmx_ele *e, coord_t *o, double scale)
{
    coord_t o0 = *o;
    double kx_tot = 0, ky_tot = 0, kx, ky;
    double r = e->p0c / o->p0c;
    for (int n = 0; n <= ix_pole_max; n++) {
        ab_multipole_kick_t(r * an[n], r * bn[n], n, o, scale, 1, &kx, &ky);
        kx_tot = kx_tot + kx;
        ky_tot = ky_tot + ky;
    }
    o->vec[1] = o->vec[1] + kx_tot;
    o->vec[3] = o->vec[3] + ky_tot;
    double alpha = (kx_tot * (2 * o0.vec[1] + kx_tot) + ky_tot * (2 * o0.vec[3] + ky_tot)) /
                   sq(1 + o0.vec[5]);
    if (alpha < -1) {
        o->state = BMX_STATE_LOST_PZ;
        return;
    }
    double dpz = (1 + o0.vec[5]) * bmo_sqrt_one(alpha);
    o->vec[5] = o0.vec[5] + dpz;
    o->beta = (1 + o->vec[5]) / sqrt(sq(1 + o->vec[5]) + sq(o->mc2 / o->p0c));
    o->vec[4] = o0.vec[4] * o->beta / o0.beta;
}

static void apply_energy_kick(double dE, coord_t *o);

/* apply_element_edge_kick.f90:283-326 electric_longitudinal_fringe, bmad_standard field, not a
 * bend with exact multipoles: the particle crosses the hard edge of the electric multipole
 * field and its energy changes by the potential there, phi = -sum_n Re((b_n - i a_n)
 * (x + i y)^(n+1)) / (n+1), applied as apply_energy_kick(-ff phi) with ff = +charge at the
 * entrance (at_sign = +1) and -charge at the exit. Complex products written out as Fortran
 * evaluates them ((a+ib)(c+id) = (ac - bd) + i(ad + bc)). */
static void electric_longitudinal_fringe(coord_t *o, const bmx_ele *e, const double *a_elec,
                                         const double *b_elec, int ix_elec_max, double at_sign)
{
    if (e->l == 0) return;
    if (ix_elec_max < 0) return;
    double ff = at_sign * o->charge;
    double xr = 1, xi = 0;                 /* xiy */
    double cr = o->vec[0], ci = o->vec[2]; /* c_vec */
    double phi = 0;
    for (int i = 0; i <= ix_elec_max; i++) {
        double nr = xr * cr - xi * ci, ni = xr * ci + xi * cr;
        xr = nr;
        xi = ni;
        if (a_elec[i] == 0 && b_elec[i] == 0) continue;
        double abr = b_elec[i], abi = -a_elec[i];
        phi = phi - (abr * xr - abi * xi) / (i + 1);
    }
    apply_energy_kick(-ff * phi, o);
}

/* ------------------------------------------------------------------------------------------
 * quadrupole
 * ------------------------------------------------------------------------------------------ */

/* bmad/low_level/quad_mat2_calc.f90:22-81 (mat2 and z_coef). */
static void quad_mat2_calc(double k1, double length, double rel_p, double mat2[2][2],
                           double zc[3])
{
    double sqrt_k = sqrt(fabs(k1));
    double sk_l = sqrt_k * length;
    double cx, sx;
    if (fabs(sk_l) < 1e-10) {
        double k_l2 = k1 * sq(length);
        cx = 1 + k_l2 / 2;
        sx = (1 + k_l2 / 6) * length;
    } else if (k1 < 0) {
        cx = cos(sk_l);
        sx = sin(sk_l) / sqrt_k;
    } else {
        cx = cosh(sk_l);
        sx = sinh(sk_l) / sqrt_k;
    }
    mat2[0][0] = cx;
    mat2[0][1] = sx / rel_p;
    mat2[1][0] = k1 * sx * rel_p;
    mat2[1][1] = cx;
    zc[0] = k1 * (-cx * sx + length) / 4;
    zc[1] = -k1 * sq(sx) / (2 * rel_p);
    zc[2] = -(cx * sx + length) / (4 * sq(rel_p));
}

/* bmad/low_level/low_energy_z_correction.f90:22-58 (no matrix). */
static double low_energy_z_correction(const coord_t *o, const bmx_ele *e, double ds)
{
    double mass = o->mc2;
    double p0c = o->p0c;
    double e_tot = e->e_tot;
    double beta0 = p0c / e_tot;
    double pz = o->vec[5];
    if (mass * sq(beta0 * pz) < 3e-7 * e_tot) {
        double f = sq(beta0) * (2 * sq(beta0) - sq(mass / e_tot) / 2);
        return ds * pz * (1 - 3 * (pz * sq(beta0)) / 2 + sq(pz) * f) * sq(mass / e_tot);
    }
    return ds * (o->beta - beta0) / beta0;
}

/* bmad/modules/fringe_mod.f90:692-865 hard_multipole_edge_kick for a quadrupole or a bend
 * (n_max = 1, ap = 0, bp = k1; orbit only, no derivatives of second order, which only the
 * matrix uses), fringe_type hard_edge_only (quadrupole) or full (:715-722), charge_dir = 1. Complex products are written
 * out in the order Fortran evaluates them, left to right. */
static void hard_multipole_edge_kick(coord_t *o, const bmx_ele *e, int particle_at)
{
    if (e->fringe_type != BMX_FRINGE_HARD_EDGE_ONLY && e->fringe_type != BMX_FRINGE_FULL) return;
    /* a bend's hard edge with k1 is in hwang_bend_edge_kick (:716-717) */
    if (e->key == BMX_SBEND && e->fringe_type == BMX_FRINGE_HARD_EDGE_ONLY) return;
    const double x = o->vec[0], y = o->vec[2];
    const double rel_p = 1 + o->vec[5];
    double fx = 0, dfx_dx = 0, dfx_dy = 0, fy = 0, dfy_dx = 0, dfy_dy = 0;
    /* poly_n1 = 1, poly = xy, then per n: poly_n1 = poly, poly = poly * xy */
    double pr = x, pi = y;           /* poly */
    double n1r = 1, n1i = 0;         /* poly_n1 */
    const int n_max = 1;
    for (int n = 1; n <= n_max; n++) {
        n1r = pr;
        n1i = pi;
        double t_r = pr * x - pi * y, t_i = pr * y + pi * x;
        pr = t_r;
        pi = t_i;
        const double ap = 0, bp = e->k1;
        /* dpoly_dx = (n+1) poly_n1, dpoly_dy = i dpoly_dx */
        const double dxr = (n + 1) * n1r, dxi = (n + 1) * n1i;
        const double dyr = -dxi, dyi = dxr;
        /* cab = charge_dir * cmplx(bp, ap) / (4 (n+2) rel_p), negated at the first edge */
        const double den = 4 * (n + 2) * rel_p;
        double cr = 1.0 * bp / den, ci = 1.0 * ap / den;
        if (particle_at == FIRST_TRACK_EDGE) {
            cr = -cr;
            ci = -ci;
        }
        const double cn = (double)(n + 3) / (n + 1);
        /* x part: xny = cmplx(x, -cn y), dxny_dy = cmplx(0, -cn) */
        double xr = x, xi = -cn * y;
        double ur, ui, vr, vi;
        /* fx += Re(cab * poly * xny) */
        ur = cr * pr - ci * pi;
        ui = cr * pi + ci * pr;
        fx = fx + (ur * xr - ui * xi);
        /* dfx_dx += Re(cab * (dpoly_dx * xny + poly)) */
        vr = dxr * xr - dxi * xi + pr;
        vi = dxr * xi + dxi * xr + pi;
        dfx_dx = dfx_dx + (cr * vr - ci * vi);
        /* dfx_dy += Re(cab * (dpoly_dy * xny + poly * dxny_dy)) */
        vr = (dyr * xr - dyi * xi) + (pr * 0.0 - pi * (-cn));
        vi = (dyr * xi + dyi * xr) + (pr * (-cn) + pi * 0.0);
        dfx_dy = dfx_dy + (cr * vr - ci * vi);
        /* y part: xny = cmplx(y, cn x), dxny_dx = cmplx(0, cn) */
        xr = y;
        xi = cn * x;
        fy = fy + (ur * xr - ui * xi);
        /* dfy_dx += Re(cab * (dpoly_dx * xny + poly * dxny_dx)) */
        vr = (dxr * xr - dxi * xi) + (pr * 0.0 - pi * cn);
        vi = (dxr * xi + dxi * xr) + (pr * cn + pi * 0.0);
        dfy_dx = dfy_dx + (cr * vr - ci * vi);
        /* dfy_dy += Re(cab * (dpoly_dy * xny + poly)) */
        vr = dyr * xr - dyi * xi + pr;
        vi = dyr * xi + dyi * xr + pi;
        dfy_dy = dfy_dy + (cr * vr - ci * vi);
    }
    const double px = o->vec[1], py = o->vec[3];
    const double denom = (1 - dfx_dx) * (1 - dfy_dy) - dfx_dy * dfy_dx;
    o->vec[0] = o->vec[0] - fx;
    o->vec[1] = px + ((1.0 - dfy_dy - denom) * px + dfy_dx * py) / denom;
    o->vec[2] = o->vec[2] - fy;
    o->vec[3] = py + (dfx_dy * px + (1.0 - dfx_dx - denom) * py) / denom;
    o->vec[4] = o->vec[4] + (o->vec[1] * fx + o->vec[3] * fy) / rel_p;
}

/* bmad/modules/fringe_mod.f90:556-661 soft_quadrupole_edge_kick for a quadrupole, fringe_type
 * soft_edge_only or full, time_dir = +1, charge_dir = 1 (orbit only). */
static void soft_quadrupole_edge_kick(coord_t *o, const bmx_ele *e, int particle_at)
{
    if (e->fringe_type != BMX_FRINGE_SOFT_EDGE_ONLY && e->fringe_type != BMX_FRINGE_FULL) return;
    const double rel_p = 1 + o->vec[5];
    const double k1_rel = 1.0 * e->k1 / rel_p;
    double f1 = k1_rel * e->fq1;
    const double f2 = k1_rel * e->fq2;
    if (f1 == 0 && f2 == 0) return;
    if (particle_at == SECOND_TRACK_EDGE) f1 = -f1;
    const double ef1 = exp(f1);
    const double v0 = o->vec[0], v1 = o->vec[1], v2 = o->vec[2], v3 = o->vec[3];
    const double vx = v1 / rel_p, vy = v3 / rel_p;
    o->vec[4] = o->vec[4] - (f1 * v0 + f2 * (1 + f1 / 2) * vx / ef1) * vx +
                (f1 * v2 + f2 * (1 - f1 / 2) * vy * ef1) * vy;
    o->vec[0] = v0 * ef1 + vx * f2;
    o->vec[1] = v1 / ef1;
    o->vec[2] = v2 / ef1 - vy * f2;
    o->vec[3] = v3 * ef1;
}

/* apply_element_edge_kick.f90:64-178,270-271 for a quadrupole with a bmad_standard field: at an
 * end where fringe_at puts a fringe (at_this_ele_end), the electric longitudinal fringe after
 * the magnetic edges at the entrance (hard multipole edge, then the soft quadrupole edge) and
 * before them at the exit (soft, then hard). */
static void quadrupole_edge(coord_t *o, const bmx_ele *e, int particle_at, const double *an_elec,
                            const double *bn_elec, int ix_elec_max)
{
    const int physical_end = particle_at == FIRST_TRACK_EDGE ? BMX_ENTRANCE_END : BMX_EXIT_END;
    if (!(e->fringe_at == BMX_BOTH_ENDS || e->fringe_at == physical_end)) return;
    if (particle_at == FIRST_TRACK_EDGE) {
        hard_multipole_edge_kick(o, e, particle_at);
        soft_quadrupole_edge_kick(o, e, particle_at);
        electric_longitudinal_fringe(o, e, an_elec, bn_elec, ix_elec_max, 1.0);
    } else {
        electric_longitudinal_fringe(o, e, an_elec, bn_elec, ix_elec_max, -1.0);
        soft_quadrupole_edge_kick(o, e, particle_at);
        hard_multipole_edge_kick(o, e, particle_at);
    }
}

/* bmad/low_level/track_a_quadrupole.f90:17-148 (magnetic and electric multipoles, hard and
 * soft edges). init_fringe_info.f90:44-54: the edges are tracked when fringe_type is not none
 * and fringe_at is not no_end, or when the element has electric multipoles (whatever its
 * fringe_type); the track returns at a lost particle after either edge (:64-66,136-137),
 * before offset_particle's exit transform. */
static int track_a_quadrupole(coord_t *o, const bmx_ele *e)
{
    const double charge_dir = 1.0;
    double length = 1 * e->l;
    double an[BMX_N_POLE], bn[BMX_N_POLE], b1;
    double an_elec[BMX_N_POLE], bn_elec[BMX_N_POLE];
    int err;
    int ix_mag_max = multipole_ele_to_ab(e, 1, an, bn, &b1, &err);
    if (err) return err;
    int ix_elec_max = multipole_ele_to_ab_elec(e, an_elec, bn_elec);
    const int ft = e->fringe_type == 0 ? BMX_FRINGE_NONE : e->fringe_type;
    if (ft != BMX_FRINGE_NONE && ft != BMX_FRINGE_SOFT_EDGE_ONLY && ft != BMX_FRINGE_HARD_EDGE_ONLY &&
        ft != BMX_FRINGE_FULL)
        return BMX_ERR_UNSUPPORTED;
    const int has_fringe = ((ft != BMX_FRINGE_NONE && e->fringe_at != BMX_NO_END) || ix_elec_max > -1);
    int n_step = 1;
    if (ix_mag_max > -1 || ix_elec_max > -1) {
        long ns = f_nint(fabs(length) / e->ds_step);
        n_step = ns > 1 ? (int)ns : 1;
    }
    double r_step = 1.0 / n_step;
    double step_len = e->l * r_step;
    if (length == 0) n_step = 0;

    if (is_misaligned(e) || e->tilt != 0) offset_particle(e, 1, o, 1, 0, 0, 0);
    if (has_fringe) {
        quadrupole_edge(o, e, FIRST_TRACK_EDGE, an_elec, bn_elec, ix_elec_max);
        if (o->state != BMX_STATE_ALIVE) return 0;
    }
    if (ix_mag_max > -1) ab_multipole_kicks(an, bn, ix_mag_max, e, o, r_step / 2);
    if (ix_elec_max > -1) ab_multipole_kicks_elec(an_elec, bn_elec, ix_elec_max, e, o, step_len / 2);

    for (int i = 1; i <= n_step; i++) {
        double rel_p = 1 + o->vec[5];
        double k1 = charge_dir * b1 / (length * rel_p);
        double mx[2][2], my[2][2], dzx[3], dzy[3];
        quad_mat2_calc(-k1, step_len, rel_p, mx, dzx);
        quad_mat2_calc(k1, step_len, rel_p, my, dzy);
        double *v = o->vec;
        v[4] = v[4] + (dzx[0] * sq(v[0]) + dzx[1] * v[0] * v[1] + dzx[2] * sq(v[1]) +
                       dzy[0] * sq(v[2]) + dzy[1] * v[2] * v[3] + dzy[2] * sq(v[3]));
        double x = v[0], px = v[1], y = v[2], py = v[3];
        v[0] = mx[0][0] * x + mx[0][1] * px;
        v[1] = mx[1][0] * x + mx[1][1] * px;
        v[2] = my[0][0] * y + my[0][1] * py;
        v[3] = my[1][0] * y + my[1][1] * py;
        v[4] = v[4] + low_energy_z_correction(o, e, step_len);
        if (ix_mag_max > -1)
            ab_multipole_kicks(an, bn, ix_mag_max, e, o, i == n_step ? r_step / 2 : r_step);
        if (ix_elec_max > -1)
            ab_multipole_kicks_elec(an_elec, bn_elec, ix_elec_max, e, o,
                                    i == n_step ? step_len / 2 : step_len);
    }
    if (has_fringe) {
        quadrupole_edge(o, e, SECOND_TRACK_EDGE, an_elec, bn_elec, ix_elec_max);
        if (o->state != BMX_STATE_ALIVE) return 0;
    }
    if (is_misaligned(e) || e->tilt != 0) offset_particle(e, 0, o, 1, 0, 0, 0);
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * thick multipole (sextupole, octupole, kickers, collimators, instrument, monitor, pipe)
 * ------------------------------------------------------------------------------------------ */

/* bmad/low_level/track_a_thick_multipole.f90:18-133 (no elseparator, no electric
 * multipoles, ac_kicker amplitude = 1). */
static int track_a_thick_multipole(coord_t *o, const bmx_ele *e)
{
    double length = e->l * 1;
    double an[BMX_N_POLE], bn[BMX_N_POLE];
    int err;
    int ix_mag_max = multipole_ele_to_ab(e, 1, an, bn, NULL, &err);
    if (err) return err;
    int n_step;
    if (ix_mag_max == -1) {
        n_step = 1;
    } else {
        long ns = f_nint(fabs(length / e->ds_step));
        n_step = ns > 1 ? (int)ns : 1;
    }
    double r_step = 1.0 / n_step;
    double step_len = length / n_step;

    if (is_misaligned(e) || e->tilt != 0) offset_particle(e, 1, o, 1, 0, 0, 0);
    /* init_fringe_info: only sextupole/quad hard-edge fringes exist here (fringe_type none by
     * default, set_ele_defaults.f90:37); anything else is outside the restatement. */
    if (e->fringe_type != BMX_FRINGE_NONE && e->fringe_type != 0 &&
        e->fringe_at != BMX_NO_END)
        return BMX_ERR_UNSUPPORTED;

    if (ix_mag_max > -1) ab_multipole_kicks(an, bn, ix_mag_max, e, o, 1.0 * r_step / 2);
    for (int i = 1; i <= n_step; i++) {
        track_a_drift(o, step_len);
        if (ix_mag_max > -1)
            ab_multipole_kicks(an, bn, ix_mag_max, e, o,
                               i == n_step ? 1.0 * r_step / 2 : 1.0 * r_step);
    }
    if (is_misaligned(e) || e->tilt != 0) offset_particle(e, 0, o, 1, 0, 0, 0);
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * sbend
 * ------------------------------------------------------------------------------------------ */

/* bmad/modules/fringe_mod.f90:241-419 hwang_bend_edge_kick, time_dir = +1 (orbit only). */
void bmo_hwang_bend_edge_kick(const bmx_ele *e, int32_t charge, int32_t first_edge, double *vec)
{
    (void)charge;
    const double c_dir = 1.0; /* rel_tracking_charge * orientation * direction */
    double e_factor = 1 / (1 + vec[5]);
    double g_tot, k1;
    if (e->is_on) {
        g_tot = (e->g + e->dg) * c_dir;
        k1 = e->k1;
    } else {
        g_tot = 0;
        k1 = 0;
    }
    double ang, fint_gap;
    if (first_edge) {
        ang = e->e1;
        fint_gap = e->fint * e->hgap;
    } else {
        ang = e->e2;
        fint_gap = e->fintx * e->hgapx;
    }
    if (e->fringe_type == BMX_FRINGE_HARD_EDGE_ONLY) fint_gap = 0;
    double cos_e = cos(ang), sin_e = sin(ang);
    double tan_e = sin_e / cos_e, sec_e = 1 / cos_e;
    double gt = g_tot * tan_e;
    double gt2 = g_tot * sq(tan_e);
    double gs2 = g_tot * sq(sec_e);
    double k1_tane = k1 * tan_e * c_dir;
    double fg_factor = 2 * fint_gap * gs2 * g_tot * sec_e * (1 + sq(sin_e));
    double v[6];
    for (int i = 0; i < 6; i++) v[i] = vec[i];
    double dx, dpx, dy, dpy, dz;
    if (first_edge) {
        dx = (-gt2 * sq(v[0]) + gs2 * sq(v[2])) * e_factor / 2;
        dpx = (gt * g_tot * (1 + 2 * sq(tan_e)) * sq(v[2]) / 2 +
               gt2 * (v[0] * v[1] - v[2] * v[3]) + k1_tane * (sq(v[0]) - sq(v[2]))) *
              e_factor;
        dy = gt2 * v[0] * v[2] * e_factor;
        dpy = (fg_factor * v[2] - gt2 * v[0] * v[3] - (g_tot + gt2) * v[1] * v[2] -
               2 * k1_tane * v[0] * v[2]) *
              e_factor;
        dz = sq(e_factor) * 0.5 *
             (sq(v[2]) * fg_factor + v[0] * v[0] * v[0] * (4.0 * k1_tane - gt * gt2) / 6.0 +
              0.5 * v[0] * sq(v[2]) * (-4.0 * k1_tane + gt * gs2) +
              (sq(v[0]) * v[1] - 2.0 * v[0] * v[2] * v[3]) * gt2 - v[1] * sq(v[2]) * gs2);
    } else {
        dx = (gt2 * sq(v[0]) - gs2 * sq(v[2])) * e_factor / 2;
        dpx = (gt2 * (v[2] * v[3] - v[0] * v[1]) + k1_tane * (sq(v[0]) - sq(v[2])) -
               gt * gt2 * (sq(v[0]) + sq(v[2])) / 2) *
              e_factor;
        dy = -gt2 * v[0] * v[2] * e_factor;
        dpy = (fg_factor * v[2] + gt2 * v[0] * v[3] + (g_tot + gt2) * v[1] * v[2] +
               (gt * gs2 - 2 * k1_tane) * v[0] * v[2]) *
              e_factor;
        dz = sq(e_factor) * 0.5 *
             (sq(v[2]) * fg_factor + v[0] * v[0] * v[0] * (4.0 * k1_tane - gt * gt2) / 6.0 +
              0.5 * v[0] * sq(v[2]) * (-4.0 * k1_tane + gt * gs2) -
              (sq(v[0]) * v[1] - 2.0 * v[0] * v[2] * v[3]) * gt2 + v[1] * sq(v[2]) * gs2);
    }
    vec[0] = v[0] + 1 * dx;
    vec[1] = v[1] + 1 * (dpx + gt * v[0]);
    vec[2] = v[2] + 1 * dy;
    vec[3] = v[3] + 1 * (dpy - gt * v[2]);
    vec[4] = v[4] + 1 * dz;
}

/* bmad/modules/fringe_mod.f90:1713-1771 bmad_rot_xz (orbit only): the frame rotated by a
 * about the y axis, no field. Returns nonzero on err_flag. */
static int bmad_rot_xz(double a, double *vec)
{
    double x = vec[0], px = vec[1], py = vec[3], delta = vec[5];
    double pp = 1 + delta;
    double ps = sq(pp) - sq(px) - sq(py);
    if (ps <= 0) return 1;
    ps = sqrt(ps);
    double ca = cos(a), sa = sin(a);
    double den = ps * ca - px * sa;
    double L = x * sa / den;
    vec[0] = x * ps / den;
    vec[1] = px * ca + ps * sa;
    vec[2] = vec[2] + py * L;
    vec[4] = vec[4] - pp * L;
    return 0;
}

/* bmad/modules/fringe_mod.f90:1588-1690 bmad_wedger (orbit only): a wedge of angle a of the
 * dipole field g_tot (bmad_rot_xz when g_tot = 0). Returns nonzero on err_flag. */
static int bmad_wedger(double a, double g_tot, double *vec)
{
    double b1 = g_tot;
    if (b1 == 0) return bmad_rot_xz(a, vec);
    double x = vec[0], px = vec[1], py = vec[3], delta = vec[5];
    double pp = 1 + delta;
    double ps = sq(pp) - sq(px) - sq(py);
    if (ps <= 0) return 1;
    ps = sqrt(ps);
    double pt = sq(pp) - sq(py);
    if (pt < 1e-10) return 1;
    pt = sqrt(pt);
    double ca = cos(a), sa = sin(a), s2a = sin(2.0 * a), s2 = sq(sa);
    double pxn = px * ca + (ps - b1 * x) * sa;
    double psn = sq(pt) - sq(pxn);
    if (psn < 1e-10) return 1;
    psn = sqrt(psn);
    double den = psn + ps * ca - px * sa;
    double num = x * px * s2a + s2 * (2 * x * ps - b1 * sq(x));
    double L = (a + asin(px / pt) - asin(pxn / pt)) / b1;
    vec[0] = x * ca + num / den;
    vec[1] = pxn;
    vec[2] = vec[2] + py * L;
    vec[4] = vec[4] - pp * L;
    return 0;
}

/* bmad/modules/fringe_mod.f90:1797-1933 bmad_fringe_dipoler (orbit only): the hard-edge dipole
 * fringe with fint, hgap in the frame of the edge. Returns nonzero on err_flag. */
static int bmad_fringe_dipoler(double *vec, double g_tot, double beta0, double fint, double hgap,
                               int particle_at)
{
    double b = particle_at == SECOND_TRACK_EDGE ? -g_tot : g_tot;
    double x = vec[0], px = vec[1], y = vec[2], py = vec[3], z = vec[4], delta = vec[5];
    double pp = 1 + delta;
    double ps = sq(pp) - sq(px) - sq(py);
    if (ps <= 0) return 1;
    ps = sqrt(ps);
    double ee = sqrt(sq(pp) + 1 / sq(beta0) - 1);
    double beta = pp / ee;
    double xp = px / ps, yp = py / ps;
    double kap = b * fint * hgap;
    double fac2 = 1 + sq(yp);
    double d[3][3];
    d[0][0] = (1 + sq(xp)) / ps;  d[1][0] = xp * yp / ps;  d[2][0] = -xp;
    d[0][1] = xp * yp / ps;       d[1][1] = fac2 / ps;     d[2][1] = -yp;
    d[0][2] = -ee * xp / sq(ps);  d[1][2] = -ee * yp / sq(ps); d[2][2] = ee / ps;
    double fi0 = atan(xp / fac2) - 2 * kap * (1 + sq(xp) * (2 + sq(yp))) * ps;
    double co2 = b / sq(cos(fi0));
    double co1 = co2 / (1 + sq(xp / fac2));
    double fi[3];
    fi[0] = co1 / fac2 - 4 * kap * co2 * xp * (2 + sq(yp)) * ps;
    fi[1] = -2 * co1 * xp * yp / sq(fac2) - 4 * kap * co2 * sq(xp) * yp * ps;
    fi[2] = -2 * kap * co2 * (1 + sq(xp) * (2 + sq(yp)));
    double fi0t = b * tan(fi0);
    double bb2 = fi[0] * d[0][1] + fi[1] * d[1][1] + fi[2] * d[2][1];
    double sqv = 1 - 2 * bb2 * y;
    if (sqv < 0) return 1;
    sqv = sqrt(sqv);
    double ynew = 2 * y / (1 + sqv);
    double bb1 = fi[0] * d[0][0] + fi[1] * d[1][0] + fi[2] * d[2][0];
    double bb3 = fi[0] * d[0][2] + fi[1] * d[1][2] + fi[2] * d[2][2];
    vec[0] = x + 0.5 * bb1 * sq(ynew);
    vec[2] = ynew;
    vec[3] = py - fi0t * ynew;
    vec[4] = z + 0.5 * beta * bb3 * sq(ynew);
    return 0;
}

/* bmad/modules/fringe_mod.f90:1474-1562 exact_bend_edge_kick (orbit only, time_dir = +1):
 * at the entrance the wedge of the edge angle without field, the fringe kick, then the wedge
 * back through the field; at the exit the reverse. g_tot here is g + the element's dg
 * attribute (0 when switched off). lost_pz, coordinates unchanged, on any err_flag. */
static void exact_bend_edge_kick(coord_t *o, const bmx_ele *e, int particle_at)
{
    double beta0 = e->p0c / e->e_tot;
    double g_tot = e->is_on ? e->g + e->dg : 0;
    double edge_angle, fint, hgap;
    if (particle_at == FIRST_TRACK_EDGE) {
        edge_angle = 1 * e->e1;
        fint = e->fint;
        hgap = e->hgap;
    } else {
        edge_angle = 1 * e->e2;
        fint = e->fintx;
        hgap = e->hgapx;
    }
    double vec[6];
    for (int i = 0; i < 6; i++) vec[i] = o->vec[i];
    int err;
    if (particle_at == FIRST_TRACK_EDGE) {
        err = bmad_wedger(edge_angle, 0.0, vec);
        if (!err) err = bmad_fringe_dipoler(vec, g_tot, beta0, 1 * fint, hgap, particle_at);
        if (!err) err = bmad_wedger(-edge_angle, g_tot, vec);
    } else {
        err = bmad_wedger(-edge_angle, g_tot, vec);
        if (!err) err = bmad_fringe_dipoler(vec, g_tot, beta0, 1 * fint, hgap, particle_at);
        if (!err) err = bmad_wedger(edge_angle, 0.0, vec);
    }
    if (err) {
        o->state = BMX_STATE_LOST_PZ;
        return;
    }
    for (int i = 0; i < 6; i++) o->vec[i] = vec[i];
}

/* fringe_here (bmad/code/fringe_here.f90:15-35) and bend_edge_kick (fringe_mod.f90:40-128): the
 * hard multipole edge before the bend's edge kick at the entrance and after it at the exit,
 * the edge kick by fringe_type (full: exact_bend_edge_kick; basic_bend, hard_edge_only:
 * hwang_bend_edge_kick; none: nothing). Other fringe types are outside this restatement. */
static int bend_edge_kick(coord_t *o, const bmx_ele *e, int particle_at)
{
    int fringe_at = e->fringe_at;
    int physical_end = particle_at == FIRST_TRACK_EDGE ? BMX_ENTRANCE_END : BMX_EXIT_END;
    int here = (fringe_at == BMX_BOTH_ENDS) || (fringe_at == physical_end);
    if (!here || e->fringe_type == BMX_FRINGE_NONE) return 0;
    if (particle_at == FIRST_TRACK_EDGE) hard_multipole_edge_kick(o, e, particle_at);
    switch (e->fringe_type) {
    case BMX_FRINGE_FULL:
        exact_bend_edge_kick(o, e, particle_at);
        break;
    case BMX_FRINGE_BASIC_BEND: case BMX_FRINGE_HARD_EDGE_ONLY:
        bmo_hwang_bend_edge_kick(e, o->charge, particle_at == FIRST_TRACK_EDGE, o->vec);
        break;
    default: return BMX_ERR_UNSUPPORTED;
    }
    if (particle_at == SECOND_TRACK_EDGE) hard_multipole_edge_kick(o, e, particle_at);
    return 0;
}

/* apply_element_edge_kick.f90:64-189,270-271 for a bend with a bmad_standard field and
 * exact_multipoles off or g = 0: at an end fringe_at names, the electric longitudinal fringe
 * before the magnetic edge at the exit and after it at the entrance. */
static int bend_element_edge(coord_t *o, const bmx_ele *e, int particle_at, const double *an_elec,
                             const double *bn_elec, int ix_elec_max)
{
    const int physical_end = particle_at == FIRST_TRACK_EDGE ? BMX_ENTRANCE_END : BMX_EXIT_END;
    if (!(e->fringe_at == BMX_BOTH_ENDS || e->fringe_at == physical_end)) return 0;
    if (particle_at == SECOND_TRACK_EDGE)
        electric_longitudinal_fringe(o, e, an_elec, bn_elec, ix_elec_max, -1.0);
    int err = bend_edge_kick(o, e, particle_at);
    if (err) return err;
    if (particle_at == FIRST_TRACK_EDGE)
        electric_longitudinal_fringe(o, e, an_elec, bn_elec, ix_elec_max, 1.0);
    return 0;
}

static int orbit_too_large(coord_t *o, double *unstable);

/* bmad/code/elec_multipole_field.f90:1-84 (field only): Ex, Ey of the electric multipole of
 * order n with strengths a, b at the particle. */
static void elec_multipole_field(double a, double b, int n, const coord_t *o, double *Ex,
                                 double *Ey)
{
    *Ex = 0;
    *Ey = 0;
    if (a == 0 && b == 0) return;
    double x = o->vec[0], y = o->vec[2];
    for (int m = 0; m <= n; m += 2) {
        double f = bmo_c_multi(n, m) * mexp(x, n - m) * mexp(y, m);
        *Ex = *Ex - b * f;
        *Ey = *Ey - a * f;
    }
    for (int m = 1; m <= n; m += 2) {
        double f = bmo_c_multi(n, m) * mexp(x, n - m) * mexp(y, m);
        *Ex = *Ex + a * f;
        *Ey = *Ey - b * f;
    }
}

/* track_a_bend.f90:336-467 apply_multipole_kicks, not exact (exact_multipoles off or g = 0):
 * the magnetic multipoles as a field B (from ab_multipole_kick with ele_orientation = 0, scaled
 * by r_step p0c / (c charge_ref)) and the kick coef c_dir (1 + g x) c / p0c x B; the electric
 * multipoles as a field E and the kick step_len coef (1 + g x) charge / (ps beta p0c) x E with
 * the energy change of a transverse kick. */
static void bend_multipole_kicks(coord_t *o, const bmx_ele *e, double coef, double step_len,
                                 double r_step, int ix_mag_max, const double *an,
                                 const double *bn, int ix_elec_max, const double *an_elec,
                                 const double *bn_elec)
{
    const double charge = o->charge;
    const double c_dir = 1.0 * charge;
    double B1 = 0, B2 = 0, E1 = 0, E2 = 0;
    double f_p0c = r_step * e->p0c / (C_LIGHT * o->charge);
    for (int i = 0; i <= ix_mag_max; i++) {
        if (an[i] == 0 && bn[i] == 0) continue;
        double kx, ky;
        ab_multipole_kick(an[i], bn[i], i, o, 1.0, &kx, &ky);
        B1 = B1 + f_p0c * ky;
        B2 = B2 - f_p0c * kx;
    }
    for (int i = 0; i <= ix_elec_max; i++) {
        if (an_elec[i] == 0 && bn_elec[i] == 0) continue;
        double Ex, Ey;
        elec_multipole_field(an_elec[i], bn_elec[i], i, o, &Ex, &Ey);
        E1 = E1 + Ex;
        E2 = E2 + Ey;
    }
    if (ix_mag_max > -1) {
        double f_coef = coef * c_dir * (1 + e->g * o->vec[0]) * C_LIGHT / o->p0c;
        o->vec[1] = o->vec[1] - f_coef * B2;
        o->vec[3] = o->vec[3] + f_coef * B1;
    }
    if (ix_elec_max > -1) {
        coord_t o0 = *o;
        double rel_p0 = 1 + o0.vec[5];
        double ps2 = sq(rel_p0) - sq(o0.vec[1]) - sq(o0.vec[3]);
        double ps = sqrt(ps2) / rel_p0;
        double f_coef = step_len * coef * (1 + e->g * o0.vec[0]) * charge / (ps * o0.beta * o0.p0c);
        double Kx = f_coef * E1, Ky = f_coef * E2;
        double alpha = (Kx * (2 * o0.vec[1] + Kx) + Ky * (2 * o0.vec[3] + Ky)) / sq(rel_p0);
        o->vec[1] = o0.vec[1] + Kx;
        o->vec[3] = o0.vec[3] + Ky;
        o->vec[5] = o0.vec[5] + rel_p0 * bmo_sqrt_one(alpha);
        o->beta = (1 + o->vec[5]) /
                  sqrt(ps2 + sq(o->vec[1]) + sq(o->vec[3]) + sq(o->mc2 / o->p0c));
        o->vec[4] = o0.vec[4] * o->beta / o0.beta;
    }
}

/* track_a_bend.f90:500-659 sbend_body_with_k1_map (orbit only, time_dir = direction =
 * orientation = +1): one step of a bend with a quadrupole gradient, b1 = k1 L. */
static void sbend_body_with_k1_map(coord_t *o, const bmx_ele *e, double dg, double b1, int n_step)
{
    const double rel_charge_dir = 1.0;
    double g = e->g;
    double length = 1 * e->l / n_step;
    double k1 = b1 / e->l;
    double g_tot = (g + dg) * rel_charge_dir;
    double rel_p = 1 + o->vec[5];
    double rel_p2 = sq(rel_p);
    double k_x = k1 + g * g_tot;
    double x_c = (g * rel_p - g_tot) / k_x;
    double om_x = sqrt(fabs(k_x) / rel_p);
    double om_y = sqrt(fabs(k1) / rel_p);
    double tau_x = -copysign(1.0, k_x);
    double tau_y = copysign(1.0, k1);
    double s_x, c_x, z2, s_y, c_y;
    double arg = om_x * length;
    if (arg < 1e-6) {
        s_x = (1 + tau_x * sq(arg) / 6) * length;
        c_x = 1 + tau_x * sq(arg) / 2;
        z2 = g * sq(length) / (2 * rel_p);
    } else if (k_x > 0) {
        s_x = sin(arg) / om_x;
        c_x = cos(arg);
        z2 = tau_x * g * (1 - c_x) / (rel_p * sq(om_x));
    } else {
        s_x = sinh(arg) / om_x;
        c_x = cosh(arg);
        z2 = tau_x * g * (1 - c_x) / (rel_p * sq(om_x));
    }
    arg = om_y * length;
    if (arg < 1e-6) {
        s_y = (1 + tau_y * sq(arg) / 6) * length;
        c_y = 1 + tau_y * sq(arg) / 2;
    } else if (k1 < 0) {
        s_y = sin(om_y * length) / om_y;
        c_y = cos(om_y * length);
    } else {
        s_y = sinh(om_y * length) / om_y;
        c_y = cosh(om_y * length);
    }
    double r[6];
    for (int i = 0; i < 6; i++) r[i] = o->vec[i];
    r[0] = r[0] - x_c;
    double z0 = -g * x_c * length;
    double z1 = -g * s_x;
    double z11 = tau_x * sq(om_x) * (length - c_x * s_x) / 4;
    double z12 = -tau_x * sq(om_x) * sq(s_x) / (2 * rel_p);
    double z22 = -(length + c_x * s_x) / (4 * rel_p2);
    double z33 = tau_y * sq(om_y) * (length - c_y * s_y) / 4;
    double z34 = -tau_y * sq(om_y) * sq(s_y) / (2 * rel_p);
    double z44 = -(length + c_y * s_y) / (4 * rel_p2);
    o->vec[0] = c_x * r[0] + s_x * r[1] / rel_p + x_c;
    o->vec[1] = tau_x * sq(om_x) * rel_p * s_x * r[0] + c_x * r[1];
    o->vec[2] = c_y * r[2] + s_y * r[3] / rel_p;
    o->vec[3] = tau_y * sq(om_y) * rel_p * s_y * r[2] + c_y * r[3];
    o->vec[4] = r[4] + 1 * z0 + z1 * r[0] + z2 * r[1] + z11 * sq(r[0]) + z12 * r[0] * r[1] +
                z22 * sq(r[1]) + z33 * sq(r[2]) + z34 * r[2] * r[3] + z44 * sq(r[3]);
}

/* One step of the bend body without a gradient (track_a_bend.f90:120-209): the semi-linear map
 * near the zero orbit, else the exact circle. Returns nonzero when the body loses the particle
 * (the reference returns from track_a_bend there). */
static int bend_body_step(coord_t *o, const bmx_ele *e, double g, double dg, double g_tot,
                          double angle, double step_len)
{
    double *v = o->vec;
    double x = v[0], px = v[1], y = v[2], py = v[3], z = v[4], pz = v[5];
    double rel_p = 1 + pz;
    double rel_p2 = sq(rel_p);
    if (dg == 0 && fabs(x * g) < 1e-9 && fabs(px) < 1e-9 && fabs(py) < 1e-9 && fabs(pz) < 1e-9) {
        /* semi-linear map near the zero orbit, track_a_bend.f90:136-159 */
        double ll = step_len;
        double cos_a = cos(angle), sin_a = sin(angle);
        double sinc_a = bmo_sinc(angle), cosc_a = bmo_cosc(angle);
        double gam2 = sq(o->mc2) / (sq(rel_p * o->p0c) + sq(o->mc2));
        double m56 = ll * (gam2 - sq(g * ll) * bmo_sincc(angle));
        v[0] = cos_a * x + ll * sinc_a * px + g * sq(ll) * cosc_a * pz;
        v[1] = -g * sin_a * x + cos_a * px + g * ll * sinc_a * pz;
        v[2] = y + ll * py;
        v[4] = -g * ll * sinc_a * x - g * sq(ll) * cosc_a * px + z + m56 * pz;
        return 0;
    }
    /* exact body, track_a_bend.f90:163-209 */
    double beta_ref = e->p0c / e->e_tot;
    double sinc_a = bmo_sinc(angle);
    double pt = sqrt(rel_p2 - sq(py));
    if (fabs(px) > pt) {
        o->state = BMX_STATE_LOST_PZ;
        return 1;
    }
    double g_p = g_tot / pt;
    double phi_1 = asin(px / pt);
    double cos_a = cos(angle), sin_a = sin(angle);
    double cosc_a = bmo_cosc(angle);
    double cos_plus = cos(angle + phi_1);
    double sin_plus = sin(angle + phi_1);
    double alpha = 2 * (1 + g * x) * sin_plus * step_len * sinc_a -
                   g_p * sq((1 + g * x) * step_len * sinc_a);
    double r = sq(cos_plus) + g_p * alpha;
    if (r < 0 || (fabs(g_p) < 1e-5 && fabs(cos_plus) < 1e-5)) {
        o->state = BMX_STATE_LOST;
        v[0] = 2 * MAX_APERTURE_LIMIT;
        return 1;
    }
    double rad = sqrt(r), denom, xi;
    if (cos_plus > 0) {
        denom = rad + cos_plus;
        xi = alpha / denom;
    } else {
        denom = rad - cos_plus;
        xi = denom / g_p;
    }
    v[0] = x * cos_a - sq(step_len) * g * cosc_a + xi;
    if (fabs(v[0]) > MAX_APERTURE_LIMIT) {
        o->state = BMX_STATE_LOST;
        return 1;
    }
    double L_u = xi;
    double L_v = -1 * (step_len * sinc_a + x * sin_a);
    double L_c = sqrt(sq(L_v) + sq(L_u));
    double angle_p = 2 * (angle + phi_1 - 1 * atan2(L_u, -L_v));
    double L_p = 1 * L_c / bmo_sinc(angle_p / 2);
    v[1] = pt * sin(phi_1 + angle - angle_p);
    v[2] = y + py * L_p / pt;
    v[4] = z + o->beta * step_len / beta_ref - rel_p * L_p / pt;
    return 0;
}

/* bmad/low_level/track_a_bend.f90:18-333 (time_dir = direction = orientation = +1): the
 * offset_particle frame change (bend branch), the entrance edge, the multipoles (magnetic, with
 * b1 pulled out for the k1 map unless exact_multipoles is set, and the b0 part as the field
 * error dg; electric) as kicks around n_step body steps, orbit_too_large, the exit edge.
 * exact_multipoles with g /= 0 (bend_exact_multipole_field) is outside this restatement. */
static int track_a_bend(coord_t *o, const bmx_ele *e)
{
    const int exact = e->exact_multipoles > 1;
    if (exact && e->g != 0) return BMX_ERR_UNSUPPORTED;
    if (e->fringe_type != BMX_FRINGE_NONE && e->fringe_type != BMX_FRINGE_BASIC_BEND &&
        e->fringe_type != BMX_FRINGE_HARD_EDGE_ONLY && e->fringe_type != BMX_FRINGE_FULL)
        return BMX_ERR_UNSUPPORTED;
    const int has_misalign = is_misaligned(e) || e->ref_tilt != 0;
    if (has_misalign) offset_particle(e, 1, o, 1, 0, 0, 0);
    const double rel_charge_dir = 1.0;

    double an_elec[BMX_N_POLE], bn_elec[BMX_N_POLE];
    int ix_elec_max = multipole_ele_to_ab_elec(e, an_elec, bn_elec);
    int err = bend_element_edge(o, e, FIRST_TRACK_EDGE, an_elec, bn_elec, ix_elec_max);
    if (err) return err;
    if (orbit_too_large(o, NULL)) return 0;

    double an[BMX_N_POLE], bn[BMX_N_POLE], b1;
    int ix_mag_max;
    if (exact) {
        ix_mag_max = multipole_ele_to_ab(e, 1, an, bn, NULL, &err);
        b1 = 0;
    } else {
        ix_mag_max = multipole_ele_to_ab(e, 1, an, bn, &b1, &err);
        b1 = b1 * rel_charge_dir;
        if (fabs(b1) < 1e-10) {
            bn[1] = b1;
            b1 = 0;
        }
    }
    if (err) return err;
    double length = 1 * e->l;
    double g = e->g, dg;
    if (length == 0) {
        dg = 0;
    } else {
        dg = bn[0] / e->l;
        bn[0] = 0;
    }
    double g_tot = (g + dg) * rel_charge_dir;
    if (!e->is_on) {
        dg = -g;
        g_tot = 0;
        b1 = 0;
    }
    int n_step = 1;
    if (ix_mag_max > -1 || ix_elec_max > -1) {
        long ns = f_nint(fabs(length) / e->ds_step);
        n_step = ns > 1 ? (int)ns : 1;
    }
    double r_step = 1.0 / n_step;
    double step_len = length / n_step;
    double angle = g * step_len;
    if (ix_mag_max > -1 || ix_elec_max > -1)
        bend_multipole_kicks(o, e, 0.5, step_len, r_step, ix_mag_max, an, bn, ix_elec_max,
                             an_elec, bn_elec);
    for (int n = 1; n <= n_step; n++) {
        if (b1 != 0) {
            sbend_body_with_k1_map(o, e, dg, b1, n_step);
            o->vec[4] = o->vec[4] + low_energy_z_correction(o, e, step_len);
        } else if ((g == 0 && dg == 0) || step_len == 0) {
            track_a_drift(o, step_len);
        } else {
            if (bend_body_step(o, e, g, dg, g_tot, angle, step_len)) return 0;
        }
        if (ix_mag_max > -1 || ix_elec_max > -1)
            bend_multipole_kicks(o, e, n == n_step ? 0.5 : 1.0, step_len, r_step, ix_mag_max, an,
                                 bn, ix_elec_max, an_elec, bn_elec);
    }
    if (orbit_too_large(o, NULL)) return 0;
    err = bend_element_edge(o, e, SECOND_TRACK_EDGE, an_elec, bn_elec, ix_elec_max);
    if (err) return err;
    if (has_misalign) offset_particle(e, 0, o, 1, 0, 0, 0);
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * rfcavity
 * ------------------------------------------------------------------------------------------ */

/* bmad/low_level/apply_energy_kick.f90:23-75 (ddE_dr = 0, no matrix). */
static void apply_energy_kick(double dE, coord_t *o)
{
    double mc2 = o->mc2;
    double p0c = o->p0c;
    double pc = (1 + o->vec[5]) * p0c;
    double beta_old = o->beta;
    double E_old = pc / o->beta;
    double E_new = E_old + dE;
    if (E_new < mc2) {
        o->vec[5] = -1;
        o->beta = 0;
        o->state = BMX_STATE_LOST_PZ;
        return;
    }
    o->vec[5] = o->vec[5] + (1 + o->vec[5]) * bmo_sqrt_one((2 * E_old * dE + sq(dE)) / sq(pc));
    double pc_new = p0c * (1 + o->vec[5]);
    double beta_new = pc_new / E_new;
    o->vec[4] = o->vec[4] * beta_new / beta_old;
    o->beta = beta_new;
}

/* bmad/low_level/track_a_rfcavity.f90:17-146 with relative time tracking
 * (bmad_struct.f90:2350), coupler_strength = 0 (rf_coupler_kick.f90:41), orientation = +1.
 * The RF phase follows particle_rf_time.f90:32-120 (time = -z/(beta c) wrapped by
 * modulo2 in quad precision) and rf_ref_time_offset.f90:19-45 (= 0 here); e_accel_field.f90:15-58
 * gives voltage with field_autoscale ignored. */
static int track_a_rfcavity(coord_t *o, const bmx_ele *e)
{
    if (e->rf_frequency == 0 && e->voltage != 0) {
        o->state = BMX_STATE_LOST;
        return 0;
    }
    double length = e->l * 1;
    double r_step = 1.0 / 1;
    double step_len = length / 1;
    (void)step_len;
    long ns = f_nint(fabs(length) / e->ds_step);
    int n_slice = ns > 1 ? (int)ns : 1;
    const double charge_dir = 1.0;
    double an[BMX_N_POLE], bn[BMX_N_POLE];
    int err;
    int ix_mag_max = multipole_ele_to_ab(e, 1, an, bn, NULL, &err);
    if (err) return err;

    offset_particle(e, 1, o, 1, 1, 0, 0);
    if (ix_mag_max > -1) ab_multipole_kicks(an, bn, ix_mag_max, e, o, 0.5 * r_step);

    double voltage = (e->is_on ? e->voltage : 0.0) * charge_dir;
    /* particle_rf_time: relative time, real(qp) result */
    __float128 time = (__float128)(-o->vec[4] / (o->beta * C_LIGHT));
    if (e->rf_frequency != 0) time = modulo2_q(time, (__float128)0.5 / (__float128)e->rf_frequency);
    __float128 ph = (__float128)TWOPI *
                    ((__float128)e->phi0 - (time - (__float128)0.0) * (__float128)e->rf_frequency);
    double phase0 = (double)ph;
    phase0 = phase0 + TWOPI * e->phi0_multipass;
    double phase = bmo_modulo2(phase0, PI);

    for (int i = 0; i <= n_slice; i++) {
        double dE = 1 * voltage * sin(phase) / n_slice;
        if (i == 0 || i == n_slice) dE = dE / 2;
        apply_energy_kick(dE, o);
        if (o->vec[5] == -1) {
            o->state = BMX_STATE_LOST_PZ;
            return 0;
        }
        if (i != n_slice) {
            double z = o->vec[4];
            track_a_drift(o, length / n_slice);
            phase = phase + TWOPI * e->rf_frequency * (o->vec[4] - z) / (C_LIGHT * o->beta);
        }
    }
    if (ix_mag_max > -1) ab_multipole_kicks(an, bn, ix_mag_max, e, o, 0.5 * r_step);
    offset_particle(e, 0, o, 1, 1, 0, 0);
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * apertures and validity
 * ------------------------------------------------------------------------------------------ */

/* bmad/low_level/orbit_too_large.f90:26-120 (charged particle, check_momentum = T). */
static int orbit_too_large(coord_t *o, double *unstable)
{
    double *v = o->vec;
    double u = 0;
    int too = 1;
    if (v[0] > MAX_APERTURE_LIMIT) {
        o->state = BMX_STATE_LOST_POS_X;
        u = (fabs(v[0]) - MAX_APERTURE_LIMIT) / MAX_APERTURE_LIMIT;
    } else if (-v[0] > MAX_APERTURE_LIMIT) {
        o->state = BMX_STATE_LOST_NEG_X;
        u = (fabs(v[0]) - MAX_APERTURE_LIMIT) / MAX_APERTURE_LIMIT;
    } else if (v[2] > MAX_APERTURE_LIMIT) {
        o->state = BMX_STATE_LOST_POS_Y;
        u = (fabs(v[2]) - MAX_APERTURE_LIMIT) / MAX_APERTURE_LIMIT;
    } else if (-v[2] > MAX_APERTURE_LIMIT) {
        o->state = BMX_STATE_LOST_NEG_Y;
        u = (fabs(v[2]) - MAX_APERTURE_LIMIT) / MAX_APERTURE_LIMIT;
    } else {
        double rel_p = 1 + v[5];
        if (rel_p < 0) {
            o->state = BMX_STATE_LOST_PZ;
            u = fabs(rel_p) + 1e-6;
        } else {
            double f_unstable = sq(v[1]) + sq(v[3]) - sq(rel_p);
            if (f_unstable > 0) {
                u = sqrt(f_unstable);
                if (fabs(v[1]) > fabs(v[3])) o->state = v[1] > 0 ? BMX_STATE_LOST_POS_X : BMX_STATE_LOST_NEG_X;
                else o->state = v[3] > 0 ? BMX_STATE_LOST_POS_Y : BMX_STATE_LOST_NEG_Y;
            } else {
                too = 0;
            }
        }
    }
    if (too && unstable) *unstable = u;
    return too;
}

/* check_aperture_limit.f90:343-353 check_this_limit */
static void check_this_limit(coord_t *o, double coord, double width2, double center, double *u)
{
    if (width2 == 0) return;
    if (coord > center - width2 && coord < center + width2) return;
    o->state = BMX_STATE_LOST;
    double a = (center - width2) - coord, b = coord - (center + width2);
    *u = a > b ? a : b;
}

/* check_aperture_limit.f90:358-385 elliptical_params_setup */
static int elliptical_params_setup(double lim1, double lim2, double *pos, double *width2)
{
    if ((lim1 == 0 && lim2 != 0) || (lim1 != 0 && lim2 == 0)) return 1;
    *width2 = (lim2 - lim1) / 2;
    double center = (lim2 + lim1) / 2;
    *pos = *pos - center;
    if (*width2 == 0) *width2 = MAX_APERTURE_LIMIT;
    if (*width2 < 0) *width2 = 0;
    return 0;
}

/* bmad/low_level/check_aperture_limit.f90:30-337, non-wall-transition apertures
 * (rectangular, elliptical), element without lords, aperture_limit_on = T. */
static void check_aperture_limit(coord_t *o, const bmx_ele *e, int particle_at, double *uf)
{
    double u = 0;
    *uf = 0;
    if (BMO_REPEATED_TEST(orbit_too_large(o, uf))) return;
    int physical_end = particle_at == FIRST_TRACK_EDGE ? BMX_ENTRANCE_END : BMX_EXIT_END;
    int at = e->aperture_at;
    int here;
    if (at == BMX_NO_END || at == 0) here = 0;
    else if (at == BMX_CONTINUOUS) here = 1;
    else here = (at == BMX_BOTH_ENDS) || (at == physical_end);
    if (!here) return;
    double *v = o->vec;
    if (e->key == BMX_ECOLLIMATOR) {
        double xw = e->px_aperture_width2, yw = e->py_aperture_width2;
        double x0 = e->px_aperture_center, y0 = e->py_aperture_center;
        if (xw == 0 || yw == 0) {
            check_this_limit(o, v[1], xw, x0, &u);
            check_this_limit(o, v[3], yw, y0, &u);
        } else {
            double r = sq((v[1] - x0) / xw) + sq((v[3] - y0) / yw);
            if (r > 1) {
                o->state = BMX_STATE_LOST;
                u = sqrt(r) - 1;
            }
        }
        xw = e->z_aperture_width2; yw = e->pz_aperture_width2;
        x0 = e->z_aperture_center; y0 = e->pz_aperture_center;
        if (xw == 0 || yw == 0) {
            check_this_limit(o, v[4], xw, x0, &u);
            check_this_limit(o, v[5], yw, y0, &u);
        } else {
            double r = sq((v[4] - x0) / xw) + sq((v[5] - y0) / yw);
            if (r > 1) {
                o->state = BMX_STATE_LOST;
                u = sqrt(r) - 1;
            }
        }
    } else if (e->key == BMX_RCOLLIMATOR) {
        check_this_limit(o, v[1], e->px_aperture_width2, e->px_aperture_center, &u);
        check_this_limit(o, v[3], e->py_aperture_width2, e->py_aperture_center, &u);
        check_this_limit(o, v[4], e->z_aperture_width2, e->z_aperture_center, &u);
        check_this_limit(o, v[5], e->pz_aperture_width2, e->pz_aperture_center, &u);
    }
    double xp = v[0], yp = v[2];
    if (e->offset_moves_aperture && (is_misaligned(e) || e->tilt != 0)) {
        coord_t o2 = *o;
        int do_tilt = (e->key == BMX_ECOLLIMATOR || e->key == BMX_RCOLLIMATOR);
        double s_here = particle_at == FIRST_TRACK_EDGE ? 0.0 : e->l;
        offset_particle(e, 1, &o2, do_tilt, 0, 1, s_here);
        xp = o2.vec[0];
        yp = o2.vec[2];
    }
    double x1 = -e->x1_limit, x2 = e->x2_limit, y1 = -e->y1_limit, y2 = e->y2_limit;
    int type = e->aperture_type;
    if (type == BMX_APERTURE_ELLIPTICAL) {
        double xw, yw;
        if (elliptical_params_setup(x1, x2, &xp, &xw)) return;
        if (elliptical_params_setup(y1, y2, &yp, &yw)) return;
        double r = sq(xp / xw) + sq(yp / yw);
        if (r < 1) {
            *uf = u;
            return;
        }
        if (fabs(xp / xw) > fabs(yp / yw)) o->state = xp > 0 ? BMX_STATE_LOST_POS_X : BMX_STATE_LOST_NEG_X;
        else o->state = yp > 0 ? BMX_STATE_LOST_POS_Y : BMX_STATE_LOST_NEG_Y;
        u = sqrt(r) - 1;
    } else {
        double f;
        if (x1 != 0 && xp < x1) {
            f = fabs((xp - x1) / (x2 - x1));
            if (f > u) { o->state = BMX_STATE_LOST_NEG_X; u = f; }
        }
        if (x2 != 0 && xp > x2) {
            f = fabs((xp - x2) / (x2 - x1));
            if (f > u) { o->state = BMX_STATE_LOST_POS_X; u = f; }
        }
        if (y1 != 0 && yp < y1) {
            f = fabs((yp - y1) / (y2 - y1));
            if (f > u) { o->state = BMX_STATE_LOST_NEG_Y; u = f; }
        }
        if (y2 != 0 && yp > y2) {
            f = fabs((yp - y2) / (y2 - y1));
            if (f > u) { o->state = BMX_STATE_LOST_POS_Y; u = f; }
        }
    }
    *uf = u;
}

void bmo_check_aperture_limit(const bmx_ele *ele, int32_t second_edge, double *vec,
                              int32_t *state, double *unstable_factor)
{
    coord_t o;
    memset(&o, 0, sizeof o);
    for (int i = 0; i < 6; i++) o.vec[i] = vec[i];
    o.state = *state;
    o.p0c = ele->p0c;
    check_aperture_limit(&o, ele, second_edge ? SECOND_TRACK_EDGE : FIRST_TRACK_EDGE,
                         unstable_factor);
    for (int i = 0; i < 6; i++) vec[i] = o.vec[i];
    *state = o.state;
}

/* ------------------------------------------------------------------------------------------
 * track1
 * ------------------------------------------------------------------------------------------ */

/* bmad/low_level/track1_bmad.f90:22-283 (charged particles). */
/* ------------------------------------------------------------------------------------------
 * thin multipoles (multipole$, ab_multipole$)
 * ------------------------------------------------------------------------------------------ */

/* sim_utils/special_functions/factorial.f90:13-48: n! from the reference's table of 13-digit
 * decimal literals (exact up to 18!, rounded from 19! on). */
static double bmo_factorial(int n)
{
    static const double f[22] = {
        1.000000000000E+00, 1.000000000000E+00, 2.000000000000E+00, 6.000000000000E+00,
        2.400000000000E+01, 1.200000000000E+02, 7.200000000000E+02, 5.040000000000E+03,
        4.032000000000E+04, 3.628800000000E+05, 3.628800000000E+06, 3.991680000000E+07,
        4.790016000000E+08, 6.227020800000E+09, 8.717829120000E+10, 1.307674368000E+12,
        2.092278988800E+13, 3.556874280960E+14, 6.402373705728E+15, 1.216451004088E+17,
        2.432902008177E+18, 5.109094217171E+19};
    return f[n];
}

/* sim_utils/special_functions/sign_of_mod.f90:41-61 with zero_is_zero = F. */
static double sign_of_nz(double x) { return x < 0 ? -1.0 : 1.0; }

/* bmad/low_level/multipole1_kt_to_ab.f90:18-41. */
static void multipole1_kt_to_ab(double knl, double knsl, double tn, int n, double *an, double *bn)
{
    if (knl == 0 && knsl == 0) {
        *bn = 0;
        *an = 0;
    } else {
        double angle = -tn * (n + 1);
        *bn = (knl * cos(angle) - knsl * sin(angle)) / bmo_factorial(n);
        *an = (knl * sin(angle) + knsl * cos(angle)) / bmo_factorial(n);
    }
}

/* bmad/low_level/multipole1_ab_to_kt.f90:17-47. */
static void multipole1_ab_to_kt(double an, double bn, int n, double *knl, double *tn)
{
    if (an == 0 && bn == 0) {
        *knl = 0;
        *tn = 0;
    } else {
        double a = an, b = bn;
        *knl = bmo_factorial(n) * sqrt(a * a + b * b);
        *tn = -atan2(a, b) / (n + 1);
        if (2 * (n + 1) * fabs(*tn) > PI) {
            *knl = -*knl;
            *tn = sign_of_nz(*tn) * (fabs(*tn) - PI / (n + 1));
        }
    }
}

/* bmad/code/multipole_kt_to_ab.f90:17-34 over orders 0 .. 21. */
void bmo_multipole_kt_to_ab(const double *knl, const double *knsl, const double *tn, double *an,
                            double *bn)
{
    for (int n = 0; n < BMX_N_POLE; n++) multipole1_kt_to_ab(knl[n], knsl[n], tn[n], n, &an[n], &bn[n]);
}

/* bmad/code/multipole_ele_to_kt.f90:29-95 (magnetic, field_master = F). multipole$: the
 * element's knl (a_pole) and tn (b_pole) plus its tilt, or, with skew strengths ksl
 * (a_pole_elec, pointer_to_ele_multipole.f90:30-34), knl = sqrt(knl^2 + ksl^2) and tilt = tn -
 * atan2(ksl, knl) / (n + 1) folded into |tilt| <= pi / (2 (n + 1)), where the element tilt is not
 * added (:57-66). ab_multipole$: multipole_ele_to_ab with use_ele_tilt = T and no kicks (the a/b
 * values unscaled, convert_this_ab.f90:575, rotated by tilt_this_multipole, :441-467), then
 * multipole_ab_to_kt. Returns ix_pole_max (-1 if no order is set). */
static int thin_multipole_ele_to_kt(const bmx_ele *e, double *knl, double *tilt)
{
    for (int n = 0; n < BMX_N_POLE; n++) knl[n] = tilt[n] = 0;
    if (e->key == BMX_MULTIPOLE) {
        int skew = 0;
        if (e->n_multipole_elec > 0)
            for (int n = 0; n < BMX_N_POLE; n++) skew |= e->a_pole_elec[n] != 0;
        for (int n = 0; n < BMX_N_POLE; n++) {
            double k = e->n_multipole > 0 ? e->a_pole[n] : 0;
            double t = e->n_multipole > 0 ? e->b_pole[n] : 0;
            if (!skew) {
                knl[n] = k;
                tilt[n] = t + e->tilt;
            } else {
                double s = e->a_pole_elec[n];
                knl[n] = sqrt(k * k + s * s);
                tilt[n] = t - atan2(s, k) / (n + 1);
                if (2 * (n + 1) * fabs(tilt[n]) > PI) {
                    knl[n] = -knl[n];
                    tilt[n] = sign_of_nz(tilt[n]) * (fabs(tilt[n]) - PI / (n + 1));
                }
            }
        }
        for (int n = BMX_N_POLE - 1; n >= 0; n--)
            if (knl[n] != 0) return n;
        return -1;
    }
    double a[BMX_N_POLE], b[BMX_N_POLE];
    for (int n = 0; n < BMX_N_POLE; n++) {
        a[n] = e->n_multipole > 0 ? e->a_pole[n] : 0;
        b[n] = e->n_multipole > 0 ? e->b_pole[n] : 0;
    }
    int ix = max_nonzero(a, b, BMX_N_POLE - 1);
    if (ix < 0) return -1;
    double t = e->tilt;
    if (t != 0) {
        for (int n = 0; n <= ix; n++) {
            if (a[n] == 0 && b[n] == 0) continue;
            double an = a[n], bn = b[n];
            double cos_t = cos((n + 1) * t), sin_t = sin((n + 1) * t);
            b[n] = bn * cos_t + an * sin_t;
            a[n] = -bn * sin_t + an * cos_t;
        }
    }
    for (int n = 0; n < BMX_N_POLE; n++) multipole1_ab_to_kt(a[n], b[n], n, &knl[n], &tilt[n]);
    return ix;
}

/* bmad/modules/multipole_mod.f90:203-286 multipole_kick, magnetic, for the tracked species equal
 * to the reference species (charge = 1). ref_orb_offset: the dipole order of a multipole$ bends
 * the reference orbit (:250-254). */
static void multipole_kick_kt(double knl, double tilt, int n, coord_t *o, int ref_orb_offset)
{
    if (knl == 0) return;
    const double charge = 1.0;
    double t = tilt, sin_ang, cos_ang, x, y;
    if (t == 0) {
        sin_ang = 0;
        cos_ang = 1;
        x = o->vec[0];
        y = o->vec[2];
    } else {
        sin_ang = sin(t);
        cos_ang = cos(t);
        x = o->vec[0] * cos_ang + o->vec[2] * sin_ang;
        y = -o->vec[0] * sin_ang + o->vec[2] * cos_ang;
    }
    if (n == 0 && ref_orb_offset) {
        o->vec[1] = o->vec[1] + charge * knl * cos_ang * o->vec[5];
        o->vec[3] = o->vec[3] + charge * knl * sin_ang * o->vec[5];
        o->vec[4] = o->vec[4] - charge * knl * (cos_ang * o->vec[0] + sin_ang * o->vec[2]);
        return;
    }
    double x_terms[BMX_N_POLE], y_terms[BMX_N_POLE];
    x_terms[n] = 1.0;
    y_terms[0] = 1.0;
    for (int m = 1; m <= n; m++) {
        x_terms[n - m] = x_terms[n - m + 1] * x;
        y_terms[m] = y_terms[m - 1] * y;
    }
    /* cc = c_multi's table with the 1/n! factor (c_full) */
    double x_value = 0, y_value = 0;
    for (int m = 0; m <= n; m += 2) x_value = x_value + c_multi_tab[n][m] * x_terms[m] * y_terms[m];
    for (int m = 1; m <= n; m += 2) y_value = y_value + c_multi_tab[n][m] * x_terms[m] * y_terms[m];
    double x_vel = charge * knl * x_value;
    double y_vel = charge * knl * y_value;
    if (t == 0) {
        o->vec[1] = o->vec[1] + x_vel;
        o->vec[3] = o->vec[3] + y_vel;
    } else {
        o->vec[1] = o->vec[1] + x_vel * cos_ang - y_vel * sin_ang;
        o->vec[3] = o->vec[3] + x_vel * sin_ang + y_vel * cos_ang;
    }
}

/* bmad/low_level/track1_bmad.f90:142-168: offset_particle without the tilt (the tilt is in the
 * kicks' angles), multipole_kicks (multipole_mod.f90:32-51) over every order with knl /= 0,
 * offset_particle unset. */
static int track_a_thin_multipole(coord_t *o, const bmx_ele *e)
{
    double knl[BMX_N_POLE], tilt[BMX_N_POLE];
    int ix = thin_multipole_ele_to_kt(e, knl, tilt);
    int mis = is_misaligned(e);
    if (mis) offset_particle(e, 1, o, 0, 0, 0, 0);
    if (ix > -1)
        for (int n = 0; n < BMX_N_POLE; n++)
            multipole_kick_kt(knl[n], tilt[n], n, o, e->key == BMX_MULTIPOLE);
    if (mis) offset_particle(e, 0, o, 0, 0, 0, 0);
    return 0;
}

static int track1_bmad(coord_t *o, const bmx_ele *e)
{
    int key = e->key;
    if (!e->is_on && key != BMX_SBEND) key = BMX_DRIFT;
    if (is_thick_multipole_key(key)) return track_a_thick_multipole(o, e);
    switch (key) {
    case BMX_DRIFT:
        /* track1_bmad.f90:93-94: no offset_particle call for drifts (a misaligned drift
         * tracks exactly like an aligned one). */
        track_a_drift(o, e->l * 1);
        return 0;
    case BMX_MARKER: return 0;
    case BMX_QUADRUPOLE: return track_a_quadrupole(o, e);
    case BMX_RFCAVITY: return track_a_rfcavity(o, e);
    case BMX_SBEND: return track_a_bend(o, e);
    case BMX_MULTIPOLE: case BMX_AB_MULTIPOLE: return track_a_thin_multipole(o, e);
    default: return BMX_ERR_UNSUPPORTED;
    }
}

/* ------------------------------------------------------------------------------------------
 * radiation
 * ------------------------------------------------------------------------------------------ */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC11): ten rounds of (hi, lo) = M0 c0, M1 c2; c = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0),
 * the key bumped by the Weyl constants between rounds. */
void bmo_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out)
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* The fluctuation deviates of one particle at one edge of one element (ran_gauss (ran6) in
 * radiation_mod.f90:116): three Philox blocks under the key (seed) with the counter (particle
 * index, turn, element index 1-based, 4 edge + block); from each block two uniforms in (0, 1]
 * of 53 bits, ((w0 << 21 | w1 >> 11) + 1/2) 2^-53 and likewise from w2, w3, and the two
 * Box-Muller deviates sqrt(-2 ln u1) (cos, sin)(2 pi u2). The reference's generator is a
 * sequential one; this stream has the same distribution, not the same numbers. */
void bmo_rad_gauss6(uint64_t seed, uint32_t id, int32_t turn, int32_t ix1, int32_t edge, double *g)
{
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int b = 0; b < 3; b++) {
        const uint32_t ctr[4] = {id, (uint32_t)turn, (uint32_t)ix1, (uint32_t)(4 * edge + b)};
        uint32_t w[4];
        bmo_philox4x32_10(ctr, key, w);
        const uint64_t m1 = ((uint64_t)w[0] << 21) | (w[1] >> 11);
        const uint64_t m2 = ((uint64_t)w[2] << 21) | (w[3] >> 11);
        const double u1 = (m1 + 0.5) * 0x1p-53;  /* in (0, 1]: never 0, so ln u1 is finite */
        const double u2 = (m2 + 0.5) * 0x1p-53;
        const double r = sqrt(-2 * log(u1));
        g[2 * b] = r * cos(TWOPI * u2);
        g[2 * b + 1] = r * sin(TWOPI * u2);
    }
}

/* Where a radiation kick is made: the element, the particle's index and the turn. */
typedef struct {
    const bmx_rad_map *maps;       /* rm0, rm1 of the element */
    const bmx_rad_config *cfg;
    uint32_t id;
    int32_t turn, ix1;
} rad_at_t;

/* bmad/modules/radiation_mod.f90:61-147 track1_radiation for a charged particle, time_dir = 1,
 * the element's maps as given (no radiation_map_setup): the damping kick about ref_orb, the
 * constant part unless zero_average (:109-112), the fluctuation kick stoc_mat g with the lower
 * triangle of stoc_mat (:115-118), lost_pz below pz = -1 (:122-125). vec only: beta is not
 * updated (no statement of the routine touches it). */
static void track1_radiation(coord_t *o, const bmx_ele *e, const rad_at_t *ra, int edge)
{
    const bmx_rad_config *c = ra->cfg;
    if (!c->damping_on && !c->fluctuations_on) return;
    if (e->l == 0) return;
    switch (e->key) {
    case BMX_DRIFT: case BMX_PIPE: case BMX_MARKER: case BMX_MULTIPOLE: case BMX_AB_MULTIPOLE:
        return;
    default: break;
    }
    const bmx_rad_map *m = &ra->maps[edge];
    double *v = o->vec;
    if (c->damping_on) {
        const double rr = c->synch_rad_scale;
        double d[6], w[6];
        for (int j = 0; j < 6; j++) d[j] = v[j] - m->ref_orb[j];
        for (int i = 0; i < 6; i++) {
            double s = 0;
            for (int j = 0; j < 6; j++) s = s + m->damp_dmat[i][j] * d[j];
            w[i] = v[i] + rr * s;
        }
        for (int i = 0; i < 6; i++) v[i] = w[i];
        if (!c->zero_average)
            for (int i = 0; i < 6; i++) v[i] = v[i] + rr * m->xfer_damp_vec[i];
    }
    if (c->fluctuations_on) {
        double g[6];
        bmo_rad_gauss6(c->seed, ra->id, ra->turn, ra->ix1, edge, g);
        for (int i = 0; i < 6; i++) {
            double s = 0;
            for (int j = 0; j <= i; j++) s = s + m->stoc_mat[i][j] * g[j];
            v[i] = v[i] + c->synch_rad_scale * s;
        }
    }
    if (v[5] < -1.0) o->state = BMX_STATE_LOST_PZ;
}

/* bmad/code/track1.f90:44-343, bmad_standard, no spin/space charge/hooks; radiation kicks
 * (:197-202, :319-324) when ra is given and the element is on. */
static int track1_rad(coord_t *o, const bmx_ele *e, const rad_at_t *ra)
{
    if (o->state != BMX_STATE_ALIVE) return 0;
    if (e->p0c != 0 && fabs(e->p0c - o->p0c) > 1e-12 * fabs(e->p0c)) return BMX_ERR_P0C;
    const int rad = ra && e->is_on;
    double u;
    check_aperture_limit(o, e, FIRST_TRACK_EDGE, &u);
    if (rad) track1_radiation(o, e, ra, 0);
    if (o->state != BMX_STATE_ALIVE) return 0;
    int err = track1_bmad(o, e);
    if (err) return err;
    const int no_momentum_change = e->key == BMX_DRIFT || e->key == BMX_MARKER;
    if (no_momentum_change ? BMO_REPEATED_TEST(orbit_too_large(o, NULL)) : orbit_too_large(o, NULL))
        return 0;
    if (o->state != BMX_STATE_ALIVE) return 0;
    if (rad) {
        track1_radiation(o, e, ra, 1);
        /* the kick moved the particle: the orbit test inside check_aperture_limit is a new one */
        if (orbit_too_large(o, &u)) return 0;
    }
    check_aperture_limit(o, e, SECOND_TRACK_EDGE, &u);
    return 0;
}

static int track1(coord_t *o, const bmx_ele *e) { return track1_rad(o, e, NULL); }

int32_t bmo_track1(const bmx_ele *ele, double mc2, int32_t charge, double p0c, double *vec,
                   double *beta, int32_t *state)
{
    coord_t o;
    for (int i = 0; i < 6; i++) o.vec[i] = vec[i];
    o.beta = *beta;
    o.p0c = p0c;
    o.mc2 = mc2;
    o.charge = charge;
    o.state = *state;
    int err = track1(&o, ele);
    for (int i = 0; i < 6; i++) vec[i] = o.vec[i];
    *beta = o.beta;
    *state = o.state;
    return err;
}

/* Bunch loop: ltt_run_beam_mode (lt_tracking_mod.f90:1631) -> track_bunch (beam_mod.f90:159-166)
 * -> track1_bunch_hom (beam_utils.f90:81-86). */
int32_t bmo_track_bunch_rad(const bmx_ele *eles, int32_t n_ele, double mc2, int32_t charge,
                            double p0c, int64_t n, double *x, double *px, double *y, double *py,
                            double *z, double *pz, double *beta, int32_t *state, int32_t *loss_ele,
                            int32_t *loss_turn, int32_t n_turns, int32_t turn0, int32_t n_threads,
                            int64_t *n_tracked, const bmx_rad_map *maps, const bmx_rad_config *cfg)
{
    int32_t err_all = 0;
    int64_t tracked = 0;
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    for (int32_t it = 0; it < n_turns; it++) {
        for (int32_t ie = 0; ie < n_ele; ie++) {
            const bmx_ele *e = &eles[ie];
            int64_t cnt = 0;
            int32_t err_e = 0;
#pragma omp parallel for schedule(static) reduction(+ : cnt) reduction(| : err_e)
            for (int64_t ip = 0; ip < n; ip++) {
                if (state[ip] != BMX_STATE_ALIVE) continue;
                coord_t o;
                o.vec[0] = x[ip]; o.vec[1] = px[ip]; o.vec[2] = y[ip];
                o.vec[3] = py[ip]; o.vec[4] = z[ip]; o.vec[5] = pz[ip];
                o.beta = beta[ip];
                o.p0c = p0c;
                o.mc2 = mc2;
                o.charge = charge;
                o.state = BMX_STATE_ALIVE;
                rad_at_t ra = {maps ? maps + 2 * ie : NULL, cfg, (uint32_t)ip, turn0 + it, ie + 1};
                int err = track1_rad(&o, e, maps ? &ra : NULL);
                err_e |= err;
                cnt++;
                x[ip] = o.vec[0]; px[ip] = o.vec[1]; y[ip] = o.vec[2];
                py[ip] = o.vec[3]; z[ip] = o.vec[4]; pz[ip] = o.vec[5];
                beta[ip] = o.beta;
                if (o.state != BMX_STATE_ALIVE) {
                    state[ip] = o.state;
                    loss_ele[ip] = ie + 1;
                    loss_turn[ip] = turn0 + it;
                }
            }
            tracked += cnt;
            if (err_e) err_all = err_e;
        }
        if (err_all) break;
    }
    if (n_tracked) *n_tracked = tracked;
    return err_all;
}

int32_t bmo_track_bunch(const bmx_ele *eles, int32_t n_ele, double mc2, int32_t charge,
                        double p0c, int64_t n, double *x, double *px, double *y, double *py,
                        double *z, double *pz, double *beta, int32_t *state, int32_t *loss_ele,
                        int32_t *loss_turn, int32_t n_turns, int32_t turn0, int32_t n_threads,
                        int64_t *n_tracked)
{
    return bmo_track_bunch_rad(eles, n_ele, mc2, charge, p0c, n, x, px, y, py, z, pz, beta, state,
                               loss_ele, loss_turn, n_turns, turn0, n_threads, n_tracked, NULL, NULL);
}

int32_t bmo_track_trace(const bmx_ele *eles, int32_t n_ele, double mc2, int32_t charge,
                        double p0c, double *vec, double *beta, int32_t *state, int32_t n_turns,
                        double *out)
{
    coord_t o;
    for (int i = 0; i < 6; i++) o.vec[i] = vec[i];
    o.beta = *beta;
    o.p0c = p0c;
    o.mc2 = mc2;
    o.charge = charge;
    o.state = *state;
    int64_t k = 0;
    for (int32_t it = 0; it < n_turns; it++) {
        for (int32_t ie = 0; ie < n_ele; ie++) {
            int err = track1(&o, &eles[ie]);
            if (err) return err;
            for (int i = 0; i < 6; i++) out[k++] = o.vec[i];
        }
    }
    for (int i = 0; i < 6; i++) vec[i] = o.vec[i];
    *beta = o.beta;
    *state = o.state;
    return 0;
}
