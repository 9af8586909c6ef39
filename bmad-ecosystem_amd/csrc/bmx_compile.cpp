// Lattice table compiler: raw bmx_ele records -> compact device table.
//
// Per element this does, once, the element-constant work that Bmad's element routines redo per
// particle per element: multipole normalisation and kick folding (multipole_ele_to_ab.f90:
// 31-236, 326-405, 472-600), step counts (track_a_thick_multipole.f90:73-79,
// track_a_rfcavity.f90:54), Hwang edge constants (fringe_mod.f90:264-288), bend-angle
// functions (track_a_bend.f90:139-175) and aperture geometry (check_aperture_limit.f90:358-385).
// Anything the GPU maps do not implement is rejected with BMX_ERR_UNSUPPORTED, mirroring
// track1_bmad's "not implemented" error (track1_bmad.f90:263-270).
#include <cmath>
#include <cstdio>
#include <cstring>

#include "bmx_internal.hpp"

namespace bmx {
namespace {

constexpr double kMaxAp = 1e3;

// sim_utils/special_functions/sinc.f90, cosc.f90, sincc.f90 (nd = 0)
double h_sinc(double x) { return std::fabs(x) < 1e-8 ? 1.0 : std::sin(x) / x; }
double h_cosc(double x)
{
    if (std::fabs(x) < 1e-8) return 0.5;
    double s = h_sinc(x / 2);
    return 0.5 * (s * s);
}
double h_sincc(double x)
{
    const double c0 = 1.0 / 6.0, c1 = -1.0 / 120.0, c2 = 1.0 / 5040.0, c3 = -1.0 / 362880.0;
    double x2 = x * x;
    if (std::fabs(x) < 0.1) return c0 + x2 * (c1 + x2 * (c2 + x2 * c3));
    return (x - std::sin(x)) / (x * x2);
}

long f_nint(double x) { return std::lround(x); }

bool misaligned(const bmx_ele &e)
{
    return e.x_offset != 0 || e.y_offset != 0 || e.z_offset != 0 || e.x_pitch != 0 ||
           e.y_pitch != 0 || e.tilt != 0 || e.ref_tilt != 0;
}

// bmad/geometry/floor_angles_to_w_mat.f90:17-57 for (theta, phi, psi = 0): the rotation of
// offset_particle's frame change (inverse: entering the element), row-major.
void pitch_matrix(double theta, double phi, bool inverse, double *w)
{
    const double s_the = std::sin(theta), c_the = std::cos(theta);
    const double s_phi = std::sin(phi), c_phi = std::cos(phi);
    const double s_psi = std::sin(0.0), c_psi = std::cos(0.0);
    if (!inverse) {
        const double m[9] = {c_the * c_psi - s_the * s_phi * s_psi, -c_the * s_psi - s_the * s_phi * c_psi,
                             s_the * c_phi, c_phi * s_psi, c_phi * c_psi, s_phi,
                             -s_the * c_psi - c_the * s_phi * s_psi, s_the * s_psi - c_the * s_phi * c_psi,
                             c_the * c_phi};
        std::memcpy(w, m, sizeof m);
    } else {
        const double m[9] = {c_the * c_psi - s_the * s_phi * s_psi, c_phi * s_psi,
                             -s_the * c_psi - c_the * s_phi * s_psi, -c_the * s_psi - s_the * s_phi * c_psi,
                             c_phi * c_psi, s_the * s_psi - c_the * s_phi * c_psi, s_the * c_phi, s_phi,
                             c_the * c_phi};
        std::memcpy(w, m, sizeof m);
    }
}

// The offset block of a misaligned or tilted element (bmx_table.hpp), offset_particle's
// non-bend branch (offset_particle.f90:183-492).
void offset_block(const bmx_ele &e, int32_t &flags, double *ob)
{
    const bool mis = e.x_offset != 0 || e.y_offset != 0 || e.z_offset != 0 || e.x_pitch != 0 ||
                     e.y_pitch != 0;
    for (int k = 0; k < kOffsetLen; k++) ob[k] = 0;
    ob[0] = e.x_offset;
    ob[1] = e.y_offset;
    ob[2] = e.z_offset;
    ob[3] = 0.5 * e.l;
    pitch_matrix(e.x_pitch, e.y_pitch, true, ob + 4);
    pitch_matrix(e.x_pitch, e.y_pitch, false, ob + 13);
    ob[22] = std::cos(e.tilt);
    ob[23] = std::sin(e.tilt);
    flags |= F_OFFSET;
    if (mis) flags |= F_MISALIGN;
    if (e.tilt != 0) flags |= F_TILT;
}

std::string where(int i, const bmx_ele &e)
{
    char buf[96];
    std::snprintf(buf, sizeof buf, "element %d (key %d): ", i + 1, e.key);
    return buf;
}

// Magnetic a/b strengths in the body frame incl. kicks (multipole_ele_to_ab with
// include_kicks$, use_ele_tilt = F). Returns ix_pole_max or -2 on unsupported scaling.
int ele_to_ab(const bmx_ele &e, double *a, double *b)
{
    for (int n = 0; n < BMX_N_POLE; n++) a[n] = b[n] = 0;
    if (!e.is_on || e.key == BMX_PIPE) return -1;
    if (e.n_multipole > 0) {
        for (int n = 0; n < BMX_N_POLE; n++) {
            a[n] = e.a_pole[n];
            b[n] = e.b_pole[n];
        }
        bool any = false;
        for (int n = 0; n < BMX_N_POLE; n++) any |= (a[n] != 0 || b[n] != 0);
        if (any) {
            double cnst = 0;
            int ref_exp = 0;
            bool scale = true;
            switch (e.key) {
            case BMX_SBEND: cnst = e.l * (e.g + e.dg); break;
            case BMX_KICKER:
                cnst = e.hkick == 0 ? e.vkick : (e.vkick == 0 ? e.hkick : std::sqrt(e.hkick * e.hkick + e.vkick * e.vkick));
                break;
            case BMX_HKICKER: case BMX_VKICKER: cnst = e.kick; break;
            case BMX_QUADRUPOLE: cnst = e.k1 * e.l; ref_exp = 1; break;
            case BMX_SEXTUPOLE: cnst = e.k2 * e.l; ref_exp = 2; break;
            case BMX_OCTUPOLE: cnst = e.k3 * e.l; ref_exp = 3; break;
            case BMX_THICK_MULTIPOLE: case BMX_RFCAVITY: scale = false; break;
            default: return -2;
            }
            if (scale) {
                if (e.r0_mag == 0) {
                    for (int n = 0; n < BMX_N_POLE; n++) { a[n] = cnst * a[n]; b[n] = cnst * b[n]; }
                } else {
                    double factor = cnst * std::pow(e.r0_mag, ref_exp + 1);
                    for (int n = 0; n < BMX_N_POLE; n++) {
                        factor = factor / e.r0_mag;
                        a[n] = factor * a[n];
                        b[n] = factor * b[n];
                    }
                }
            }
        }
    }
    double ak[4] = {0, 0, 0, 0}, bk[4] = {0, 0, 0, 0};
    switch (e.key) {
    case BMX_HKICKER: bk[0] = -e.kick; break;
    case BMX_VKICKER: ak[0] = e.kick; break;
    case BMX_KICKER: bk[0] = -e.hkick; ak[0] = e.vkick; break;
    case BMX_DRIFT: case BMX_MARKER: break;
    default:
        if (e.hkick != 0 || e.vkick != 0) {
            double t = e.key == BMX_SBEND ? e.ref_tilt : e.tilt;
            if (t != 0) {
                double c = std::cos(t), s = std::sin(t);
                bk[0] = -e.hkick * c - e.vkick * s;
                ak[0] = -e.hkick * s + e.vkick * c;
            } else {
                bk[0] = -e.hkick;
                ak[0] = e.vkick;
            }
        }
    }
    switch (e.key) {
    case BMX_QUADRUPOLE: bk[1] = e.k1 * e.l; break;
    case BMX_SBEND:
        bk[0] = bk[0] + e.dg * e.l;
        bk[1] = e.k1 * e.l;
        bk[2] = e.k2 * e.l / 2.0;
        break;
    case BMX_SEXTUPOLE: bk[2] = e.k2 * e.l / 2.0; break;
    case BMX_OCTUPOLE: bk[3] = e.k3 * e.l / 6.0; break;
    default: break;
    }
    for (int i = 0; i < 4; i++) {
        a[i] = a[i] + ak[i];
        b[i] = b[i] + bk[i];
    }
    for (int n = BMX_N_POLE - 1; n >= 0; n--)
        if (a[n] != 0 || b[n] != 0) return n;
    return -1;
}

// The highest order below n_max with a nonzero strength, or -1 (max_nonzero).
int max_order(const double *a, const double *b, int n_max)
{
    for (int n = n_max - 1; n >= 0; n--)
        if (a[n] != 0 || b[n] != 0) return n;
    return -1;
}

// Thin multipole strengths as a/b values (track1_bmad.f90:142-168). multipole$: knl and the
// order tilts from multipole_ele_to_kt (multipole_ele_to_kt.f90:46-76): knl = a_pole, tilt =
// b_pole + the element tilt, or, with skew strengths knsl (a_pole_elec), knl = sqrt(knl^2 +
// knsl^2) and tilt = tn - atan2(knsl, knl) / (n + 1) folded into |tilt| <= pi / (2 (n + 1)),
// where the element tilt is not added. The kick of an order (multipole_kick, multipole_mod.f90:
// 203-286) is then the a/b kick with b_n = knl cos((n+1) tilt) / n!, a_n = -knl sin((n+1) tilt)
// / n! (multipole1_kt_to_ab.f90), n! by repeated multiplication as c_multi's 1/n! factor;
// k0 / t0 return the dipole order's knl and tilt (its reference-orbit form). ab_multipole$:
// a_pole / b_pole unscaled (convert_this_ab.f90:575) and rotated by the element tilt
// (tilt_this_multipole, multipole_ele_to_ab.f90:441-467). Returns the highest order present
// (-1 none).
int thin_multipole_ab(const bmx_ele &e, double *a, double *b, double &k0, double &t0)
{
    const double pi = 3.141592653589793238462643383279;
    for (int n = 0; n < BMX_N_POLE; n++) a[n] = b[n] = 0;
    k0 = t0 = 0;
    if (e.key == BMX_MULTIPOLE) {
        bool skew = false;
        if (e.n_multipole_elec > 0)
            for (int n = 0; n < BMX_N_POLE; n++) skew |= e.a_pole_elec[n] != 0;
        double fact = 1;
        int ix = -1;
        for (int n = 0; n < BMX_N_POLE; n++) {
            if (n > 0) fact = n * fact;
            double knl = e.n_multipole > 0 ? e.a_pole[n] : 0;
            double t = e.n_multipole > 0 ? e.b_pole[n] : 0;
            if (!skew) {
                t = t + e.tilt;
            } else {
                const double ks = e.a_pole_elec[n];
                const double k = knl;
                knl = std::sqrt(k * k + ks * ks);
                t = t - std::atan2(ks, k) / (n + 1);
                if (2 * (n + 1) * std::fabs(t) > pi) {
                    knl = -knl;
                    t = (t < 0 ? -1.0 : 1.0) * (std::fabs(t) - pi / (n + 1));
                }
            }
            if (knl == 0) continue;
            ix = n;
            if (n == 0) {
                k0 = knl;
                t0 = t;
            }
            const double angle = -t * (n + 1);
            b[n] = knl * std::cos(angle) / fact;
            a[n] = knl * std::sin(angle) / fact;
        }
        return ix;
    }
    for (int n = 0; n < BMX_N_POLE; n++) {
        a[n] = e.n_multipole > 0 ? e.a_pole[n] : 0;
        b[n] = e.n_multipole > 0 ? e.b_pole[n] : 0;
    }
    int ix = max_order(a, b, BMX_N_POLE);
    if (e.tilt != 0) {
        for (int n = 0; n <= ix; n++) {
            if (a[n] == 0 && b[n] == 0) continue;
            const double an = a[n], bn = b[n];
            const double c = std::cos((n + 1) * e.tilt), sn = std::sin((n + 1) * e.tilt);
            b[n] = bn * c + an * sn;
            a[n] = -bn * sn + an * c;
        }
    }
    return ix;
}

// Electric a/b strengths (multipole_ele_to_ab.f90:31-236 with pole_type = electric$):
// a_pole_elec / b_pole_elec, order n scaled by r0_elec^-n unless r0_elec is 0 or 1
// (convert_this_ab, :519-530). Returns ix_pole_max.
int ele_to_ab_elec(const bmx_ele &e, double *a, double *b)
{
    for (int n = 0; n < BMX_N_POLE; n++) a[n] = b[n] = 0;
    if (!e.is_on || e.key == BMX_PIPE || e.n_multipole_elec <= 0) return -1;
    for (int n = 0; n < BMX_N_POLE; n++) {
        a[n] = e.a_pole_elec[n];
        b[n] = e.b_pole_elec[n];
    }
    const double radius = e.r0_elec;
    if (radius != 0 && radius != 1) {
        double factor = radius;
        for (int n = 0; n < BMX_N_POLE; n++) {
            factor = factor / radius;
            a[n] = factor * a[n];
            b[n] = factor * b[n];
        }
    }
    return max_order(a, b, BMX_N_POLE);
}

// Series terms for the quadrupole's fast form over a step with b1 s^2/L = w_l: W = w_l/(1+pz),
// sized for 1+pz >= 0.5 (further off momentum the kernel reduces the argument); the fewest
// terms whose first omitted term t^M/(4M)! is below 1e-18, and the t range they cover.
double quad_series_limit(double w_l, int &nterms)
{
    const double wmax = std::fabs(w_l) / 0.5;
    const double tmax = wmax * wmax;
    double fact[40];
    fact[0] = 1;
    for (int k = 1; k < 40; k++) fact[k] = fact[k - 1] * k;
    nterms = 8;
    for (int m = 2; m <= 8; m++) {
        if (std::pow(tmax, m) / fact[4 * m] <= 1e-18) { nterms = m; break; }
    }
    return std::pow(1e-18 * fact[4 * nterms], 1.0 / nterms);
}

// Aperture flags and the 4 aperture values; returns false if unsupported.
bool aperture_setup(const bmx_ele &e, int32_t &flags, double ap[4], double mom[8], std::string &why)
{
    int at = e.aperture_at;
    if (at == BMX_CONTINUOUS) { why = "continuous apertures"; return false; }
    if (at != BMX_ENTRANCE_END && at != BMX_EXIT_END && at != BMX_BOTH_ENDS) return true;
    int32_t f = 0;
    // collimator momentum windows (check_aperture_limit.f90, tested before the x-y aperture)
    const bool windows = (e.key == BMX_ECOLLIMATOR || e.key == BMX_RCOLLIMATOR) &&
                         (e.px_aperture_width2 != 0 || e.py_aperture_width2 != 0 ||
                          e.z_aperture_width2 != 0 || e.pz_aperture_width2 != 0);
    if (windows) {
        const double m[8] = {e.px_aperture_width2, e.px_aperture_center, e.py_aperture_width2,
                             e.py_aperture_center, e.z_aperture_width2,  e.z_aperture_center,
                             e.pz_aperture_width2, e.pz_aperture_center};
        std::memcpy(mom, m, sizeof m);
        f |= F_APE_MOMENTUM;
        if (e.key == BMX_ECOLLIMATOR) f |= F_APE_MOM_ELLIP;
    }
    const double x1 = -e.x1_limit, x2 = e.x2_limit, y1 = -e.y1_limit, y2 = e.y2_limit;
    if (e.aperture_type == BMX_APERTURE_ELLIPTICAL) {
        // elliptical_params_setup: a half-set pair makes the reference skip the x-y test.
        if ((x1 == 0) != (x2 == 0) || (y1 == 0) != (y2 == 0)) {
            if (!windows) return true;
            why = "momentum windows with a half-set elliptical aperture";
            return false;
        }
        double xw = (x2 - x1) / 2, xc = (x2 + x1) / 2;
        double yw = (y2 - y1) / 2, yc = (y2 + y1) / 2;
        if (xw == 0) xw = kMaxAp;
        if (yw == 0) yw = kMaxAp;
        xw = xw > 0 ? xw : 0;
        yw = yw > 0 ? yw : 0;
        ap[0] = xc;
        ap[1] = 1 / xw;
        ap[2] = yc;
        ap[3] = 1 / yw;
        f |= F_APE_ELLIP;
        if (xc == 0 && yc == 0) f |= F_APE_CENTERED;
        if (std::fabs(xc) + xw < 0.5 * kMaxAp && std::fabs(yc) + yw < 0.5 * kMaxAp) f |= F_APE_EXIT_COVERS;
    } else if (e.aperture_type == BMX_APERTURE_RECTANGULAR || e.aperture_type == 1 /* auto */ ||
               e.aperture_type == 0) {
        if (x1 == 0 && x2 == 0 && y1 == 0 && y2 == 0 && !windows) return true;
        if (x1 != 0 && x2 != 0 && y1 != 0 && y2 != 0 && std::fabs(x1) < 0.5 * kMaxAp &&
            std::fabs(x2) < 0.5 * kMaxAp && std::fabs(y1) < 0.5 * kMaxAp && std::fabs(y2) < 0.5 * kMaxAp)
            f |= F_APE_EXIT_COVERS;
        ap[0] = x1;
        ap[1] = x2;
        ap[2] = y1;
        ap[3] = y2;
    } else {
        why = "aperture type";
        return false;
    }
    if (at == BMX_ENTRANCE_END || at == BMX_BOTH_ENDS) f |= F_APE_ENT;
    if (at == BMX_EXIT_END || at == BMX_BOTH_ENDS) f |= F_APE_EXIT;
    if (!(f & F_APE_EXIT)) f &= ~F_APE_EXIT_COVERS;
    flags |= f;
    return true;
}

void edge_constants(const bmx_ele &e, bool entrance, double *out)
{
    const double c_dir = 1.0;
    double g_tot = (e.g + e.dg) * c_dir;
    double k1 = e.k1;
    double ang = entrance ? e.e1 : e.e2;
    double fint_gap = entrance ? e.fint * e.hgap : e.fintx * e.hgapx;
    if (e.fringe_type == BMX_FRINGE_HARD_EDGE_ONLY) fint_gap = 0;
    double cos_e = std::cos(ang), sin_e = std::sin(ang);
    double tan_e = sin_e / cos_e, sec_e = 1 / cos_e;
    out[0] = g_tot * tan_e;                                             // gt
    out[1] = g_tot * (tan_e * tan_e);                                   // gt2
    out[2] = g_tot * (sec_e * sec_e);                                   // gs2
    out[3] = k1 * tan_e * c_dir;                                        // k1_tane
    out[4] = 2 * fint_gap * out[2] * g_tot * sec_e * (1 + sin_e * sin_e);  // fg_factor
    out[5] = tan_e * tan_e;
}

// The body block of a sector bend for one step of length L (the K_BEND parameter layout):
// g_tot = g + dg of the body; the edge constants are those of the element.
void bend_body_block(const bmx_ele &e, double L, double g, double g_tot, double *q)
{
    const double angle = g * L;
    q[0] = L;
    q[1] = g;
    q[2] = g_tot * 1.0;
    q[3] = angle;
    q[4] = std::cos(angle);
    q[5] = std::sin(angle);
    q[6] = h_sinc(angle);
    q[7] = h_cosc(angle);
    q[8] = e.p0c / e.e_tot;
    q[9] = h_sincc(angle);
    edge_constants(e, true, &q[10]);
    edge_constants(e, false, &q[16]);
    q[22] = L * L * g * q[7];
    q[23] = L / q[8];
    q[24] = L * q[6];
    q[25] = 0.5 * q[12];
    q[26] = 0.5 * q[18];
    q[27] = angle * angle * q[7];  // 1 - cos(angle) without cancellation
    q[28] = g != 0 ? 1.0 / g : 0.0;
}

// ---- the bend branch of offset_particle (offset_particle.f90:206-232, 404-430) ----
// A point and frame; every step below maps r -> S r + t and w -> S w, so the whole frame change
// of a particle at (x, y, 0) with the unit frame is r = W (x, y, 0) + c, w = W: W and c follow
// from composing the steps once, here.
struct Frame {
    double r[3];
    double w[3][3];
};

void frame_apply(Frame &f, const double S[3][3], const double *t)
{
    double r[3], w[3][3];
    for (int i = 0; i < 3; i++) {
        r[i] = S[i][0] * f.r[0] + S[i][1] * f.r[1] + S[i][2] * f.r[2] + (t ? t[i] : 0.0);
        for (int j = 0; j < 3; j++) w[i][j] = S[i][0] * f.w[0][j] + S[i][1] * f.w[1][j] + S[i][2] * f.w[2][j];
    }
    std::memcpy(f.r, r, sizeof r);
    std::memcpy(f.w, w, sizeof w);
}

// Rotation by angle about axis (0 = x, 1 = y, 2 = z), as applied by rotate_vec / rotate_mat
// (sim_utils/math/rotation_3d_mod.f90:737-835).
void axis_rotation(int axis, double angle, double R[3][3])
{
    const double c = std::cos(angle), s = std::sin(angle);
    const double I[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    std::memcpy(R, I, sizeof I);
    if (axis == 0) { R[1][1] = c; R[1][2] = -s; R[2][1] = s; R[2][2] = c; }
    else if (axis == 1) { R[0][0] = c; R[0][2] = s; R[2][0] = -s; R[2][2] = c; }
    else { R[0][0] = c; R[0][1] = -s; R[1][0] = s; R[1][1] = c; }
}

void mat3_mul(const double A[3][3], const double B[3][3], double C[3][3])
{
    double T[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) T[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
    std::memcpy(C, T, sizeof T);
}

// bmad/geometry/bend_shift.f90:19-61: along an arc of curvature g by delta_s (with the bend's
// ref_tilt when with_tilt).
void bend_shift_frame(Frame &f, double g, double delta_s, bool with_tilt, double ref_tilt)
{
    const double angle = delta_s * g;
    if (angle == 0) {
        f.r[2] = f.r[2] - delta_s;
        return;
    }
    double S[3][3], R[3][3];
    axis_rotation(1, angle, S);
    if (with_tilt && ref_tilt != 0) {
        axis_rotation(2, -ref_tilt, R);
        double T[3][3];
        mat3_mul(S, R, T);
        axis_rotation(2, ref_tilt, R);
        mat3_mul(R, T, S);
    }
    double L[3] = {-2 * std::pow(std::sin(angle / 2), 2) / g, 0.0, -std::sin(angle) / g};
    if (with_tilt && ref_tilt != 0) {
        const double c = std::cos(ref_tilt), sn = std::sin(ref_tilt);
        const double l0 = c * L[0] - sn * L[1], l1 = sn * L[0] + c * L[1];
        L[0] = l0;
        L[1] = l1;
    }
    frame_apply(f, S, L);
}

// out[0..8] W and [9..11] d of the entering frame change (r = W ((x, y, 0) - d)), [12..20] W and
// [21..23] c of the leaving one (r = W (x, y, 0) + c, including the move to the exit face), for
// a misaligned bend (orientation +1, drift to the edges).
void bend_frame_blocks(const bmx_ele &e, double *out)
{
    const double g = e.g, rt = e.ref_tilt, L_half = 0.5 * e.l;
    double Sm[3][3], St[3][3], wm[9];
    pitch_matrix(e.x_pitch, e.y_pitch, false, wm);  // floor_angles_to_w_mat(x_pitch, y_pitch, 0)
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            Sm[i][j] = wm[3 * i + j];
            St[j][i] = wm[3 * i + j];
        }
    const double offs[3] = {e.x_offset, e.y_offset, e.z_offset};
    // entering: ds_center = L/2
    Frame f = {{0, 0, 0}, {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    bend_shift_frame(f, g, L_half, true, rt);
    {
        double t[3];
        for (int i = 0; i < 3; i++) t[i] = -(St[i][0] * offs[0] + St[i][1] * offs[1] + St[i][2] * offs[2]);
        frame_apply(f, St, t);
    }
    if (rt != 0) {
        bend_shift_frame(f, g, -L_half, true, rt);
        double R[3][3];
        axis_rotation(2, -rt, R);
        frame_apply(f, R, nullptr);
        bend_shift_frame(f, g, L_half, false, 0.0);
    }
    bend_shift_frame(f, g, -L_half, false, 0.0);
    // r = W (x, y, 0) + c with W = f.w, c = f.r; d = -W^T c
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[3 * i + j] = f.w[i][j];
    for (int j = 0; j < 3; j++) out[9 + j] = -(f.w[0][j] * f.r[0] + f.w[1][j] * f.r[1] + f.w[2][j] * f.r[2]);
    // leaving: ds_center = -L/2
    Frame u = {{0, 0, 0}, {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    bend_shift_frame(u, g, -L_half, false, 0.0);
    if (rt != 0) {
        bend_shift_frame(u, g, -L_half, false, 0.0);
        double R[3][3];
        axis_rotation(2, rt, R);
        frame_apply(u, R, nullptr);
        bend_shift_frame(u, g, L_half, true, rt);
    }
    frame_apply(u, Sm, offs);
    bend_shift_frame(u, g, L_half, true, rt);
    u.r[2] = u.r[2] + 2 * L_half;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[12 + 3 * i + j] = u.w[i][j];
    for (int i = 0; i < 3; i++) out[21 + i] = u.r[i];
}

}  // namespace

RadCfg rad_settings(const bmx_rad_config &c)
{
    RadCfg r{};
    r.scale = c.synch_rad_scale;
    r.flags = (c.damping_on ? RAD_DAMP : 0u) | (c.fluctuations_on ? RAD_FLUCT : 0u) |
              (c.zero_average ? RAD_ZERO_AVERAGE : 0u);
    r.key0 = (uint32_t)c.seed;
    r.key1 = (uint32_t)(c.seed >> 32);
    return r;
}

namespace {

// Kinds that radiate in track1_radiation (radiation_mod.f90:86-93): all with a length but
// drifts, pipes, markers and thin multipoles (the other kinds it leaves out are not on this
// path).
bool radiates(const bmx_ele &e)
{
    if (!e.is_on || e.l == 0) return false;
    switch (e.key) {
    case BMX_DRIFT: case BMX_PIPE: case BMX_MARKER: case BMX_MULTIPOLE: case BMX_AB_MULTIPOLE:
        return false;
    default: return true;
    }
}

bool map_is_zero(const bmx_rad_map &m)
{
    for (int i = 0; i < 6; i++) {
        if (m.xfer_damp_vec[i] != 0) return false;
        for (int j = 0; j < 6; j++)
            if (m.damp_dmat[i][j] != 0 || (j <= i && m.stoc_mat[i][j] != 0)) return false;
    }
    return true;
}

// One half's map in the radiation block layout (bmx_table.hpp).
void rad_edge_block(const bmx_rad_map &m, double *out)
{
    for (int i = 0; i < 6; i++) out[i] = m.ref_orb[i];
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) out[6 + 6 * i + j] = m.damp_dmat[i][j];
    for (int i = 0; i < 6; i++) out[42 + i] = m.xfer_damp_vec[i];
    int k = 48;
    for (int i = 0; i < 6; i++)
        for (int j = 0; j <= i; j++) out[k++] = m.stoc_mat[i][j];
}

bool finite_map(const bmx_rad_map &m)
{
    const double *v = &m.ref_orb[0];
    for (size_t i = 0; i < sizeof(bmx_rad_map) / sizeof(double); i++)
        if (!std::isfinite(v[i])) return false;
    return true;
}

}  // namespace

int compile_table(const bmx_ele *eles, int32_t n_ele, double mc2, int32_t charge, HostTable &out,
                  std::string &err, const bmx_rad_map *rad, const bmx_rad_config *rad_cfg)
{
    // the tracked species is the reference species: magnetic charge ratios are 1; the charge
    // enters only the electric multipole strengths
    out.hdr.clear();
    out.prm.clear();
    out.kind_mask = 0;
    out.rare_q = false;
    out.rad = false;
    if (n_ele <= 0 || !eles) { err = "empty lattice"; return BMX_ERR_ARG; }
    if (!(mc2 > 0)) { err = "mass must be positive"; return BMX_ERR_ARG; }
    const double p0c = eles[0].p0c;
    if (!(p0c > 0)) { err = "element 1 has no reference momentum"; return BMX_ERR_ARG; }
    out.sp.mc2 = mc2;
    out.sp.p0c = p0c;
    out.sp.mc2_sq = mc2 * mc2;
    const double e_tot = eles[0].e_tot;
    if (!(e_tot > 0)) { err = "element 1 has no reference energy"; return BMX_ERR_ARG; }
    out.sp.beta0 = p0c / e_tot;
    {
        const double me = mc2 / e_tot;
        out.sp.me2 = me * me;
        out.sp.lez_f = out.sp.beta0 * out.sp.beta0 * (2 * (out.sp.beta0 * out.sp.beta0) - (me * me) / 2);
    }
    out.sp.lez_thr = 3e-7 * e_tot;
    out.sp.rad = RadCfg{};
    if (rad) {
        if (!rad_cfg) { err = "radiation maps without settings"; return BMX_ERR_ARG; }
        if (rad_cfg->reserved != 0 || !std::isfinite(rad_cfg->synch_rad_scale)) {
            err = "bad radiation settings";
            return BMX_ERR_ARG;
        }
        out.sp.rad = rad_settings(*rad_cfg);
    }

    for (int i = 0; i < n_ele; i++) {
        const bmx_ele &e = eles[i];
        if (e.key > 0 && e.key < 64) out.kind_mask |= (uint64_t)1 << e.key;
        if (e.p0c != p0c || e.e_tot != e_tot) {
            err = where(i, e) + "reference momentum differs from element 1 (energy changing lattices are outside this path)";
            return BMX_ERR_P0C;
        }
        EleHdr h{};
        std::vector<double> q;
        std::string why;
        int key = e.key;
        if (!e.is_on) {
            if (key == BMX_SBEND) { err = where(i, e) + "switched-off bends are not supported"; return BMX_ERR_UNSUPPORTED; }
            key = BMX_DRIFT;  // track1_bmad.f90:51-60
        }
        // Misalignments: drifts and markers ignore them in the reference (track1_bmad.f90:93-94);
        // the other non-bend kinds get offset_particle's frame change around their map; bends
        // take the bend branch inside their own map (K_BMULT).
        // a thin multipole's frame change leaves out its tilt (offset_particle with set_tilt =
        // F, track1_bmad.f90:144): the tilt is in its kicks
        // radiation kicks at the element's ends (track1_radiation), when its maps are not zero
        bool rad_here = false;
        if (rad && radiates(e)) {
            if (!finite_map(rad[2 * i]) || !finite_map(rad[2 * i + 1])) {
                err = where(i, e) + "radiation map with a value that is not finite";
                return BMX_ERR_ARG;
            }
            rad_here = !map_is_zero(rad[2 * i]) || !map_is_zero(rad[2 * i + 1]);
        }
        const bool thin = key == BMX_MULTIPOLE || key == BMX_AB_MULTIPOLE;
        bmx_ele e_frame = e;
        if (thin) e_frame.tilt = 0;
        const bool offset = key != BMX_DRIFT && key != BMX_MARKER && key != BMX_SBEND && misaligned(e_frame);
        if (offset && e.ref_tilt != 0) {
            err = where(i, e) + "ref_tilt on an element that is not a bend";
            return BMX_ERR_UNSUPPORTED;
        }
        if ((offset || (key == BMX_SBEND && misaligned(e))) && e.offset_moves_aperture &&
            (e.x1_limit != 0 || e.x2_limit != 0 || e.y1_limit != 0 || e.y2_limit != 0)) {
            err = where(i, e) + "apertures that move with the element offsets are not supported";
            return BMX_ERR_UNSUPPORTED;
        }
        double a[BMX_N_POLE], b[BMX_N_POLE];
        // (a multipole$ keeps its skew strengths in a_pole_elec)
        if (key != BMX_QUADRUPOLE && key != BMX_SBEND && key != BMX_MULTIPOLE &&
            ele_to_ab_elec(e, a, b) >= 0) {
            err = where(i, e) + "electric multipoles are supported on quadrupoles and bends only";
            return BMX_ERR_UNSUPPORTED;
        }
        switch (key) {
        case BMX_DRIFT:
        case BMX_MARKER:
        case BMX_PIPE:
        case BMX_MONITOR:
        case BMX_INSTRUMENT:
        case BMX_ECOLLIMATOR:
        case BMX_RCOLLIMATOR:
        case BMX_SEXTUPOLE:
        case BMX_OCTUPOLE:
        case BMX_THICK_MULTIPOLE:
        case BMX_KICKER:
        case BMX_HKICKER:
        case BMX_VKICKER: {
            int ix = (key == BMX_DRIFT || key == BMX_MARKER) ? -1 : ele_to_ab(e, a, b);
            if (ix == -2) { err = where(i, e) + "multipole scaling"; return BMX_ERR_UNSUPPORTED; }
            if (ix > 3) { err = where(i, e) + "multipole orders above 3"; return BMX_ERR_UNSUPPORTED; }
            if (key == BMX_MARKER || (ix == -1 && e.l == 0)) {
                h.kind = K_MARKER;
            } else if (ix == -1) {
                h.kind = K_DRIFT;
                q.push_back(e.l);
            } else {
                h.kind = K_MULT;
                long ns = f_nint(std::fabs(e.l / e.ds_step));
                int n_step = ns > 1 ? (int)ns : 1;
                double r_step = 1.0 / n_step;
                double half = 1.0 * r_step / 2;
                q.push_back(e.l / n_step);
                int mask = 0;
                for (int n = 0; n < 4; n++) {
                    q.push_back(half * a[n]);
                    if (a[n] != 0 || b[n] != 0) mask |= 1 << n;
                }
                for (int n = 0; n < 4; n++) q.push_back(half * b[n]);
                h.n = n_step;
                h.flags |= mask << F_ORDER_SHIFT;
                if (a[0] == 0 && a[1] == 0 && a[2] == 0 && a[3] == 0) h.flags |= F_MULT_NORMAL;
            }
            break;
        }
        case BMX_QUADRUPOLE: {
            int ix = ele_to_ab(e, a, b);
            if (ix == -2) { err = where(i, e) + "multipole scaling"; return BMX_ERR_UNSUPPORTED; }
            // pull_this_b1 (multipole_ele_to_ab.f90:664-676): k1 L comes out of the multipoles
            const double b1 = b[1];
            b[1] = 0;
            if (ix <= 1) ix = max_order(a, b, 2);
            double ae[BMX_N_POLE], be[BMX_N_POLE];
            const int ixe = ele_to_ab_elec(e, ae, be);
            const int ft = e.fringe_type == 0 ? BMX_FRINGE_NONE : e.fringe_type;
            if (ft != BMX_FRINGE_NONE && ft != BMX_FRINGE_SOFT_EDGE_ONLY &&
                ft != BMX_FRINGE_HARD_EDGE_ONLY && ft != BMX_FRINGE_FULL) {
                err = where(i, e) + "quadrupole fringe_type";
                return BMX_ERR_UNSUPPORTED;
            }
            if (ix > 3 || ixe > 3) { err = where(i, e) + "quadrupole multipole orders above 3"; return BMX_ERR_UNSUPPORTED; }
            // edge fields (init_fringe_info.f90:44-54), at the ends fringe_at names
            const bool hard = ft == BMX_FRINGE_HARD_EDGE_ONLY || ft == BMX_FRINGE_FULL;
            const bool soft = (ft == BMX_FRINGE_SOFT_EDGE_ONLY || ft == BMX_FRINGE_FULL) &&
                              (e.fq1 != 0 || e.fq2 != 0) && e.k1 != 0;
            const bool has_fringe = (ft != BMX_FRINGE_NONE && e.fringe_at != BMX_NO_END) || ixe > -1;
            const bool edge_ent = has_fringe && (e.fringe_at == BMX_BOTH_ENDS || e.fringe_at == BMX_ENTRANCE_END);
            const bool edge_exit = has_fringe && (e.fringe_at == BMX_BOTH_ENDS || e.fringe_at == BMX_EXIT_END);
            const bool edges = (edge_ent || edge_exit) && (hard || soft || ixe > -1);
            if (ix == -1 && ixe == -1 && !edges) {
                if (e.l == 0) {
                    h.kind = K_MARKER;
                    break;
                }
                h.kind = K_QUAD;
                int nterms = 0;
                const double t_lim = quad_series_limit(b1 * e.l, nterms);
                q = {e.l, 1.0 * b1, b1 * e.l, b1 / e.l, t_lim};
                h.n = nterms;
                break;
            }
            // multipoles or kicks beside k1, electric multipoles, or edge fields: stepped
            // (track_a_quadrupole.f90:46-53,70-130; one step when only the edges are present)
            if (e.l == 0) { err = where(i, e) + "zero-length quadrupole with multipoles or edges"; return BMX_ERR_UNSUPPORTED; }
            h.kind = K_QMULT;
            h.flags |= F_RARE_MAP;
            out.rare_q = true;
            long ns = (ix > -1 || ixe > -1) ? f_nint(std::fabs(e.l) / e.ds_step) : 1;
            const int n_step = ns > 1 ? (int)ns : 1;
            const double r_step = 1.0 / n_step;
            const double step_len = e.l * r_step;
            int nterms = 0;
            const double t_lim = quad_series_limit(b1 * step_len * r_step, nterms);
            q = {step_len, b1 * r_step, b1 * step_len * r_step, b1 / e.l, t_lim};
            int mmask = 0, emask = 0;
            const double half = r_step / 2;
            for (int n = 0; n < 4; n++) q.push_back(ix >= 0 ? half * a[n] : 0.0);
            for (int n = 0; n < 4; n++) q.push_back(ix >= 0 ? half * b[n] : 0.0);
            if (ix >= 0)
                for (int n = 0; n < 4; n++)
                    if (a[n] != 0 || b[n] != 0) mmask |= 1 << n;
            // electric: ab_multipole_kick's f = charge/(beta p0c) without the particle's beta,
            // times the scale step_len/2 (the element p0c equals the tracked reference p0c)
            const double fe = charge / e.p0c * (step_len / 2);
            for (int n = 0; n < 4; n++) q.push_back(ixe >= 0 ? ae[n] * fe : 0.0);
            for (int n = 0; n < 4; n++) q.push_back(ixe >= 0 ? -be[n] * fe : 0.0);
            for (int n = 0; n < 4; n++) q.push_back(ixe >= 0 ? charge * ae[n] / (n + 1) : 0.0);
            for (int n = 0; n < 4; n++) q.push_back(ixe >= 0 ? charge * be[n] / (n + 1) : 0.0);
            if (ixe >= 0)
                for (int n = 0; n < 4; n++)
                    if (ae[n] != 0 || be[n] != 0) emask |= 1 << n;
            // hard_multipole_edge_kick's cab = k1/(4 (n+2) rel_p) for n = 1; the soft edge's
            // f1 = k1 fq1/rel_p, f2 = k1 fq2/rel_p (fringe_mod.f90:600-620,760-790)
            q.push_back(e.k1 / 12.0);
            q.push_back(e.k1 * e.fq1);
            q.push_back(e.k1 * e.fq2);
            h.n = n_step;
            h.flags |= (mmask << F_ORDER_SHIFT) | (emask << F_ELEC_SHIFT);
            if (edge_ent) h.flags |= F_QEDGE_ENT;
            if (edge_exit) h.flags |= F_QEDGE_EXIT;
            if (hard) h.flags |= F_QHARD;
            if (soft) h.flags |= F_QSOFT;
            break;
        }
        case BMX_SBEND: {
            int ix = ele_to_ab(e, a, b);
            if (ix == -2) { err = where(i, e) + "multipole scaling"; return BMX_ERR_UNSUPPORTED; }
            double ae[BMX_N_POLE], be[BMX_N_POLE];
            const int ixe = ele_to_ab_elec(e, ae, be);
            const int ft = e.fringe_type == 0 ? BMX_FRINGE_NONE : e.fringe_type;
            const bool kick_type = ft == BMX_FRINGE_BASIC_BEND || ft == BMX_FRINGE_HARD_EDGE_ONLY;
            if (!kick_type && ft != BMX_FRINGE_NONE && ft != BMX_FRINGE_FULL) { err = where(i, e) + "bend fringe_type"; return BMX_ERR_UNSUPPORTED; }
            const bool plain = ix == -1 && ixe == -1 && ft != BMX_FRINGE_FULL && !misaligned(e);
            const bool at_ent = e.fringe_at == BMX_ENTRANCE_END || e.fringe_at == BMX_BOTH_ENDS;
            const bool at_exit = e.fringe_at == BMX_EXIT_END || e.fringe_at == BMX_BOTH_ENDS;
            double fg_ent = e.fint * e.hgap, fg_exit = e.fintx * e.hgapx;
            if (ft == BMX_FRINGE_HARD_EDGE_ONLY) fg_ent = fg_exit = 0;
            if (ft != BMX_FRINGE_NONE && at_ent) h.flags |= F_EDGE_ENT;
            if (ft != BMX_FRINGE_NONE && at_exit) h.flags |= F_EDGE_EXIT;
            if (e.e1 == 0 && e.e2 == 0 && fg_ent == 0 && fg_exit == 0 && e.k1 == 0) h.flags |= F_EDGE_SIMPLE;
            if (plain) {
                h.kind = K_BEND;
                const double g = e.g, L = e.l;
                if (g == 0 || L == 0) h.flags |= F_BEND_DRIFT;
                q.resize(kParams[K_BEND]);
                bend_body_block(e, L, g, g, q.data());
                if (g != 0 && q[2] == g && std::fabs(g * L) <= 0.1) h.flags |= F_BEND_SERIES;
                break;
            }
            // the general bend (track_a_bend.f90:18-333)
            const bool exact = e.exact_multipoles > 1;
            if (exact && e.g != 0 && (ix > -1 || ixe > -1)) {
                err = where(i, e) + "exact multipoles in a curved bend";
                return BMX_ERR_UNSUPPORTED;
            }
            if (e.l == 0) { err = where(i, e) + "zero-length bend with multipoles, offsets or edges"; return BMX_ERR_UNSUPPORTED; }
            // b1 pulled out for the k1 map unless exact_multipoles is set (:55-70)
            double b1 = 0;
            if (!exact) {
                b1 = b[1];
                b[1] = 0;
                if (ix <= 1) ix = max_order(a, b, 2);
                if (std::fabs(b1) < 1e-10) {
                    b[1] = b1;
                    b1 = 0;
                }
            }
            const bool stepped = ix > -1 || ixe > -1;
            const double L = e.l, g = e.g;
            const double dg = b[0] / L;  // the b0 part is the field error (:74-80)
            b[0] = 0;
            if (max_order(a, b, BMX_N_POLE) > 3 || max_order(ae, be, BMX_N_POLE) > 3) {
                err = where(i, e) + "bend multipole orders above 3";
                return BMX_ERR_UNSUPPORTED;
            }
            long ns = stepped ? f_nint(std::fabs(L) / e.ds_step) : 1;
            const int n_step = ns > 1 ? (int)ns : 1;
            const double r_step = 1.0 / n_step;
            const double step_len = L / n_step;
            h.kind = K_BMULT;
            h.flags |= F_RARE_MAP;
            out.rare_q = true;
            h.n = n_step;
            q.resize(kParams[K_BMULT]);
            bend_body_block(e, step_len, g, g + dg, q.data());
            if (g != 0 && q[2] == g && std::fabs(g * step_len) <= 0.1) h.flags |= F_BEND_SERIES;
            if (b1 != 0) h.flags |= F_BK1;
            else if (g == 0 && dg == 0) h.flags |= F_BEND_DRIFT;
            if (dg == 0) h.flags |= F_BSEMI;
            q[29] = e.g + e.dg;
            q[30] = b1 / L;
            int mmask = 0, emask = 0;
            const double fe = charge / e.p0c * (step_len / 2);
            for (int n = 0; n < 4; n++) {
                if (ix > -1 && (a[n] != 0 || b[n] != 0)) mmask |= 1 << n;
                if (ixe > -1 && (ae[n] != 0 || be[n] != 0)) emask |= 1 << n;
                q[31 + n] = ix > -1 ? a[n] * (r_step / 2) : 0.0;
                q[35 + n] = ix > -1 ? b[n] * (r_step / 2) : 0.0;
                q[39 + n] = ixe > -1 ? ae[n] * fe : 0.0;
                q[43 + n] = ixe > -1 ? -be[n] * fe : 0.0;
                q[47 + n] = ixe > -1 ? charge * ae[n] / (n + 1) : 0.0;
                q[51 + n] = ixe > -1 ? charge * be[n] / (n + 1) : 0.0;
            }
            h.flags |= (mmask << F_ORDER_SHIFT) | (emask << F_ELEC_SHIFT);
            if (at_ent) h.flags |= F_QEDGE_ENT;
            if (at_exit) h.flags |= F_QEDGE_EXIT;
            if (ft == BMX_FRINGE_FULL) h.flags |= F_BEXACT;
            q[55] = e.k1 / 12.0;
            q[56] = std::cos(e.e1);
            q[57] = std::sin(e.e1);
            q[58] = std::sin(2.0 * e.e1);
            q[59] = e.fint * e.hgap;
            q[60] = std::cos(e.e2);
            q[61] = std::sin(e.e2);
            q[62] = std::sin(2.0 * e.e2);
            q[63] = e.fintx * e.hgapx;
            const double beta0 = e.p0c / e.e_tot;
            q[64] = 1 / (beta0 * beta0);
            q[65] = e.e1;
            q[66] = e.e2;
            if (e.x_offset != 0 || e.y_offset != 0 || e.z_offset != 0 || e.x_pitch != 0 || e.y_pitch != 0) {
                h.flags |= F_MISALIGN;
                bend_frame_blocks(e, &q[67]);
            } else if (e.ref_tilt != 0) {
                h.flags |= F_TILT;
            }
            q[91] = std::cos(e.ref_tilt);
            q[92] = std::sin(e.ref_tilt);
            q[93] = L;
            break;
        }
        case BMX_RFCAVITY: {
            int ix = ele_to_ab(e, a, b);
            if (ix != -1) { err = where(i, e) + "rf cavity multipoles/kicks are not supported"; return BMX_ERR_UNSUPPORTED; }
            if (e.rf_frequency == 0) {
                if (e.voltage != 0) { err = where(i, e) + "rf cavity with zero frequency"; return BMX_ERR_UNSUPPORTED; }
            }
            h.kind = K_RF;
            const double length = e.l * 1;
            long ns = f_nint(std::fabs(length) / e.ds_step);
            const int n_slice = ns > 1 ? (int)ns : 1;
            h.n = n_slice;
            const double twopi = 2 * 3.141592653589793238462643383279;
            const double f = e.rf_frequency;
            q = {length / n_slice, e.voltage * 1.0, e.phi0, f, f != 0 ? 0.5 / f : 1e300, twopi * f,
                 twopi * e.phi0_multipass};
            break;
        }
        case BMX_MULTIPOLE:
        case BMX_AB_MULTIPOLE: {
            double k0, t0;
            thin_multipole_ab(e, a, b, k0, t0);
            h.kind = K_THIN;
            h.flags |= F_RARE_MAP;
            const bool ref0 = key == BMX_MULTIPOLE && k0 != 0;
            if (ref0) {
                h.flags |= F_THIN_REF0;
                a[0] = b[0] = 0;
            }
            const int n_ord = max_order(a, b, BMX_N_POLE) + 1;
            h.n = n_ord;
            q.push_back(ref0 ? k0 * std::cos(t0) : 0.0);
            q.push_back(ref0 ? k0 * std::sin(t0) : 0.0);
            for (int n = 0; n < n_ord; n++) q.push_back(a[n]);
            for (int n = 0; n < n_ord; n++) q.push_back(b[n]);
            break;
        }
        default:
            err = where(i, e) + "element kind is not implemented on the GPU path";
            return BMX_ERR_UNSUPPORTED;
        }
        double ap[4] = {0, 0, 0, 0}, mom[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (!aperture_setup(e, h.flags, ap, mom, why)) {
            err = where(i, e) + why + " are not supported";
            return BMX_ERR_UNSUPPORTED;
        }
        const int n_prm = h.kind == K_THIN ? kParams[K_THIN] + 2 * h.n : kParams[h.kind];
        if ((int)q.size() != n_prm) { err = where(i, e) + "internal parameter count"; return BMX_ERR_ARG; }
        h.off = (int32_t)out.prm.size();
        if (offset) {
            double ob[kOffsetLen];
            offset_block(e_frame, h.flags, ob);
            out.prm.insert(out.prm.end(), ob, ob + kOffsetLen);
        }
        if (h.flags & (F_APE_ENT | F_APE_EXIT)) out.prm.insert(out.prm.end(), ap, ap + 4);
        if (h.flags & F_APE_MOMENTUM) out.prm.insert(out.prm.end(), mom, mom + 8);
        out.prm.insert(out.prm.end(), q.begin(), q.end());
        if (rad_here) {
            h.flags |= F_RAD | F_RARE_MAP;
            out.rare_q = true;  // the kernels that carry the kicks
            out.rad = true;
            double rb[kRadLen];
            rad_edge_block(rad[2 * i], rb);
            rad_edge_block(rad[2 * i + 1], rb + kRadEdge);
            out.prm.insert(out.prm.end(), rb, rb + kRadLen);
        }
        out.hdr.push_back(h);
    }
    // Entrance tests equal to the exit test just before (a particle does not move between
    // elements, so it passes the second if it passed the first). Element 0 keeps its test.
    for (size_t i = 1; i < out.hdr.size(); i++) {
        EleHdr &h = out.hdr[i];
        const EleHdr &g = out.hdr[i - 1];
        if (!(h.flags & F_APE_ENT) || !(g.flags & F_APE_EXIT)) continue;
        if ((h.flags | g.flags) & F_APE_MOMENTUM) continue;
        if ((h.flags & F_APE_ELLIP) != (g.flags & F_APE_ELLIP)) continue;
        const int32_t ha = h.off + ((h.flags & F_OFFSET) ? kOffsetLen : 0);
        const int32_t ga = g.off + ((g.flags & F_OFFSET) ? kOffsetLen : 0);
        if (std::memcmp(&out.prm[ha], &out.prm[ga], 4 * sizeof(double)) != 0) continue;
        h.flags = (h.flags & ~F_APE_ENT) | F_APE_ENT_SAME;
    }
    return BMX_OK;
}

}  // namespace bmx
