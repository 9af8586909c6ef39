// Per-particle element maps for gfx950, FP64.
//
// One lane tracks one particle; the particle lives in VGPRs for the whole call. Element
// parameters are wave-uniform and are read through scalar loads (SGPR operands), so every
// branch on the element kind is uniform across the wavefront. The maps restate Bmad's
// bmad_standard tracking (file:line citations below); operation order may differ from the
// reference at the rounding level (FMA contraction, reciprocal reuse), which the parity
// tests bound at 1e-12 relative per element.
#pragma once
#include <hip/hip_runtime.h>

#include "bmx_table.hpp"

namespace bmx {

constexpr double kPi = 3.141592653589793238462643383279;
constexpr double kTwoPi = 2 * kPi;
constexpr double kCLight = 2.99792458e8;
constexpr double kMaxAperture = 1e3;  // bmad_com%max_aperture_limit, bmad_struct.f90:2314

enum : int32_t {
    ST_ALIVE = 1, ST_LOST = 2, ST_LOST_NEG_X = 3, ST_LOST_POS_X = 4,
    ST_LOST_NEG_Y = 5, ST_LOST_POS_Y = 6, ST_LOST_PZ = 8,
};

struct Particle {
    double x, px, y, py, z, pz, beta;
    int32_t state;
    // Invariants of the energy, refreshed whenever pz changes (only in rf cavities here):
    double rel_p;  // 1 + pz
    double ez;     // sqrt_one(mc2^2 (2 pz + pz^2) / ((p0c rel_p)^2 + mc2^2)), track_a_drift.f90:69
};

__device__ __forceinline__ double sqr(double v) { return v * v; }

// sim_utils/special_functions/sqrt_one.f90:18-41
__device__ __forceinline__ double sqrt_one(double v) { return v / (sqrt(1 + v) + 1); }

// sim_utils/special_functions/sinc.f90:15-40
__device__ __forceinline__ double sinc(double v) { return fabs(v) < 1e-8 ? 1.0 : sin(v) / v; }

__device__ __forceinline__ void refresh_energy(Particle &p, const Species &sp)
{
    p.rel_p = 1 + p.pz;
    p.ez = sqrt_one((sp.mc2_sq * (2 * p.pz + p.pz * p.pz)) / (sqr(sp.p0c * p.rel_p) + sp.mc2_sq));
}

// bmad/low_level/orbit_too_large.f90:26-120 (charged particle, momentum check on).
__device__ __forceinline__ bool orbit_too_large(Particle &p)
{
    const double f_unstable = p.px * p.px + p.py * p.py - p.rel_p * p.rel_p;
    const bool bad = (p.x > kMaxAperture) | (-p.x > kMaxAperture) | (p.y > kMaxAperture) |
                     (-p.y > kMaxAperture) | (p.rel_p < 0) | (f_unstable > 0);
    if (!bad) return false;
    if (p.x > kMaxAperture) p.state = ST_LOST_POS_X;
    else if (-p.x > kMaxAperture) p.state = ST_LOST_NEG_X;
    else if (p.y > kMaxAperture) p.state = ST_LOST_POS_Y;
    else if (-p.y > kMaxAperture) p.state = ST_LOST_NEG_Y;
    else if (p.rel_p < 0) p.state = ST_LOST_PZ;
    else if (fabs(p.px) > fabs(p.py)) p.state = p.px > 0 ? ST_LOST_POS_X : ST_LOST_NEG_X;
    else p.state = p.py > 0 ? ST_LOST_POS_Y : ST_LOST_NEG_Y;
    return true;
}

// Aperture test of check_aperture_limit.f90:256-324 (elliptical with strict r < 1, or
// rectangular). ap[] layout in bmx_table.hpp.
__device__ __forceinline__ void aperture(Particle &p, int32_t flags, const double *__restrict__ ap)
{
    if (flags & F_APE_ELLIP) {
        const double xp = p.x - ap[0], yp = p.y - ap[2];
        const double xr = xp * ap[1], yr = yp * ap[3];
        const double r = xr * xr + yr * yr;
        if (r < 1) return;
        if (fabs(xr) > fabs(yr)) p.state = xp > 0 ? ST_LOST_POS_X : ST_LOST_NEG_X;
        else p.state = yp > 0 ? ST_LOST_POS_Y : ST_LOST_NEG_Y;
    } else {
        const double x1 = ap[0], x2 = ap[1], y1 = ap[2], y2 = ap[3];
        const bool out = (x1 != 0 && p.x < x1) | (x2 != 0 && p.x > x2) | (y1 != 0 && p.y < y1) |
                         (y2 != 0 && p.y > y2);
        if (!out) return;
        double u = 0, f;
        if (x1 != 0 && p.x < x1) { f = fabs((p.x - x1) / (x2 - x1)); if (f > u) { p.state = ST_LOST_NEG_X; u = f; } }
        if (x2 != 0 && p.x > x2) { f = fabs((p.x - x2) / (x2 - x1)); if (f > u) { p.state = ST_LOST_POS_X; u = f; } }
        if (y1 != 0 && p.y < y1) { f = fabs((p.y - y1) / (y2 - y1)); if (f > u) { p.state = ST_LOST_NEG_Y; u = f; } }
        if (y2 != 0 && p.y > y2) { f = fabs((p.y - y2) / (y2 - y1)); if (f > u) { p.state = ST_LOST_POS_Y; u = f; } }
    }
}

// bmad/low_level/track_a_drift.f90:25-116 (direction = orientation = +1, beta > 0).
// sqrt_one(-pxy2) reuses ps_rel = sqrt(1 - pxy2), the same rounded square root.
__device__ __forceinline__ void drift(Particle &p, double L)
{
    if (L == 0) return;
    const double px_rel = p.px / p.rel_p;
    const double py_rel = p.py / p.rel_p;
    const double pxy2 = px_rel * px_rel + py_rel * py_rel;
    if (pxy2 >= 1) {
        p.state = ST_LOST_PZ;
        return;
    }
    const double ps = sqrt(1 - pxy2);
    p.x = p.x + L * px_rel / ps;
    p.y = p.y + L * py_rel / ps;
    p.z = p.z + L * (p.ez + (-pxy2 / (ps + 1)) / ps);
}

// bmad/low_level/quad_mat2_calc.f90:22-81: 2x2 map and z coefficients of one plane with
// strength k (k > 0 defocusing). cx/sx are precomputed by the caller for |k|.
struct Plane {
    double m11, m12, m21, m22, z1, z2, z3;
};

__device__ __forceinline__ Plane quad_plane(double k, double L, double rel_p, double cx, double sx)
{
    Plane q;
    q.m11 = cx;
    q.m12 = sx / rel_p;
    q.m21 = k * sx * rel_p;
    q.m22 = cx;
    q.z1 = k * (-cx * sx + L) / 4;
    q.z2 = -k * sx * sx / (2 * rel_p);
    q.z3 = -(cx * sx + L) / (4 * rel_p * rel_p);
    return q;
}

// bmad/low_level/track_a_quadrupole.f90:17-148 for n_step = 1 (no multipoles other than k1),
// no fringe, no misalignment; low_energy_z_correction.f90:22-58.
__device__ __forceinline__ void quadrupole(Particle &p, const double *__restrict__ q)
{
    const double L = q[0], b1 = q[1];
    const double k1 = b1 / (L * p.rel_p);
    const double sqrt_k = sqrt(fabs(k1));
    const double sk_l = sqrt_k * L;
    double cx_x, sx_x, cx_y, sx_y;  // planes: x uses k = -k1, y uses k = +k1
    if (fabs(sk_l) < 1e-10) {
        const double kl2x = -k1 * L * L, kl2y = k1 * L * L;
        cx_x = 1 + kl2x / 2;
        sx_x = (1 + kl2x / 6) * L;
        cx_y = 1 + kl2y / 2;
        sx_y = (1 + kl2y / 6) * L;
    } else {
        double s, c;
        sincos(sk_l, &s, &c);
        const double ch = cosh(sk_l), sh = sinh(sk_l);
        const double sf = s / sqrt_k, sd = sh / sqrt_k;
        if (k1 > 0) {  // x: k = -k1 < 0 focusing; y: defocusing
            cx_x = c; sx_x = sf; cx_y = ch; sx_y = sd;
        } else {
            cx_x = ch; sx_x = sd; cx_y = c; sx_y = sf;
        }
    }
    const Plane a = quad_plane(-k1, L, p.rel_p, cx_x, sx_x);
    const Plane b = quad_plane(k1, L, p.rel_p, cx_y, sx_y);
    const double x = p.x, px = p.px, y = p.y, py = p.py;
    p.z = p.z + (a.z1 * x * x + a.z2 * x * px + a.z3 * px * px + b.z1 * y * y + b.z2 * y * py +
                 b.z3 * py * py);
    p.x = a.m11 * x + a.m12 * px;
    p.px = a.m21 * x + a.m22 * px;
    p.y = b.m11 * y + b.m12 * py;
    p.py = b.m21 * y + b.m22 * py;
    // low_energy_z_correction
    const double beta0 = q[2], me2 = q[3], f = q[4], thr = q[5];
    const double pz = p.pz;
    if (q[6] * sqr(beta0 * pz) < thr)
        p.z = p.z + L * pz * (1 - 3 * (pz * beta0 * beta0) / 2 + pz * pz * f) * me2;
    else
        p.z = p.z + L * (p.beta - beta0) / beta0;
}

// bmad/modules/multipole_mod.f90:82-174,314-401 (ab_multipole_kicks, magnetic): kick from
// orders n in mask with strengths (a_n, b_n) * scale; c_multi's signed binomials summed as the
// real and imaginary parts of (x + i y)^n: kx = -b Re + a Im, ky = a Re + b Im.
__device__ __forceinline__ void multipole_kick(Particle &p, const double *__restrict__ ah,
                                               const double *__restrict__ bh, int32_t mask,
                                               double factor)
{
    double kx = 0, ky = 0;
    double zr = 1, zi = 0;
    const double x = p.x, y = p.y;
#pragma unroll
    for (int n = 0; n < 4; n++) {
        if (mask & (1 << n)) {
            const double an = ah[n] * factor, bn = bh[n] * factor;
            kx += -bn * zr + an * zi;
            ky += an * zr + bn * zi;
        }
        if (n < 3 && (mask >> (n + 1))) {
            const double nr = zr * x - zi * y;
            zi = zr * y + zi * x;
            zr = nr;
        }
    }
    p.px = p.px + kx;
    p.py = p.py + ky;
}

// bmad/low_level/track_a_thick_multipole.f90:18-133 (magnetic, no misalignment):
// half kick, then n_step x (drift, kick), the last kick halved. As in the reference, a drift
// that loses the particle does not end the loop (track_a_drift does not test the state).
__device__ __forceinline__ void thick_multipole(Particle &p, const double *__restrict__ m,
                                                int32_t n_step, int32_t mask)
{
    const double step = m[0];
    multipole_kick(p, m + 1, m + 5, mask, 1.0);
    for (int i = 1; i <= n_step; i++) {
        drift(p, step);
        multipole_kick(p, m + 1, m + 5, mask, i == n_step ? 1.0 : 2.0);
    }
}

// bmad/modules/fringe_mod.f90:241-419 hwang_bend_edge_kick (time_dir = +1), orbit only.
// e[] = gt, gt2, gs2, k1_tane, fg_factor of that face; g_tot of the bend.
__device__ __forceinline__ void hwang_edge(Particle &p, const double *__restrict__ e, double g_tot,
                                           bool entering, bool simple)
{
    const double ef = 1 / p.rel_p;
    const double x = p.x, px = p.px, y = p.y, py = p.py;
    if (simple) {
        // e = 0, fint*hgap = 0, k1 = 0: gt = gt2 = k1_tane = fg = 0, gs2 = g_tot.
        const double gs2 = e[2];
        if (entering) {
            p.x = x + (gs2 * y * y) * ef / 2;
            p.py = py + (-(g_tot)*px * y) * ef;
            p.z = p.z + ef * ef * 0.5 * (-(px * y * y * gs2));
        } else {
            p.x = x + (-gs2 * y * y) * ef / 2;
            p.py = py + ((g_tot)*px * y) * ef;
            p.z = p.z + ef * ef * 0.5 * (px * y * y * gs2);
        }
        return;
    }
    const double gt = e[0], gt2 = e[1], gs2 = e[2], k1_tane = e[3], fg = e[4], tan_e2 = e[5];
    double dx, dpx, dy, dpy, dz;
    if (entering) {
        dx = (-gt2 * x * x + gs2 * y * y) * ef / 2;
        dpx = (gt * g_tot * (1 + 2 * tan_e2) * y * y / 2 + gt2 * (x * px - y * py) +
               k1_tane * (x * x - y * y)) * ef;
        dy = gt2 * x * y * ef;
        dpy = (fg * y - gt2 * x * py - (g_tot + gt2) * px * y - 2 * k1_tane * x * y) * ef;
        dz = ef * ef * 0.5 *
             (y * y * fg + x * x * x * (4.0 * k1_tane - gt * gt2) / 6.0 +
              0.5 * x * y * y * (-4.0 * k1_tane + gt * gs2) + (x * x * px - 2.0 * x * y * py) * gt2 -
              px * y * y * gs2);
    } else {
        dx = (gt2 * x * x - gs2 * y * y) * ef / 2;
        dpx = (gt2 * (y * py - x * px) + k1_tane * (x * x - y * y) - gt * gt2 * (x * x + y * y) / 2) * ef;
        dy = -gt2 * x * y * ef;
        dpy = (fg * y + gt2 * x * py + (g_tot + gt2) * px * y + (gt * gs2 - 2 * k1_tane) * x * y) * ef;
        dz = ef * ef * 0.5 *
             (y * y * fg + x * x * x * (4.0 * k1_tane - gt * gt2) / 6.0 +
              0.5 * x * y * y * (-4.0 * k1_tane + gt * gs2) - (x * x * px - 2.0 * x * y * py) * gt2 +
              px * y * y * gs2);
    }
    p.x = x + dx;
    p.px = px + (dpx + gt * x);
    p.y = y + dy;
    p.py = py + (dpy - gt * y);
    p.z = p.z + dz;
}

// bmad/low_level/track_a_bend.f90:18-333: exact_multipoles = off, no body multipoles, b1 = 0,
// n_step = 1, no misalignment. Exact sector-bend body (:163-209), semi-linear body near the
// zero orbit (:136-159), drift when g = dg = 0 (:117-118).
__device__ __forceinline__ void sbend(Particle &p, int32_t flags, const double *__restrict__ b,
                                      const Species &sp)
{
    const double L = b[0], g = b[1], g_tot = b[2], angle = b[3];
    const bool simple = flags & F_EDGE_SIMPLE;
    if (flags & F_EDGE_ENT) hwang_edge(p, b + 10, g_tot, true, simple);
    if (orbit_too_large(p)) return;
    if (flags & F_BEND_DRIFT) {
        drift(p, L);
    } else {
        const double x = p.x, px = p.px, y = p.y, py = p.py, z = p.z;
        const double rel_p = p.rel_p;
        const double cos_a = b[4], sin_a = b[5], sinc_a = b[6], cosc_a = b[7];
        if (fabs(x * g) < 1e-9 && fabs(px) < 1e-9 && fabs(py) < 1e-9 && fabs(p.pz) < 1e-9) {
            const double ll = L;
            const double gam2 = sp.mc2_sq / (sqr(rel_p * sp.p0c) + sp.mc2_sq);
            const double m56 = ll * (gam2 - sqr(g * ll) * b[9]);
            p.x = cos_a * x + ll * sinc_a * px + g * ll * ll * cosc_a * p.pz;
            p.px = -g * sin_a * x + cos_a * px + g * ll * sinc_a * p.pz;
            p.y = y + ll * py;
            p.z = -g * ll * sinc_a * x - g * ll * ll * cosc_a * px + z + m56 * p.pz;
        } else {
            const double beta_ref = b[8];
            const double pt = sqrt(rel_p * rel_p - py * py);
            if (fabs(px) > pt) {
                p.state = ST_LOST_PZ;
                return;
            }
            const double g_p = g_tot / pt;
            const double phi_1 = asin(px / pt);
            double sin_plus, cos_plus;
            sincos(angle + phi_1, &sin_plus, &cos_plus);
            const double u = (1 + g * x) * L * sinc_a;
            const double alpha = 2 * (1 + g * x) * sin_plus * L * sinc_a - g_p * u * u;
            const double r = cos_plus * cos_plus + g_p * alpha;
            if (r < 0 || (fabs(g_p) < 1e-5 && fabs(cos_plus) < 1e-5)) {
                p.state = ST_LOST;
                p.x = 2 * kMaxAperture;
                return;
            }
            const double rad = sqrt(r);
            double xi;
            if (cos_plus > 0) xi = alpha / (rad + cos_plus);
            else xi = (rad - cos_plus) / g_p;
            p.x = x * cos_a - L * L * g * cosc_a + xi;
            if (fabs(p.x) > kMaxAperture) {
                p.state = ST_LOST;
                return;
            }
            const double L_u = xi;
            const double L_v = -(L * sinc_a + x * sin_a);
            const double L_c = sqrt(L_v * L_v + L_u * L_u);
            const double angle_p = 2 * (angle + phi_1 - atan2(L_u, -L_v));
            const double L_p = L_c / sinc(angle_p / 2);
            p.px = pt * sin(phi_1 + angle - angle_p);
            p.y = y + py * L_p / pt;
            p.z = z + p.beta * L / beta_ref - rel_p * L_p / pt;
        }
    }
    if (orbit_too_large(p)) return;
    if (flags & F_EDGE_EXIT) hwang_edge(p, b + 16, g_tot, false, simple);
}

// Fortran MODULO for reals followed by modulo2_mod.f90:85-99: result in [-amp, amp).
__device__ __forceinline__ double modulo2(double x, double amp)
{
    const double p2 = 2 * amp;
    double r;
    if (x >= 0 && x < p2) r = x;
    else if (x < 0 && x >= -p2) r = x + p2;
    else {
        r = fmod(x, p2);
        if (r != 0 && (r < 0)) r += p2;
    }
    if (r >= amp) r = r - p2;
    return r;
}

// bmad/low_level/apply_energy_kick.f90:23-75 (ddE_dr = 0). Returns false if lost.
__device__ __forceinline__ bool energy_kick(Particle &p, double dE, const Species &sp)
{
    const double pc = p.rel_p * sp.p0c;
    const double beta_old = p.beta;
    const double E_old = pc / p.beta;
    const double E_new = E_old + dE;
    if (E_new < sp.mc2) {
        p.pz = -1;
        p.beta = 0;
        p.state = ST_LOST_PZ;
        p.rel_p = 0;
        return false;
    }
    p.pz = p.pz + p.rel_p * sqrt_one((2 * E_old * dE + dE * dE) / (pc * pc));
    const double pc_new = sp.p0c * (1 + p.pz);
    const double beta_new = pc_new / E_new;
    p.z = p.z * beta_new / beta_old;
    p.beta = beta_new;
    return true;
}

// bmad/low_level/track_a_rfcavity.f90:17-146 with relative time tracking: phase from
// particle_rf_time.f90:86-90,116-118 (t = -z/(beta c), wrapped to [-1/2f, 1/2f)), computed in
// double-double where the reference uses quad precision, then modulo2(phase0, pi); n_slice + 1
// energy kicks (halved at the ends) between n_slice drifts (:96-133).
__device__ __forceinline__ void rfcavity(Particle &p, const double *__restrict__ r, int32_t n_slice,
                                         const Species &sp)
{
    const double dl = r[0], voltage = r[1], phi0 = r[2], freq = r[3], half_period = r[4];
    const double twopi_f = r[5];
    double t = -p.z / (p.beta * kCLight);
    if (!(fabs(t) < half_period)) t = modulo2(t, half_period);
    // phase0 = twopi * (phi0 - t * freq) with the product and difference carried exactly
    const double tf_hi = t * freq;
    const double tf_lo = fma(t, freq, -tf_hi);
    const double d_hi = phi0 - tf_hi;
    const double d_lo = ((phi0 - d_hi) - tf_hi) - tf_lo;
    const double ph_hi = kTwoPi * d_hi;
    const double ph_lo = fma(kTwoPi, d_hi, -ph_hi) + kTwoPi * d_lo;
    double phase0 = ph_hi + ph_lo;
    phase0 = phase0 + r[6];
    double phase = modulo2(phase0, kPi);
    for (int i = 0; i <= n_slice; i++) {
        double dE = voltage * sin(phase) / n_slice;
        if (i == 0 || i == n_slice) dE = dE / 2;
        if (!energy_kick(p, dE, sp)) return;
        refresh_energy(p, sp);
        if (p.pz == -1) {
            p.state = ST_LOST_PZ;
            return;
        }
        if (i != n_slice) {
            const double z0 = p.z;
            drift(p, dl);
            phase = phase + twopi_f * (p.z - z0) / (kCLight * p.beta);
        }
    }
}

// One element for one particle: the body of track1 (bmad/code/track1.f90:196-329) after the
// entrance validity check, i.e. entrance aperture, element map, orbit_too_large (:291),
// exit aperture (:329).
__device__ __forceinline__ void track_element(Particle &p, const EleHdr h,
                                              const double *__restrict__ prm, const Species &sp)
{
    const double *ap = prm + kParams[h.kind];
    if (h.flags & F_APE_ENT) {
        aperture(p, h.flags, ap);
        if (p.state != ST_ALIVE) return;
    }
    switch (h.kind) {
    case K_DRIFT: drift(p, prm[0]); break;
    case K_QUAD: quadrupole(p, prm); break;
    case K_MULT: thick_multipole(p, prm, h.n, (h.flags >> F_ORDER_SHIFT) & 0xF); break;
    case K_BEND: sbend(p, h.flags, prm, sp); break;
    case K_RF: rfcavity(p, prm, h.n, sp); break;
    default: break;
    }
    orbit_too_large(p);
    if (p.state == ST_ALIVE && (h.flags & F_APE_EXIT)) aperture(p, h.flags, ap);
}

}  // namespace bmx
