// Per-particle element maps for gfx950, FP64.
//
// One lane tracks one particle; the particle lives in VGPRs for the whole call. Element
// parameters are wave-uniform and are read through scalar loads (SGPR operands), so every
// branch on the element kind is uniform across the wavefront.
//
// Two formulations of the drift, quadrupole and bend bodies, selected by the template flag F:
//   F = false ("reference form"): Bmad's formulas as written (sqrt, divisions, sin/cos/asin/
//             atan2/sinh/cosh of the particle's coordinates per element), file:line cited.
//   F = true  ("fast form", the default): the same maps re-expressed without per-particle
//             transcendentals and with most divisions replaced by reciprocals cached per
//             particle -- convergent series with bounded truncation (< 1e-17 relative) where
//             the arguments are small, angle-addition identities in the bend, rsqrt-based
//             square roots. Outside the series domains each map falls back to the reference
//             form for that particle.
// Both agree with the CPU restatement to the rounding level (<= 1e-12 relative per element,
// checked by the GPU parity tests); element-loss decisions are identical away from the
// aperture edges.
#pragma once
#include <hip/hip_runtime.h>

#include "bmx_table.hpp"

namespace bmx {

// Element parameters are read through the constant address space: wave-uniform indices into
// it compile to scalar loads (s_load) whose results sit in SGPRs.
#define BMX_CONST __attribute__((address_space(4)))
typedef const double BMX_CONST *CP;
typedef const int32_t BMX_CONST *CHdr;

// One element's header values (bmx_table.hpp EleHdr). A specialised build passes them as
// compile-time constants, so every test on the kind and flags folds away.
struct EleRef {
    int32_t kind, flags, n, off;
};
typedef int32_t EleVec __attribute__((ext_vector_type(4)));
// The header of element h, read with one 16-byte scalar load.
__device__ __forceinline__ EleRef load_ele(CHdr h)
{
    const EleVec v = *(const EleVec BMX_CONST *)h;
    // F_APE_CENTERED only pays where the flags are constants; cleared here, its test folds away
    return {v.x, v.y & ~F_APE_CENTERED, v.z, v.w};
}

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
    // Functions of the energy, refreshed whenever pz or beta changes (only in rf cavities):
    double rel_p;   // 1 + pz
    double inv_rp;  // 1 / (1 + pz)
    double ez;      // sqrt_one(mc2^2 (2 pz + pz^2) / ((p0c rel_p)^2 + mc2^2)), track_a_drift.f90:69
    double lez;     // low_energy_z_correction per metre (low_energy_z_correction.f90:44-57)
    // Where the particle was lost (set with the loss state): element index (1-based), turn.
    int32_t loss_ele, loss_turn;
    // The particle's index in its beam: the counter of its radiation fluctuation stream.
    uint32_t id;
};

__device__ __forceinline__ double sqr(double v) { return v * v; }

// A constant table read through a pointer the compiler cannot follow: the loads of its entries
// stay where they are used instead of being hoisted out of the loops into registers held for
// the whole kernel. For coefficients of rarely taken paths (rf phase, far arguments).
__device__ __forceinline__ CP launder(CP p)
{
    asm volatile("" : "+s"(p));
    return p;
}
// A wave-uniform integer passed through an empty asm at its point of use, so that its copy into
// a vector register is made there (in a rare branch) and not hoisted into the common path.
__device__ __forceinline__ int32_t uniform_here(int32_t v)
{
    asm volatile("" : "+s"(v));
    return v;
}

// v, opaque to the optimiser. Inside a rarely taken branch it keeps the compiler from
// speculating a cheap-looking computation (a square root) onto every lane of the common path.
__device__ __forceinline__ double rare(double v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// True on every lane if c holds on any active lane. Rare per-particle cases (a coordinate
// outside a series domain, a loss) are gated by this wave-uniform test so the common path is
// straight-line code; inside the gated block each lane still takes its own branch, so a
// particle's result never depends on the other particles of its wave.
#ifndef BMX_GATE
#define BMX_GATE 1
#endif
#if BMX_GATE
__device__ __forceinline__ bool any_lane(bool c) { return __ballot(c) != 0; }

// The active lanes on which c holds. A compound rare condition is gated as the union of the
// masks of its plain comparisons (what each comparison instruction writes), and re-tested per
// lane only inside the gate, so the common path carries no per-lane form of it.
typedef unsigned long long LaneMask;
__device__ __forceinline__ LaneMask lanes(bool c) { return __ballot(c); }
#else
__device__ __forceinline__ bool any_lane(bool c) { return c; }
#endif

// sim_utils/special_functions/sqrt_one.f90:18-41
__device__ __forceinline__ double sqrt_one(double v) { return v / (sqrt(1 + v) + 1); }

// sim_utils/special_functions/sinc.f90:15-40
__device__ __forceinline__ double sinc(double v) { return fabs(v) < 1e-8 ? 1.0 : sin(v) / v; }

__device__ __forceinline__ void refresh_energy(Particle &p, const Species &sp)
{
    p.rel_p = 1 + p.pz;
    p.inv_rp = 1 / p.rel_p;
    p.ez = sqrt_one((sp.mc2_sq * (2 * p.pz + p.pz * p.pz)) / (sqr(sp.p0c * p.rel_p) + sp.mc2_sq));
    const double pz = p.pz;
    if (sp.mc2 * sqr(sp.beta0 * pz) < sp.lez_thr)
        p.lez = pz * (1 - 3 * (pz * sp.beta0 * sp.beta0) / 2 + pz * pz * sp.lez_f) * sp.me2;
    else
        p.lez = (p.beta - sp.beta0) / sp.beta0;
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

// The tests of orbit_too_large without the state assignment: true if it would flag the
// particle. |x| > limit is x > limit or -x > limit; px^2 + py^2 > rel_p^2 has the sign of
// f_unstable.
__device__ __forceinline__ bool orbit_bad(const Particle &p)
{
    const double f_unstable = p.px * p.px + p.py * p.py - p.rel_p * p.rel_p;
    return (fabs(p.x) > kMaxAperture) | (fabs(p.y) > kMaxAperture) | (p.rel_p < 0) | (f_unstable > 0);
}

// The lanes on which orbit_bad holds.
__device__ __forceinline__ LaneMask orbit_bad_lanes(const Particle &p)
{
    const double f_unstable = p.px * p.px + p.py * p.py - p.rel_p * p.rel_p;
    return lanes(fabs(p.x) > kMaxAperture) | lanes(fabs(p.y) > kMaxAperture) |
           lanes(p.rel_p < 0) | lanes(f_unstable > 0);
}

// True if the aperture test of check_aperture_limit.f90:256-324 loses the particle (no state
// change): elliptical r >= 1 (alive is strictly r < 1), rectangular outside a set limit.
// Offsets from the centre of an elliptical aperture (x - 0 is x exactly, so a centred ellipse
// skips the subtraction with identical results).
__device__ __forceinline__ void ellipse_offsets(const Particle &p, int32_t flags, CP ap, double &xp,
                                                double &yp)
{
    if (flags & F_APE_CENTERED) {
        xp = p.x;
        yp = p.y;
    } else {
        xp = p.x - ap[0];
        yp = p.y - ap[2];
    }
}

// Collimator momentum windows (check_aperture_limit.f90, ecollimator / rcollimator branches):
// m[] = px, py, z, pz (half width, centre) pairs; an ecollimator with both widths of a pair set
// tests that pair as an ellipse. True if the particle is outside; u receives the unstable
// factor as the reference computes it (each failing test overwrites it).
__device__ __forceinline__ bool window_out(double v, double w, double c, double &u)
{
    if (w == 0 || (v > c - w && v < c + w)) return false;
    const double a = (c - w) - v, b = v - (c + w);
    u = a > b ? a : b;
    return true;
}

__device__ __forceinline__ bool window_pair_out(double v0, double v1, CP m, bool ellip, double &u)
{
    const double w0 = m[0], c0 = m[1], w1 = m[2], c1 = m[3];
    if (ellip && w0 != 0 && w1 != 0) {
        const double a = (v0 - c0) / w0, b = (v1 - c1) / w1;
        const double r = a * a + b * b;
        if (!(r > 1)) return false;
        u = sqrt(r) - 1;
        return true;
    }
    const bool o0 = window_out(v0, w0, c0, u);
    const bool o1 = window_out(v1, w1, c1, u);
    return o0 | o1;
}

__device__ __forceinline__ bool momentum_out(const Particle &p, int32_t flags, CP m, double &u)
{
    const bool ellip = flags & F_APE_MOM_ELLIP;
    const bool a = window_pair_out(p.px, p.py, m, ellip, u);
    const bool b = window_pair_out(p.z, p.pz, m + 4, ellip, u);
    return a | b;
}

__device__ __forceinline__ bool aperture_fails(const Particle &p, int32_t flags, CP ap)
{
    if (flags & F_APE_MOMENTUM) {
        double u;
        if (momentum_out(p, flags, ap + 4, u)) return true;
    }
    if (flags & F_APE_ELLIP) {
        double xp, yp;
        ellipse_offsets(p, flags, ap, xp, yp);
        const double xr = xp * ap[1], yr = yp * ap[3];
        return !(xr * xr + yr * yr < 1);
    }
    const double x1 = ap[0], x2 = ap[1], y1 = ap[2], y2 = ap[3];
    return (x1 != 0 && p.x < x1) | (x2 != 0 && p.x > x2) | (y1 != 0 && p.y < y1) | (y2 != 0 && p.y > y2);
}

// The lanes on which aperture_fails holds (the limits are wave-uniform).
__device__ __forceinline__ LaneMask aperture_fail_lanes(const Particle &p, int32_t flags, CP ap)
{
    LaneMask mom = 0;
    if (flags & F_APE_MOMENTUM) {
        double u;
        mom = lanes(momentum_out(p, flags, ap + 4, u));
    }
    if (flags & F_APE_ELLIP) {
        double xp, yp;
        ellipse_offsets(p, flags, ap, xp, yp);
        const double xr = xp * ap[1], yr = yp * ap[3];
        return mom | lanes(!(xr * xr + yr * yr < 1));
    }
    const double x1 = ap[0], x2 = ap[1], y1 = ap[2], y2 = ap[3];
    LaneMask m = 0;
    if (x1 != 0) m |= lanes(p.x < x1);
    if (x2 != 0) m |= lanes(p.x > x2);
    if (y1 != 0) m |= lanes(p.y < y1);
    if (y2 != 0) m |= lanes(p.y > y2);
    return m | mom;
}

// Aperture test of check_aperture_limit.f90:256-324 (elliptical with strict r < 1, or
// rectangular). ap[] layout in bmx_table.hpp.
__device__ __forceinline__ void aperture(Particle &p, int32_t flags, CP ap)
{
    // momentum windows first: lost (state lost$) with their unstable factor, which the
    // spatial test below can override
    double u = 0;
    if (flags & F_APE_MOMENTUM) {
        if (momentum_out(p, flags, ap + 4, u)) p.state = ST_LOST;
    }
    if (flags & F_APE_ELLIP) {
        double xp, yp;
        ellipse_offsets(p, flags, ap, xp, yp);
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
        double f;
        if (x1 != 0 && p.x < x1) { f = fabs((p.x - x1) / (x2 - x1)); if (f > u) { p.state = ST_LOST_NEG_X; u = f; } }
        if (x2 != 0 && p.x > x2) { f = fabs((p.x - x2) / (x2 - x1)); if (f > u) { p.state = ST_LOST_POS_X; u = f; } }
        if (y1 != 0 && p.y < y1) { f = fabs((p.y - y1) / (y2 - y1)); if (f > u) { p.state = ST_LOST_NEG_Y; u = f; } }
        if (y2 != 0 && p.y > y2) { f = fabs((p.y - y2) / (y2 - y1)); if (f > u) { p.state = ST_LOST_POS_Y; u = f; } }
    }
}

// ---- drift -------------------------------------------------------------------------------

// q(p) = 1/sqrt(1 - p) - 1 = sum_k (2k)!/(4^k k!^2) p^k, p = (px^2 + py^2)/(1 + pz)^2.
// For p < 2^-10 the terms after p^6 are below 1e-18 relative.
constexpr double kDriftSeriesMax = 1.0 / 1024;
__device__ __forceinline__ double inv_ps_minus_1(double p)
{
    double q = 0.2255859375;              // 231/1024
    q = fma(q, p, 0.24609375);            // 63/256
    q = fma(q, p, 0.2734375);             // 35/128
    q = fma(q, p, 0.3125);                // 5/16
    q = fma(q, p, 0.375);                 // 3/8
    q = fma(q, p, 0.5);                   // 1/2
    return q * p;
}

// bmad/low_level/track_a_drift.f90:25-116 (direction = orientation = +1, beta > 0):
//   x += L px_rel/ps, y += L py_rel/ps, z += L (ez + sqrt_one(-pxy2)/ps), lost_pz if pxy2 >= 1.
// Fast form: 1/ps = 1 + q(pxy2) and sqrt_one(-pxy2)/ps = 1 - 1/ps = -q(pxy2).
// Exact drift body (track_a_drift.f90:44-80); px_rel = px/(1+pz) computed by division as in
// the reference.
__device__ __forceinline__ void drift_exact(Particle &p, double L)
{
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

// The series body of the fast form (pxy2 < kDriftSeriesMax).
__device__ __forceinline__ void drift_series(Particle &p, double L, double px_rel, double py_rel,
                                             double pxy2)
{
    const double q = inv_ps_minus_1(pxy2);
    const double lx = L * px_rel, ly = L * py_rel;
    p.x = p.x + fma(lx, q, lx);
    p.y = p.y + fma(ly, q, ly);
    p.z = p.z + L * (p.ez - q);
}

template <bool F>
__device__ __forceinline__ void drift(Particle &p, double L)
{
    // L == 0 (either sign) tested on the wave-uniform bit pattern with scalar instructions (the
    // empty asm keeps the compiler from turning the test back into a vector compare)
    unsigned long long bits = __double_as_longlong(L);
    asm volatile("" : "+s"(bits));
    if ((bits << 1) == 0) return;
    if (!F) {
        drift_exact(p, L);
        return;
    }
    const double px_rel = p.px * p.inv_rp;
    const double py_rel = p.py * p.inv_rp;
    const double pxy2 = px_rel * px_rel + py_rel * py_rel;
    const bool series = pxy2 < kDriftSeriesMax;
    // the exact body for the rare lanes outside the series domain (gated per wave), then the
    // series body, written once, for the rest
    if (any_lane(!series)) {
        if (!series) drift_exact(p, L);
    }
    if (series) drift_series(p, L, px_rel, py_rel, pxy2);
}

// ---- quadrupole --------------------------------------------------------------------------

// bmad/low_level/quad_mat2_calc.f90:22-81: 2x2 map and z coefficients of one plane with
// strength k (k > 0 defocusing), given cx = cos/cosh(sqrt|k| L) and sx = sin/sinh(..)/sqrt|k|.
struct Plane {
    double m11, m12, m21, m22, z1, z2, z3;
};

__device__ __forceinline__ Plane quad_plane(double k, double L, double rel_p, double cx, double sx)
{
    Plane q;
    q.m11 = cx;
    q.m22 = cx;
    q.m12 = sx / rel_p;
    q.m21 = k * sx * rel_p;
    q.z1 = k * (-cx * sx + L) / 4;
    q.z2 = -k * sx * sx / (2 * rel_p);
    q.z3 = -(cx * sx + L) / (4 * rel_p * rel_p);
    return q;
}

// Domain of the quadrupole's 5-term plane-function series: t = (k1 L^2/(1+pz))^2 <= 1.
constexpr double kQuadTLim = 1.0;

// The four plane functions of the fast quadrupole form, t = W^2: C(-+W) = A -+ W B and
// S(-+W) = A1 -+ W B1 with A = sum t^m/(4m)!, A1 = sum t^m/(4m+1)!, B = sum t^m/(4m+2)!,
// B1 = sum t^m/(4m+3)!, m = 0..N-1, in power form with the coefficients as multiplicands
// (scalar operands of the FMAs); B1 is summed as 6 B1 so that every sum starts from an inline
// constant. N = 5 for t <= kQuadTLim (the first omitted term t^5/20! is below 4.2e-19), N = 4
// for t <= kQuadShortMax (t^4/16! below 7.3e-19).
constexpr double kQuadShortMax = 1.0 / 16;
struct QuadFns {
    double cxx, cxy, sxx, sxy;
};
template <int N>
__device__ __forceinline__ QuadFns quad_series(double W, double t)
{
    const double t2 = t * t, t3 = t2 * t;
    double A = fma(t, 0.041666666666666664, 1.0);     // 1/4!
    A = fma(2.48015873015873e-05, t2, A);              // 1/8!
    A = fma(2.08767569878681e-09, t3, A);              // 1/12!
    double A1 = fma(t, 0.008333333333333333, 1.0);    // 1/5!
    A1 = fma(2.7557319223985893e-06, t2, A1);          // 1/9!
    A1 = fma(1.6059043836821613e-10, t3, A1);          // 1/13!
    double B = fma(t, 0.001388888888888889, 0.5);     // 1/6!, 1/2!
    B = fma(2.755731922398589e-07, t2, B);             // 1/10!
    B = fma(1.1470745597729725e-11, t3, B);            // 1/14!
    double B1 = fma(t, 0.0011904761904761906, 1.0);   // 6/7!
    B1 = fma(1.5031265031265032e-07, t2, B1);          // 6/11!
    B1 = fma(4.58829823909189e-12, t3, B1);            // 6/15!
    if constexpr (N == 5) {
        const double t4 = t2 * t2;
        A = fma(4.779477332387385e-14, t4, A);         // 1/16!
        A1 = fma(2.8114572543455206e-15, t4, A1);      // 1/17!
        B = fma(1.5619206968586225e-16, t4, B);        // 1/18!
        B1 = fma(4.932381147974598e-17, t4, B1);       // 6/19!
    }
    B1 = B1 * 0.16666666666666666;                     // 1/3!
    return {fma(-W, B, A), fma(W, B, A), fma(-W, B1, A1), fma(W, B1, A1)};
}

// bmad/low_level/track_a_quadrupole.f90:17-148 for n_step = 1 (no multipoles other than k1),
// no fringe, no misalignment; low_energy_z_correction.f90:22-58.
// Fast form: with W = k1 L^2/(1+pz) the plane functions are C(-+W), L S(-+W) where
// C(w) = sum w^n/(2n)! (= cos sqrt(-w) or cosh sqrt(w)) and S(w) = sum w^n/(2n+1)!; both
// planes come from four polynomials in t = W^2 (even / odd parts), n_terms each, valid for
// t <= t_lim (first omitted term below 1e-18); larger |W| is reduced by halving.
template <bool F>
__device__ __forceinline__ void quadrupole(Particle &p, CP q, int32_t n_terms, const Species &sp)
{
    const double L = q[0], b1 = q[1];
    double cx_x, sx_x, cx_y, sx_y;  // planes: x uses k = -k1, y uses k = +k1
    double k1;
    if constexpr (F) {
        k1 = q[3] * p.inv_rp;                 // b1/L / (1 + pz)
        double W = q[2] * p.inv_rp;           // b1 L / (1 + pz)
        double t = W * W;
        // The 4-term series for t <= kQuadShortMax, else the 5-term one, with argument
        // quartering for t > kQuadTLim (far off-momentum particles or strong quads):
        // C(16t) = 2 C(t)^2 - 1, S(16t) = S(t) C(t) in t = W^2, the double-angle identities of
        // cos/cosh and sin/sinh. Chosen per lane, gated per wave: the common path is the short
        // series alone.
        QuadFns f;
        if (any_lane(t > kQuadShortMax)) {
            if (t > kQuadShortMax) {
                int halvings = 0;
                while (t > kQuadTLim && halvings < 64) {
                    W *= 0.25;
                    t *= 0.0625;
                    halvings++;
                }
                f = quad_series<5>(W, t);
                for (int h = 0; h < halvings; h++) {
                    f.sxx = f.sxx * f.cxx;
                    f.sxy = f.sxy * f.cxy;
                    f.cxx = fma(2 * f.cxx, f.cxx, -1.0);
                    f.cxy = fma(2 * f.cxy, f.cxy, -1.0);
                }
            } else {
                f = quad_series<4>(W, t);
            }
        } else {
            f = quad_series<4>(W, t);
        }
        const double cxx = f.cxx, cxy = f.cxy, sxx = f.sxx, sxy = f.sxy;
        cx_x = cxx;
        cx_y = cxy;
        sx_x = L * sxx;
        sx_y = L * sxy;
    } else {
        k1 = b1 / (L * p.rel_p);
        const double sqrt_k = sqrt(fabs(k1));
        const double sk_l = sqrt_k * L;
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
    }
    if constexpr (F) {
        // Per plane (strength k, C = cx, S = sx) with u = p/(1+pz): x' = C x + S u,
        // p' = k S (1+pz) x + C p, dz = [k (L - C S) x^2 - 2 k S^2 x u - (C S + L) u^2] / 4 --
        // quad_mat2_calc.f90:56-61 and the update of track_a_quadrupole.f90:104-112 with the
        // (1+pz) factors of m12, z2 and z3 folded into u.
        const double x = p.x, px = p.px, y = p.y, py = p.py;
        const double ux = px * p.inv_rp, uy = py * p.inv_rp;
        const double wx = sx_x * ux, wy = sx_y * uy;
        const double kxx = -k1 * x, kyy = k1 * y;
        const double dzx = fma(kxx, fma(-2 * sx_x, wx, fma(-cx_x, sx_x, L) * x),
                               -(fma(cx_x, sx_x, L) * ux) * ux);
        const double dzy = fma(kyy, fma(-2 * sx_y, wy, fma(-cx_y, sx_y, L) * y),
                               -(fma(cx_y, sx_y, L) * uy) * uy);
        p.x = fma(cx_x, x, wx);
        p.px = fma(kxx, sx_x * p.rel_p, cx_x * px);
        p.y = fma(cx_y, y, wy);
        p.py = fma(kyy, sx_y * p.rel_p, cx_y * py);
        p.z = p.z + fma(0.25, dzx + dzy, L * p.lez);
    } else {
        const Plane a = quad_plane(-k1, L, p.rel_p, cx_x, sx_x);
        const Plane b = quad_plane(k1, L, p.rel_p, cx_y, sx_y);
        const double x = p.x, px = p.px, y = p.y, py = p.py;
        p.z = p.z + (a.z1 * x * x + a.z2 * x * px + a.z3 * px * px + b.z1 * y * y + b.z2 * y * py +
                     b.z3 * py * py);
        p.x = a.m11 * x + a.m12 * px;
        p.px = a.m21 * x + a.m22 * px;
        p.y = b.m11 * y + b.m12 * py;
        p.py = b.m21 * y + b.m22 * py;
        const double beta0 = sp.beta0, me2 = sp.me2, f = sp.lez_f, thr = sp.lez_thr;
        const double pz = p.pz;
        if (sp.mc2 * sqr(beta0 * pz) < thr)
            p.z = p.z + L * pz * (1 - 3 * (pz * beta0 * beta0) / 2 + pz * pz * f) * me2;
        else
            p.z = p.z + L * (p.beta - beta0) / beta0;
    }
}

// ---- thick multipole ---------------------------------------------------------------------

// bmad/modules/multipole_mod.f90:82-174,314-401 (ab_multipole_kicks, magnetic): kick from
// orders n in mask with strengths (a_n, b_n): c_multi's signed binomials summed as the real and
// imaginary parts of (x + i y)^n, kx = -b Re + a Im, ky = a Re + b Im. The table holds the
// half-step strengths; factor (1 or 2) scales the kick, exactly as scaling the strengths would.
// multipole_sum gives the sums at (x, y); multipole_kick adds them to (px, py).
__device__ __forceinline__ void multipole_sum(double x, double y, CP ah, CP bh, int32_t mask,
                                              double &kx, double &ky)
{
    kx = 0;
    ky = 0;
    if (mask & 1) {
        kx = -bh[0];
        ky = ah[0];
    }
    if (mask >> 1) {
        double zr = x, zi = y;
#pragma unroll
        for (int n = 1; n < 4; n++) {
            if (mask & (1 << n)) {
                kx += -bh[n] * zr + ah[n] * zi;
                ky += ah[n] * zr + bh[n] * zi;
            }
            if (n < 3 && (mask >> (n + 1))) {
                const double nr = zr * x - zi * y;
                zi = zr * y + zi * x;
                zr = nr;
            }
        }
    }
}

__device__ __forceinline__ void multipole_kick(Particle &p, CP ah, CP bh, int32_t mask,
                                               double factor)
{
    double kx, ky;
    multipole_sum(p.x, p.y, ah, bh, mask, kx, ky);
    p.px = fma(factor, kx, p.px);
    p.py = fma(factor, ky, p.py);
}

// The kick of a pure sextupole (mask = order 2 only) with its strengths held in registers;
// NORMAL: no skew strength (a = 0).
template <bool NORMAL>
__device__ __forceinline__ void sextupole_kick(Particle &p, double a, double b, double factor)
{
    const double x = p.x, y = p.y;
    const double re = x * x - y * y, im = x * y + y * x;
    if constexpr (NORMAL) {
        p.px = fma(factor, -b * re, p.px);
        p.py = fma(factor, b * im, p.py);
    } else {
        p.px = fma(factor, -b * re + a * im, p.px);
        p.py = fma(factor, a * re + b * im, p.py);
    }
}

// bmad/low_level/track_a_thick_multipole.f90:18-133 (magnetic, no misalignment):
// half kick, then n_step x (drift, kick), the last kick halved. As in the reference, a drift
// that loses the particle does not end the loop (track_a_drift does not test the state).
template <bool F, bool NORMAL>
__device__ __forceinline__ void sextupole_body(Particle &p, CP m, int32_t n_step)
{
    const double step = m[0];
    const double a = m[3], b = m[7];
    sextupole_kick<NORMAL>(p, a, b, 1.0);
    for (int i = 1; i <= n_step; i++) {
        drift<F>(p, step);
        sextupole_kick<NORMAL>(p, a, b, i == n_step ? 1.0 : 2.0);
    }
}

template <bool F>
__device__ __forceinline__ void thick_multipole(Particle &p, CP m, int32_t n_step, int32_t flags)
{
    const double step = m[0];
    const int32_t mask = (flags >> F_ORDER_SHIFT) & 0xF;
    if (mask == 4) {
        if (flags & F_MULT_NORMAL) sextupole_body<F, true>(p, m, n_step);
        else sextupole_body<F, false>(p, m, n_step);
        return;
    }
    multipole_kick(p, m + 1, m + 5, mask, 1.0);
    for (int i = 1; i <= n_step; i++) {
        drift<F>(p, step);
        multipole_kick(p, m + 1, m + 5, mask, i == n_step ? 1.0 : 2.0);
    }
}

// ---- sbend -------------------------------------------------------------------------------

// bmad/modules/fringe_mod.f90:241-419 hwang_bend_edge_kick (time_dir = +1), orbit only.
// e[] = gt, gt2, gs2, k1_tane, fg_factor, tan_e^2 of that face; g_tot of the bend.
// half_gs2 = gs2/2 of that face.
__device__ __forceinline__ void hwang_edge(Particle &p, CP e, double half_gs2, double g_tot,
                                           bool entering, bool simple)
{
    const double ef = p.inv_rp;
    const double x = p.x, px = p.px, y = p.y, py = p.py;
    if (simple) {
        // e = 0, fint*hgap = 0, k1 = 0: gt = gt2 = k1_tane = fg = 0, gs2 = g_tot:
        // x += gs2 y^2 ef/2, py -= g_tot px y ef, z -= gs2 px y^2 ef^2/2 (signs flipped at the
        // exit face)
        const double yy = y * y;
        const double a = half_gs2 * ef;
        const double gef = g_tot * ef;
        if (entering) {
            p.x = fma(a, yy, x);
            p.py = fma(-gef, px * y, py);
            p.z = fma(-(a * ef), px * yy, p.z);
        } else {
            p.x = fma(-a, yy, x);
            p.py = fma(gef, px * y, py);
            p.z = fma(a * ef, px * yy, p.z);
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

// asin(s)/s = sum_n (2n)!/(4^n n!^2 (2n+1)) s^(2n); for |s| <= 1/8 the terms after s^18 are
// below 1e-18.
constexpr double kArcSeriesMax = 0.125;
__constant__ const double kAsinCoef[10] = {
    0.009761609529194078, 0.011551800896139705, 0.01396484375, 0.017352764423076924,
    0.022372159090909092, 0.030381944444444444, 0.044642857142857144, 0.075,
    0.16666666666666666, 1.0};
__device__ __forceinline__ double asin_over(double s2)
{
    CP c = launder((CP)kAsinCoef);
    double r = c[0];
#pragma unroll
    for (int k = 1; k < 10; k++) r = fma(r, s2, c[k]);
    return r;
}

// The same series for s^2 <= 2^-10 (|s| <= 1/32, which covers the half exit angle of bends
// up to about 3.5 degrees): terms up to s^10, the next (231/13312 s^12) below 2e-20.
constexpr double kArcShortMax = 1.0 / 1024;
// Power form: the coefficients are multiplicands (scalar operands), not addends that would be
// held in vector registers.
__device__ __forceinline__ double asin_over_short(double s2)
{
    const double s4 = s2 * s2, s8 = s4 * s4;
    double r = fma(0.16666666666666666, s2, 1.0);     // 1/6
    r = fma(0.075, s4, r);                             // 3/40
    r = fma(0.044642857142857144, s4 * s2, r);         // 5/112
    r = fma(0.030381944444444444, s8, r);              // 35/1152
    return fma(0.022372159090909092, s8 * s2, r);      // 63/2816
}

// sin and cos of |phi| <= 4 (an rf phase wrapped to [-pi, pi)): quadrant reduction by pi/2
// (two-part constant) and Taylor polynomials on |r| <= pi/4 (truncation below 1e-20).
// Coefficients (Horner order): sin -1/19!, 1/17!, ..., -1/3!; cos 1/20!, -1/18!, ..., 1/4!.
__constant__ const double kSinCosCoef[18] = {
    -8.22063524662433e-18, 2.8114572543455206e-15, -7.647163731819816e-13,
    1.6059043836821613e-10, -2.505210838544172e-08, 2.7557319223985893e-06,
    -0.0001984126984126984, 0.008333333333333333, -0.16666666666666666,
    4.110317623312165e-19, -1.5619206968586225e-16, 4.779477332387385e-14,
    -1.1470745597729725e-11, 2.08767569878681e-09, -2.755731922398589e-07,
    2.48015873015873e-05, -0.001388888888888889, 0.041666666666666664};
__device__ __forceinline__ void sincos_reduced(double phi, double *sp, double *cp)
{
    const double kPio2Hi = 1.5707963267948966;      // pi/2 rounded to double
    const double kPio2Lo = 6.123233995736766e-17;   // pi/2 - kPio2Hi
    const double kq = rint(phi * 0.6366197723675814);  // 2/pi
    const double r = fma(-kq, kPio2Lo, fma(-kq, kPio2Hi, phi));
    const double r2 = r * r;
    // sin r = r (1 - r^2/3! + ... - r^18/19!), cos r = 1 - r^2/2! + ... + r^20/20!
    CP c = launder((CP)kSinCosCoef);
    double s = c[0];
#pragma unroll
    for (int k = 1; k < 9; k++) s = fma(s, r2, c[k]);
    s = r * fma(s, r2, 1.0);
    double cc = c[9];
#pragma unroll
    for (int k = 10; k < 18; k++) cc = fma(cc, r2, c[k]);
    cc = fma(cc, r2, -0.5);
    const double co = fma(cc, r2, 1.0);
    const int q = ((int)kq) & 3;
    *sp = q == 0 ? s : (q == 1 ? co : (q == 2 ? -s : -co));
    *cp = q == 0 ? co : (q == 1 ? -s : (q == 2 ? -co : s));
}

// 1/sqrt(v) for normal v > 0: the hardware estimate y refined once as y (1 + e/2 + 3 e^2/8),
// e = 1 - v y^2 (the series of (1 - e)^(-1/2), cubically convergent). Unlike the library
// rsqrt it has no zero/infinity/denormal cases (the callers' arguments are near 1 or L^2).
__device__ __forceinline__ double rsqrt_pos(double v)
{
    const double y = __builtin_amdgcn_rsq(v);
    const double e = fma(-v * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// a / d for a normal d: the hardware reciprocal estimate (relative error below 1e-7 on gfx950)
// refined by one Newton step (error e^2, below 1e-14), the quotient corrected once with its
// exact residual a - d q (error of order e^4 before the final rounding).
__device__ __forceinline__ double div_normal(double a, double d)
{
    double y = __builtin_amdgcn_rcp(d);
    y = fma(fma(-d, y, 1.0), y, y);
    const double q = a * y;
    return fma(fma(-d, q, a), y, q);
}

// sqrt(v) for normal v > 0: s = v rsqrt(v) with one Newton correction of the residual v - s^2.
__device__ __forceinline__ double sqrt_pos(double v)
{
    const double y = rsqrt_pos(v);
    const double s = v * y;
    return fma(fma(-s, s, v), 0.5 * y, s);
}

// sqrt(1 - u) for u < 2^-10 (terms after u^6 below 1e-19).
__device__ __forceinline__ double sqrt_1m(double u)
{
    double r = -0.0205078125;
    r = fma(r, u, -0.02734375);
    r = fma(r, u, -0.0390625);
    r = fma(r, u, -0.0625);
    r = fma(r, u, -0.125);
    r = fma(r, u, -0.5);
    return fma(r, u, 1.0);
}

// sqrt(1 - u) - 1 for u < 2^-10, without the cancellation of forming sqrt(1 - u) first: the
// same series times u (the first omitted term, -33/2048 u^7, is below 3e-20 of the result).
// Power form, coefficients as multiplicands (scalar operands), as asin_over_short.
__device__ __forceinline__ double sqrt_1m_minus_1(double u)
{
    const double u2 = u * u, u4 = u2 * u2;
    double r = fma(-0.125, u, -0.5);                   // -1/2, -1/8
    r = fma(-0.0625, u2, r);                           // -1/16
    r = fma(-0.0390625, u2 * u, r);                    // -5/128
    r = fma(-0.02734375, u4, r);                       // -7/256
    r = fma(-0.0205078125, u4 * u, r);                 // -21/1024
    return r * u;
}

// atan(t) for t^2 <= 2^-10: t (1 - t^2/3 + t^4/5 - t^6/7 + t^8/9 - t^10/11); the first omitted
// term, t^13/13, is below 7e-20 of the result. Power form as sqrt_1m_minus_1.
constexpr double kAtanSeriesMax = 1.0 / 1024;
__device__ __forceinline__ double atan_short(double t)
{
    const double t2 = t * t, t4 = t2 * t2, t8 = t4 * t4;
    double r = fma(-0.3333333333333333, t2, 1.0);     // -1/3
    r = fma(0.2, t4, r);                               // 1/5
    r = fma(-0.14285714285714285, t4 * t2, r);         // -1/7
    r = fma(0.1111111111111111, t8, r);                // 1/9
    r = fma(-0.09090909090909091, t8 * t2, r);         // -1/11
    return r * t;
}

// h / sin(h) for the angle h with sine sh and cosine ch (|h| < pi) outside the direct series
// domain: k half-angle steps s' = s / sqrt(2 (1 + c)), c' = sqrt((1 + c)/2) bring |s| below 1/8,
// then h / sin h = 2^k asin(s_k)/s_k * s_k / sh.
__device__ __forceinline__ double h_over_sin_far(double sh, double ch)
{
    if (sh == 0) return 1.0;
    double s = sh, c = ch, scale = 1.0;
    for (int k = 0; k < 8 && !(fabs(s) <= kArcSeriesMax && c > 0); k++) {
        const double t = sqrt(2 * (1 + c));
        s = s / t;
        c = 0.5 * t;
        scale *= 2;
    }
    return scale * asin_over(s * s) * (s / sh);
}

// Semi-linear body for particles within 1e-9 of the zero orbit (track_a_bend.f90:136-159).
__device__ __forceinline__ void sbend_semilinear(Particle &p, CP b, const Species &sp)
{
    const double L = b[0], g = b[1];
    const double cos_a = b[4], sin_a = b[5], sinc_a = b[6], cosc_a = b[7];
    const double x = p.x, px = p.px, y = p.y, py = p.py, z = p.z;
    const double ll = L;
    const double gam2 = sp.mc2_sq / (sqr(p.rel_p * sp.p0c) + sp.mc2_sq);
    const double m56 = ll * (gam2 - sqr(g * ll) * b[9]);
    p.x = cos_a * x + ll * sinc_a * px + g * ll * ll * cosc_a * p.pz;
    p.px = -g * sin_a * x + cos_a * px + g * ll * sinc_a * p.pz;
    p.y = y + ll * py;
    p.z = -g * ll * sinc_a * x - g * ll * ll * cosc_a * px + z + m56 * p.pz;
}

// Exact sector-bend body as written in the reference (track_a_bend.f90:163-209).
__device__ __forceinline__ void sbend_exact_ref(Particle &p, CP b)
{
    const double L = b[0], g = b[1], g_tot = b[2], angle = b[3];
    const double cos_a = b[4], sin_a = b[5], sinc_a = b[6], cosc_a = b[7], beta_ref = b[8];
    const double x = p.x, px = p.px, y = p.y, py = p.py, z = p.z;
    const double rel_p = p.rel_p;
    const double pt2 = rel_p * rel_p - py * py;
    const double pt = sqrt(pt2);
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

// Fast form of the exact body: with s1 = px/pt = sin(phi_1) and the bend angle's cos/sin,
// sin/cos(angle + phi_1) follow by angle addition; with psi = atan2(L_u, -L_v) known through
// its cosine -L_v/L_c and sine L_u/L_c, the half exit angle h = angle_p/2 = angle + phi_1 - psi
// has sin h and cos h by angle subtraction, L_p = L_c h / sin h from the asin(s)/s series, and
// px = pt sin(angle + phi_1 - 2h) by double-angle identities. No transcendental calls; the
// rare cases (a square root outside the series domain, a losing particle, a particle turned
// by more than 90 degrees) are gated per wave and decided per lane.
__device__ __forceinline__ void sbend_exact_fast(Particle &p, CP b)
{
    const double L = b[0], g = b[1], g_tot = b[2];
    const double cos_a = b[4], sin_a = b[5], sinc_a = b[6];
    const double x = p.x, px = p.px, y = p.y, py = p.py, z = p.z;
    const double rel_p = p.rel_p;
    const double pt2 = rel_p * rel_p - py * py;
    // pt2 <= 0 gives NaN coordinates here as sqrt does in the reference form
    const double inv_pt = rsqrt_pos(pt2);
    const double pt = pt2 * inv_pt;
    const double g_p = g_tot * inv_pt;
    const double s1 = px * inv_pt;
    const double s1s = s1 * s1;
    double c1 = sqrt_1m(s1s);
    const bool c1_far = !(s1s < kDriftSeriesMax);
    if (any_lane(c1_far)) {
        if (c1_far) c1 = sqrt(rare(1 - s1s));
    }
    const double sin_plus = fma(sin_a, c1, cos_a * s1);
    const double cos_plus = fma(cos_a, c1, -sin_a * s1);
    const double u = fma(g, x, 1.0) * b[24];                 // (1 + g x) L sinc_a
    const double alpha = u * fma(-g_p, u, 2 * sin_plus);     // 2 sin_plus u - g_p u^2
    const double r = fma(cos_plus, cos_plus, g_p * alpha);
    // r <= 0 loses the particle below; the floor keeps the square root's argument normal
    const double rad = sqrt_pos(fmax(r, 0x1p-1022));
    double xi = div_normal(alpha, rad + cos_plus);
    const bool turned = !(cos_plus > 0);
    if (any_lane(turned)) {
        if (turned) xi = (rad - cos_plus) / g_p;
    }
    const double x_new = x * cos_a - b[22] + xi;
    if (lanes(fabs(px) > pt) | lanes(r < 0) |
        (lanes(fabs(g_p) < 1e-5) & lanes(fabs(cos_plus) < 1e-5)) |
        lanes(fabs(x_new) > kMaxAperture)) {
        if (fabs(px) > pt) {
            p.state = ST_LOST_PZ;
            return;
        }
        if ((r < 0) | ((fabs(g_p) < 1e-5) & (fabs(cos_plus) < 1e-5))) {
            p.state = ST_LOST;
            p.x = 2 * kMaxAperture;
            return;
        }
        if (fabs(x_new) > kMaxAperture) {
            p.state = ST_LOST;
            p.x = x_new;
            return;
        }
    }
    const double L_u = xi;
    const double mL_v = fma(x, sin_a, b[24]);  // -L_v = L sinc_a + x sin_a
    const double lc2 = mL_v * mL_v + L_u * L_u;
    const double inv_lc = rsqrt_pos(lc2);
    const double L_c = lc2 * inv_lc;
    const double cpsi = mL_v * inv_lc, spsi = L_u * inv_lc;
    const double sh = fma(sin_plus, cpsi, -cos_plus * spsi);  // sin(angle_p/2)
    const double ch = fma(cos_plus, cpsi, sin_plus * spsi);   // cos(angle_p/2)
    // L_p = L_c / sinc(h) = L_c h / sin h
    const double sh2 = sh * sh;
    double hs = asin_over_short(sh2);
    if (lanes(!(sh2 <= kArcShortMax)) | lanes(!(ch > 0))) {
        if (!(sh2 <= kArcShortMax && ch > 0)) {
            if (fabs(sh) <= kArcSeriesMax && ch > 0) hs = asin_over(sh2);
            else hs = h_over_sin_far(sh, ch);
        }
    }
    const double L_p = L_c * hs;
    // sin(angle + phi_1 - 2h) with sin/cos 2h from sh, ch
    const double s2h = 2 * sh * ch;
    const double c2h = fma(-2 * sh, sh, 1.0);
    p.x = x_new;
    p.px = pt * fma(sin_plus, c2h, -cos_plus * s2h);
    const double lp_pt = L_p * inv_pt;
    p.y = fma(py, lp_pt, y);
    p.z = z + (p.beta * b[23] - rel_p * lp_pt);
}

// The exact body in momentum form, for g_tot = g (the same orbit as track_a_bend.f90:163-209:
// a circle of radius pt/g in the horizontal plane from the entrance face to the exit face of
// a sector bend of angle theta). With the longitudinal momenta pz_t = sqrt((1+pz)^2 - px^2 -
// py^2) at the entrance and pz_f at the exit, the circle's geometry gives in closed form
//   px_f = px cos theta - (g x - (pz_t - 1)) sin theta,
//   x_f  = x cos theta + ((pz_f - pz_t) + (pz_t - 1)(1 - cos theta) + px sin theta) / g,
// and the orbit turns by Phi = theta + phi_1 - phi_f (phi = atan2(px, pz) at each face), so
// the path length in the horizontal plane is L_p = pt Phi / g (y += py L_p/pt, z -= (1+pz)
// L_p/pt, as the reference). pz - 1 comes from the series of sqrt(1 - u) - 1 and
// tan(phi_1 - phi_f) = (px pz_f - px_f pz_t)/(pz_t pz_f + px px_f), so the body needs no
// transcendental call, one division and no square root, and every difference of nearly equal
// quantities is formed from small terms. The reference's form evaluates the same geometry
// through asin, atan2 and sin of angle sums; the two agree to a few units in the last place of
// the coordinates' scale. Returns false on lanes outside the form's domain (|p_perp|/p >=
// 2^-5 at either face, |phi_1 - phi_f| beyond the atan series, |x_f| above orbit_too_large's
// limit), whose coordinates are then untouched; the caller takes those through the general
// form, which also decides every loss inside the body.
__device__ __forceinline__ bool sbend_momentum_form(Particle &p, CP b)
{
    const double x = p.x, px = p.px, py = p.py, pz = p.pz, rel_p = p.rel_p;
    const double g = b[1], cos_a = b[4], sin_a = b[5], omc = b[27], inv_g = b[28];
    const double irp2 = p.inv_rp * p.inv_rp;
    const double py2 = py * py;
    const double su = fma(px, px, py2) * irp2;
    const double m_t = sqrt_1m_minus_1(su);
    const double pzt_m1 = fma(rel_p, m_t, pz);        // pz_t - 1
    const double pz_t = 1.0 + pzt_m1;
    const double pxf = fma(px, cos_a, -fma(g, x, -pzt_m1) * sin_a);
    const double sf = fma(pxf, pxf, py2) * irp2;
    const double m_f = sqrt_1m_minus_1(sf);
    const double pz_f = 1.0 + fma(rel_p, m_f, pz);
    const double br = fma(px, sin_a, fma(pzt_m1, omc, rel_p * (m_f - m_t)));
    const double xf = fma(br, inv_g, x * cos_a);
    const double S = fma(px, pz_f, -pxf * pz_t);
    const double C = fma(pz_t, pz_f, px * pxf);
    const double t = div_normal(S, C);
    const bool ok = (su < kDriftSeriesMax) & (sf < kDriftSeriesMax) & (t * t <= kAtanSeriesMax) &
                    (fabs(xf) <= kMaxAperture);
    if (ok) {
        const double lp_pt = (b[3] + atan_short(t)) * inv_g;  // L_p / pt = Phi / g
        p.x = xf;
        p.px = pxf;
        p.y = fma(py, lp_pt, p.y);
        p.z = p.z + (p.beta * b[23] - rel_p * lp_pt);
    }
    return ok;
}

// orbit_too_large at a point inside an element: true (and the state set) if it flags the
// particle; the test is branch-free, the state assignment gated.
__device__ __forceinline__ bool orbit_check(Particle &p)
{
    // rel_p < 0 cannot newly hold inside a bend (pz does not change there), so the gate tests
    // the rest of orbit_bad only
    const double f_unstable = p.px * p.px + p.py * p.py - p.rel_p * p.rel_p;
    if (lanes(fabs(p.x) > kMaxAperture) | lanes(fabs(p.y) > kMaxAperture) |
        lanes(f_unstable > 0)) {
        // only live particles enter an element, so a lane not alive now was flagged here
        if (orbit_bad(p)) orbit_too_large(p);
        return p.state != ST_ALIVE;
    }
    return false;
}

// bmad/low_level/track_a_bend.f90:18-333: exact_multipoles = off, no body multipoles, b1 = 0,
// n_step = 1, no misalignment. Hwang entrance kick, orbit_too_large (:54), body (exact
// :163-209, semi-linear near the zero orbit :136-159, drift when g = dg = 0 :117-118),
// orbit_too_large (:319), Hwang exit kick.
// SB: both edges have the simple Hwang kick and the body is not a drift (the flags are then
// known at compile time, so the edge code has no run-time branches to merge; F_BEND_SERIES
// stays as given).
template <bool F, bool SB>
__device__ __forceinline__ void sbend(Particle &p, int32_t flags, CP b, const Species &sp)
{
    if constexpr (SB) flags = F_EDGE_ENT | F_EDGE_EXIT | F_EDGE_SIMPLE | (flags & F_BEND_SERIES);
    const double L = b[0], g = b[1], g_tot = b[2];
    const bool simple = flags & F_EDGE_SIMPLE;
    if (flags & F_EDGE_ENT) hwang_edge(p, b + 10, b[25], g_tot, true, simple);
    if (orbit_check(p)) return;
    if (flags & F_BEND_DRIFT) {
        drift<F>(p, L);
    } else {
        // the semi-linear branch needs |pz| < 1e-9 among its conditions: gated on that alone
        bool semi = false;
        if (any_lane(fabs(p.pz) < 1e-9)) {
            semi = (fabs(p.x * g) < 1e-9) & (fabs(p.px) < 1e-9) & (fabs(p.py) < 1e-9) &
                   (fabs(p.pz) < 1e-9);
            if (semi) sbend_semilinear(p, b, sp);
        }
        if (!semi) {
            if (F && (flags & F_BEND_SERIES)) {
                // the momentum form on the common lanes, the general form on the others
                const bool done = sbend_momentum_form(p, b);
                if (any_lane(!done)) {
                    if (!done) sbend_exact_fast(p, b);
                }
            } else if (F) {
                sbend_exact_fast(p, b);
            } else {
                sbend_exact_ref(p, b);
            }
        }
        if (p.state != ST_ALIVE) return;
    }
    if (orbit_check(p)) return;
    if (flags & F_EDGE_EXIT) hwang_edge(p, b + 16, b[26], g_tot, false, simple);
}

// ---- rf cavity ---------------------------------------------------------------------------

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
template <bool F>
__device__ __forceinline__ void rfcavity(Particle &p, CP r, int32_t n_slice,
                                         const Species &sp_)
{
    const double dl = r[0], voltage = r[1], phi0 = r[2], freq = r[3], half_period = r[4];
    const double twopi_f = r[5];
    double t = -p.z / (p.beta * kCLight);
    if (!(fabs(t) < half_period)) t = modulo2(t, half_period);
    // phase0 = twopi * (phi0 - t * freq) with the product and difference carried exactly
    const double tf_hi = t * freq;
    const double tf_lo = fma(t, freq, -tf_hi);
    const double d_hi = phi0 - tf_hi;
    const double d_bb = d_hi - phi0;
    const double d_lo = ((phi0 - (d_hi - d_bb)) + (-tf_hi - d_bb)) - tf_lo;
    const double ph_hi = kTwoPi * d_hi;
    const double ph_lo = fma(kTwoPi, d_hi, -ph_hi) + kTwoPi * d_lo;
    double phase0 = ph_hi + ph_lo;
    phase0 = phase0 + r[6];
    double phase = modulo2(phase0, kPi);
    if (F) {
        // sin(phase) of each slice by rotating (sin, cos) through the per-slice phase step
        // d = 2 pi f dz/(c beta); d is tiny (|d| << 1e-2), so cos d and sin d come from short
        // series (truncation below 1e-20); a larger step re-evaluates sincos.
        double sp, cp;
        sincos_reduced(phase, &sp, &cp);
        for (int i = 0; i <= n_slice; i++) {
            double dE = voltage * sp / n_slice;
            if (i == 0 || i == n_slice) dE = dE / 2;
            if (!energy_kick(p, dE, sp_)) return;
            refresh_energy(p, sp_);
            if (p.pz == -1) {
                p.state = ST_LOST_PZ;
                return;
            }
            if (i != n_slice) {
                const double z0 = p.z;
                drift<F>(p, dl);
                const double d = twopi_f * (p.z - z0) / (kCLight * p.beta);
                phase = phase + d;
                if (fabs(d) < 1e-3) {
                    const double d2 = d * d;
                    const double cd = fma(fma(d2, 1.0 / 24, -0.5), d2, 1.0);
                    const double sd = d * fma(fma(d2, 1.0 / 120, -1.0 / 6), d2, 1.0);
                    const double ns = fma(sp, cd, cp * sd);
                    cp = fma(cp, cd, -sp * sd);
                    sp = ns;
                } else {
                    sincos_reduced(modulo2(phase, kPi), &sp, &cp);
                }
            }
        }
        return;
    }
    for (int i = 0; i <= n_slice; i++) {
        double dE = voltage * sin(phase) / n_slice;
        if (i == 0 || i == n_slice) dE = dE / 2;
        if (!energy_kick(p, dE, sp_)) return;
        refresh_energy(p, sp_);
        if (p.pz == -1) {
            p.state = ST_LOST_PZ;
            return;
        }
        if (i != n_slice) {
            const double z0 = p.z;
            drift<F>(p, dl);
            phase = phase + twopi_f * (p.z - z0) / (kCLight * p.beta);
        }
    }
}

// ---- quadrupole with multipoles ----------------------------------------------------------

// bmad/modules/multipole_mod.f90:82-174 ab_multipole_kicks, electric branch (:128-143), with
// ab_multipole_kick's electric strengths a2 = a f, b2 = -b f, f = charge/(beta p0c)
// (:343-346): the table holds a f beta and b2 beta at the half step; the kick changes the
// energy, pz += (1+pz) sqrt_one(alpha), alpha = (kx (2 px + kx) + ky (2 py + ky))/(1+pz)^2
// (lost_pz for alpha < -1, the coordinates kicked and pz left as it was), beta from the new
// momentum and z scaled by beta_new/beta_old.
__device__ __forceinline__ void elec_kick(Particle &p, CP ae, CP be, int32_t mask, double factor,
                                          const Species &sp)
{
    double kx, ky;
    multipole_sum(p.x, p.y, ae, be, mask, kx, ky);
    const double fb = factor / p.beta;
    kx = kx * fb;
    ky = ky * fb;
    const double px0 = p.px, py0 = p.py;
    p.px = px0 + kx;
    p.py = py0 + ky;
    const double alpha = (kx * (2 * px0 + kx) + ky * (2 * py0 + ky)) / (p.rel_p * p.rel_p);
    if (alpha < -1) {
        p.state = ST_LOST_PZ;
        return;
    }
    p.pz = p.pz + p.rel_p * sqrt_one(alpha);
    const double rp = 1 + p.pz;
    const double beta_new = rp / sqrt(rp * rp + sqr(sp.mc2 / sp.p0c));
    p.z = p.z * beta_new / p.beta;
    p.beta = beta_new;
    refresh_energy(p, sp);
}

// bmad/low_level/apply_element_edge_kick.f90:283-326 electric_longitudinal_fringe (bmad_standard
// field, not a bend): crossing the hard edge of the electric multipole field changes the
// energy by dE = sign * charge * sum Re((b_n - i a_n) (x + i y)^(n+1))/(n+1), sign = +1 at
// the entrance and -1 at the exit (apply_energy_kick.f90:23-75). fa, fb hold charge a_n/(n+1)
// and charge b_n/(n+1). Returns false if the particle is lost (energy below the rest mass).
__device__ __forceinline__ bool elec_fringe(Particle &p, CP fa, CP fb, int32_t mask, double sign,
                                            const Species &sp)
{
    const double x = p.x, y = p.y;
    double zr = x, zi = y, dE = 0;
#pragma unroll
    for (int n = 0; n < 4; n++) {
        if (mask & (1 << n)) dE += fb[n] * zr + fa[n] * zi;
        if (n < 3 && (mask >> (n + 1))) {
            const double nr = zr * x - zi * y;
            zi = zr * y + zi * x;
            zr = nr;
        }
    }
    if (!energy_kick(p, sign * dE, sp)) return false;
    refresh_energy(p, sp);
    return true;
}

// bmad/modules/fringe_mod.f90:692-865 hard_multipole_edge_kick of a quadrupole (n = 1, bp = k1,
// ap = 0) in closed form: with c = -+k1/(12 (1+pz)) (minus at the entrance), the complex
// polynomials of the reference reduce to fx = c x (x^2 + 3 y^2), fy = -c y (3 x^2 + y^2),
// dfx/dx = -dfy/dy = 3 c (x^2 + y^2), dfx/dy = -dfy/dx = 6 c x y; then x -= fx, y -= fy, the
// momenta from the inverse Jacobian and z += (px fx + py fy)/(1+pz) (:823-830). h12 = k1/12.
// The division by 1+pz is kept (not the stored reciprocal) so that a particle an electric edge
// has just taken to pz = -1 ends as the reference's does.
__device__ __forceinline__ void quad_hard_edge(Particle &p, double h12, bool entering)
{
    const double rel_p = 1 + p.pz;
    const double c = (entering ? -h12 : h12) / rel_p;
    const double x = p.x, y = p.y, px = p.px, py = p.py;
    const double xx = x * x, yy = y * y;
    const double fx = c * x * (xx + 3 * yy);
    const double fy = -c * y * (3 * xx + yy);
    const double d = 3 * c * (xx + yy);          // dfx/dx = -dfy/dy
    const double e = 6 * c * x * y;              // dfx/dy = -dfy/dx
    const double denom = (1 - d) * (1 + d) + e * e;
    p.x = x - fx;
    p.px = px + ((1.0 + d - denom) * px - e * py) / denom;
    p.y = y - fy;
    p.py = py + (e * px + (1.0 - d - denom) * py) / denom;
    p.z = p.z + (p.px * fx + p.py * fy) / rel_p;
}

// bmad/modules/fringe_mod.f90:556-661 soft_quadrupole_edge_kick (time_dir = +1): f1 = k1 fq1/(1+pz)
// (negated at the exit), f2 = k1 fq2/(1+pz); kf1 = k1 fq1, kf2 = k1 fq2.
__device__ __forceinline__ void quad_soft_edge(Particle &p, double kf1, double kf2, bool entering)
{
    const double rel_p = 1 + p.pz;
    double f1 = kf1 / rel_p;
    const double f2 = kf2 / rel_p;
    if (!entering) f1 = -f1;
    const double ef1 = exp(f1);
    const double x = p.x, px = p.px, y = p.y, py = p.py;
    const double vx = px / rel_p, vy = py / rel_p;
    p.z = p.z - (f1 * x + f2 * (1 + f1 / 2) * vx / ef1) * vx + (f1 * y + f2 * (1 - f1 / 2) * vy * ef1) * vy;
    p.x = x * ef1 + vx * f2;
    p.px = px / ef1;
    p.y = y / ef1 - vy * f2;
    p.py = py * ef1;
}

// bmad/low_level/track_a_quadrupole.f90:17-148 with multipoles beside k1 and edge fields: the
// entrance edge (hard multipole edge, soft quadrupole edge, then the electric edge:
// init_fringe_info.f90:44-54 gives an element with electric multipoles an edge kick whatever
// its fringe_type; apply_element_edge_kick.f90:165-178,270-271 orders the electric edge after
// the magnetic ones on entry and before them on exit), half kicks, then n_step quadrupole steps
// with k1 = b1/(L (1+pz)) taken from the momentum at each step, each followed by a full kick
// (half after the last), then the exit edge. Returns false when an edge loses the particle:
// the reference then returns before offset_particle's exit transform.
template <bool F>
__device__ __forceinline__ bool quad_multipole(Particle &p, CP q, int32_t n_step, int32_t flags,
                                               const Species &sp)
{
    const int32_t mmask = (flags >> F_ORDER_SHIFT) & 0xF;
    const int32_t emask = (flags >> F_ELEC_SHIFT) & 0xF;
    if (flags & F_QEDGE_ENT) {
        if (flags & F_QHARD) quad_hard_edge(p, q[29], true);
        if (flags & F_QSOFT) quad_soft_edge(p, q[30], q[31], true);
        if (emask && !elec_fringe(p, q + 21, q + 25, emask, 1.0, sp)) return false;
    }
    if (mmask) multipole_kick(p, q + 5, q + 9, mmask, 1.0);
    if (emask) elec_kick(p, q + 13, q + 17, emask, 1.0, sp);
    for (int i = 1; i <= n_step; i++) {
        quadrupole<F>(p, q, 0, sp);
        const double f = i == n_step ? 1.0 : 2.0;
        if (mmask) multipole_kick(p, q + 5, q + 9, mmask, f);
        if (emask) elec_kick(p, q + 13, q + 17, emask, f, sp);
    }
    if (flags & F_QEDGE_EXIT) {
        const bool kept = !emask || elec_fringe(p, q + 21, q + 25, emask, -1.0, sp);
        if (flags & F_QSOFT) quad_soft_edge(p, q[30], q[31], false);
        if (flags & F_QHARD) quad_hard_edge(p, q[29], false);
        if (!kept) return false;
    }
    return true;
}

// One element for one particle: the body of track1 (bmad/code/track1.f90:196-329) after the
// entrance validity check, i.e. entrance aperture, element map, orbit_too_large (:291),
// exit aperture (:329). h is the element's header (bmx_table.hpp); prm the lattice's parameter
// array. Returns true when the particle was lost at the entrance
// aperture (location = upstream_end$).
// MAPS selects the code compiled in: MAPS_ALL (every kind, offsets), or the pair MAPS_NO_RF /
// MAPS_RARE for loops that keep the rare, register-hungry cases out of the common element
// path: MAPS_NO_RF has no rf cavity, no offset_particle frame change and no collimator
// momentum windows; MAPS_RARE has only the kinds that can carry them (no bends) with all three;
// MAPS_ALL_Q and MAPS_RARE_Q add the stepped quadrupole (K_QMULT) and the general bend
// (K_BMULT) to MAPS_ALL and MAPS_RARE, for kernels built for lattices that hold one (their
// code would otherwise shape the register assignment of every kernel); MAPS_ALL_QR and
// MAPS_RARE_QR add the radiation kicks to those, for lattices with radiation (the kicks raise
// the register count of every kernel they are compiled into); MAPS_RARE_R is MAPS_RARE with
// the kicks, for the cell kernels of lattices with radiation and no stepped elements.
// from_prev: the particle has just left the element before this one in the same lattice, so
// an entrance test that repeats that element's exit test (F_APE_ENT_SAME, F_APE_ENT cleared)
// is skipped; otherwise it is made.
// A particle lost here gets p.loss_ele = ix1 (the element's 1-based index), p.loss_turn = turn.
enum MapSet : int {
    MAPS_ALL = 0, MAPS_NO_RF = 1, MAPS_RARE = 2, MAPS_RARE_Q = 3, MAPS_ALL_Q = 4, MAPS_RARE_QR = 5,
    MAPS_ALL_QR = 6, MAPS_RARE_R = 7
};
// The map set carries the stepped quadrupole and the general bend.
constexpr bool maps_q(int m) { return m == MAPS_RARE_Q || m == MAPS_ALL_Q || m == MAPS_RARE_QR || m == MAPS_ALL_QR; }
// The rare paths (MAPS_RARE_QR carries the plain bend all the same: a bend that radiates takes
// the rare path).
constexpr bool maps_rare(int m)
{
    return m == MAPS_RARE || m == MAPS_RARE_Q || m == MAPS_RARE_QR || m == MAPS_RARE_R;
}
// The map set carries the radiation kicks.
constexpr bool maps_rad(int m) { return m == MAPS_RARE_QR || m == MAPS_ALL_QR || m == MAPS_RARE_R; }

// ---- offset_particle, non-bend branch (bmad/code/offset_particle.f90:183-492) -------------
// Element orientation +1, particle direction +1; ob is the offset block (bmx_table.hpp).

// bmad/low_level/track_a_drift.f90:25-116 with include_ref_motion = False (dz = -L/ps): the
// drift to the element edge after the frame change.
__device__ __forceinline__ void drift_edge(Particle &p, double ds)
{
    const double px_rel = p.px / p.rel_p;
    const double py_rel = p.py / p.rel_p;
    const double pxy2 = px_rel * px_rel + py_rel * py_rel;
    if (pxy2 >= 1) {
        p.state = ST_LOST_PZ;
        return;
    }
    const double ps = sqrt(1 - pxy2);
    p.x = p.x + ds * px_rel / ps;
    p.y = p.y + ds * py_rel / ps;
    p.z = p.z + (-ds / ps);
}

// The frame change of a misaligned element (:183-230 entering, :430-470 leaving): position
// (r0, r1, r2) and momentum (px, py, ps) rotated by w (3x3, row-major), then (add0, add1) added
// to the new x, y; s_edge receives the rotated longitudinal position. A particle with
// px^2 + py^2 >= (1 + pz)^2 is lost with its coordinates unchanged.
__device__ __forceinline__ void offset_frame(Particle &p, CP w, double r0, double r1, double r2,
                                             double add0, double add1, double &s_edge)
{
    const double s0 = w[0] * r0 + w[1] * r1 + w[2] * r2;
    const double s1 = w[3] * r0 + w[4] * r1 + w[5] * r2;
    s_edge = w[6] * r0 + w[7] * r1 + w[8] * r2;
    const double pz2 = p.rel_p * p.rel_p - p.px * p.px - p.py * p.py;
    if (pz2 <= 0) {
        p.state = ST_LOST_PZ;
        return;
    }
    const double ps = sqrt(pz2);
    const double q0 = w[0] * p.px + w[1] * p.py + w[2] * ps;
    const double q1 = w[3] * p.px + w[4] * p.py + w[5] * ps;
    p.px = q0;
    p.py = q1;
    p.x = s0 + add0;
    p.y = s1 + add1;
}

// bmad/code/tilt_coords.f90:18-52 with (cos, sin) of the angle.
__device__ __forceinline__ void tilt_rotate(Particle &p, double c, double s)
{
    const double x = p.x, y = p.y, px = p.px, py = p.py;
    p.x = c * x + s * y;
    p.y = -s * x + c * y;
    p.px = c * px + s * py;
    p.py = -s * px + c * py;
}

__device__ __forceinline__ void offset_set(Particle &p, CP ob, int32_t flags)
{
    if (flags & F_MISALIGN) {
        double s_body;
        offset_frame(p, ob + 4, p.x - ob[0], p.y - ob[1], 0.0 - (ob[2] + ob[3]), 0.0, 0.0, s_body);
        s_body = s_body + ob[3];
        // drift from s_body to the entrance face (s = 0)
        const double ds = 0.0 - s_body;
        if (p.state == ST_ALIVE && fabs(ds) > 1e-14) drift_edge(p, ds);
    }
    if (flags & F_TILT) tilt_rotate(p, ob[22], ob[23]);
}

__device__ __forceinline__ void offset_unset(Particle &p, CP ob, int32_t flags)
{
    if (flags & F_TILT) tilt_rotate(p, ob[22], -ob[23]);
    if (flags & F_MISALIGN) {
        double s_lab;
        // the exit face sits at ds_center = L/2 - L = -L/2 from the element centre
        offset_frame(p, ob + 13, p.x, p.y, 0.0 - (-ob[3]), ob[0], ob[1], s_lab);
        s_lab = s_lab + (ob[2] + ob[3]);
        const double ds = (ob[3] + ob[3]) - s_lab;
        if (p.state == ST_ALIVE && fabs(ds) > 1e-14) drift_edge(p, ds);
    }
}

// ---- the general sector bend -------------------------------------------------------------

// bmad/modules/fringe_mod.f90:1713-1771 bmad_rot_xz on (x, px, y, py, z) with pp = 1 + pz: the
// frame rotated by a (ca = cos a, sa = sin a) about y. False when ps^2 <= 0.
__device__ __forceinline__ bool edge_rot_xz(double *v, double pp, double ca, double sa)
{
    const double x = v[0], px = v[1], py = v[3];
    double ps = pp * pp - px * px - py * py;
    if (ps <= 0) return false;
    ps = sqrt(ps);
    const double den = ps * ca - px * sa;
    const double L = x * sa / den;
    v[0] = x * ps / den;
    v[1] = px * ca + ps * sa;
    v[2] = v[2] + py * L;
    v[4] = v[4] - pp * L;
    return true;
}

// fringe_mod.f90:1588-1690 bmad_wedger with g_tot /= 0: a wedge of angle a (ca, sa, s2a =
// sin 2a) through the dipole field b1 = g_tot. False on its error exits.
__device__ __forceinline__ bool edge_wedger(double *v, double pp, double a, double ca, double sa,
                                            double s2a, double b1)
{
    const double x = v[0], px = v[1], py = v[3];
    double ps = pp * pp - px * px - py * py;
    if (ps <= 0) return false;
    ps = sqrt(ps);
    double pt = pp * pp - py * py;
    if (pt < 1e-10) return false;
    pt = sqrt(pt);
    const double s2 = sa * sa;
    const double pxn = px * ca + (ps - b1 * x) * sa;
    double psn = pt * pt - pxn * pxn;
    if (psn < 1e-10) return false;
    psn = sqrt(psn);
    const double den = psn + ps * ca - px * sa;
    const double num = x * px * s2a + s2 * (2 * x * ps - b1 * x * x);
    const double L = (a + asin(px / pt) - asin(pxn / pt)) / b1;
    v[0] = x * ca + num / den;
    v[1] = pxn;
    v[2] = v[2] + py * L;
    v[4] = v[4] - pp * L;
    return true;
}

// fringe_mod.f90:1797-1933 bmad_fringe_dipoler: the hard-edge dipole fringe with b = +-g_tot,
// fint hgap = fh, inv_b02 = 1/beta_ref^2. False on its error exits.
__device__ __forceinline__ bool edge_dipoler(double *v, double pp, double b, double fh,
                                             double inv_b02)
{
    const double x = v[0], px = v[1], y = v[2], py = v[3], z = v[4];
    double ps = pp * pp - px * px - py * py;
    if (ps <= 0) return false;
    ps = sqrt(ps);
    const double ee = sqrt(pp * pp + inv_b02 - 1);
    const double beta = pp / ee;
    const double xp = px / ps, yp = py / ps;
    const double kap = b * fh;
    const double fac2 = 1 + yp * yp;
    const double d11 = (1 + xp * xp) / ps, d21 = xp * yp / ps, d31 = -xp;
    const double d12 = xp * yp / ps, d22 = fac2 / ps, d32 = -yp;
    const double d13 = -ee * xp / (ps * ps), d23 = -ee * yp / (ps * ps), d33 = ee / ps;
    const double w2 = 1 + xp * xp * (2 + yp * yp);
    const double fi0 = atan(xp / fac2) - 2 * kap * w2 * ps;
    const double cf = cos(fi0);
    const double co2 = b / (cf * cf);
    const double co1 = co2 / (1 + (xp / fac2) * (xp / fac2));
    const double fi1 = co1 / fac2 - 4 * kap * co2 * xp * (2 + yp * yp) * ps;
    const double fi2 = -2 * co1 * xp * yp / (fac2 * fac2) - 4 * kap * co2 * xp * xp * yp * ps;
    const double fi3 = -2 * kap * co2 * w2;
    const double fi0t = b * tan(fi0);
    const double bb2 = fi1 * d12 + fi2 * d22 + fi3 * d32;
    double sq = 1 - 2 * bb2 * y;
    if (sq < 0) return false;
    sq = sqrt(sq);
    const double ynew = 2 * y / (1 + sq);
    const double bb1 = fi1 * d11 + fi2 * d21 + fi3 * d31;
    const double bb3 = fi1 * d13 + fi2 * d23 + fi3 * d33;
    v[0] = x + 0.5 * bb1 * ynew * ynew;
    v[2] = ynew;
    v[3] = py - fi0t * ynew;
    v[4] = z + 0.5 * beta * bb3 * ynew * ynew;
    return true;
}

// fringe_mod.f90:1474-1562 exact_bend_edge_kick: at the entrance the wedge of the edge angle
// without field, the fringe, the wedge back through the field g_tot; at the exit the reverse.
// On any error lost_pz with the coordinates unchanged.
__device__ __noinline__ void exact_bend_edge(Particle &p, CP q, bool entering)
{
    const double g_tot = q[29];
    const double pp = 1 + p.pz;
    double v[6] = {p.x, p.px, p.y, p.py, p.z, p.pz};
    const CP e = q + (entering ? 56 : 60);  // cos, sin, sin 2a, fint hgap
    const double a = entering ? q[65] : q[66];
    bool ok;
    if (entering) {
        ok = edge_rot_xz(v, pp, e[0], e[1]);  // bmad_wedger without field
        ok = ok && edge_dipoler(v, pp, g_tot, e[3], q[64]);
        ok = ok && (g_tot == 0 ? edge_rot_xz(v, pp, e[0], -e[1]) : edge_wedger(v, pp, -a, e[0], -e[1], -e[2], g_tot));
    } else {
        ok = g_tot == 0 ? edge_rot_xz(v, pp, e[0], -e[1]) : edge_wedger(v, pp, -a, e[0], -e[1], -e[2], g_tot);
        ok = ok && edge_dipoler(v, pp, -g_tot, e[3], q[64]);
        ok = ok && edge_rot_xz(v, pp, e[0], e[1]);
    }
    if (!ok) {
        p.state = ST_LOST_PZ;
        return;
    }
    p.x = v[0];
    p.px = v[1];
    p.y = v[2];
    p.py = v[3];
    p.z = v[4];
}

// The bend's edge at one end (apply_element_edge_kick.f90:64-189,270-271 with bend_edge_kick,
// fringe_mod.f90:40-128): the electric edge before the magnetic one at the exit and after it at
// the entrance; the magnetic edge is the hard multipole edge then exact_bend_edge_kick (fringe
// full; the hard multipole edge after it at the exit) or hwang_bend_edge_kick.
__device__ __forceinline__ void bend_general_edge(Particle &p, CP q, int32_t flags, int32_t emask,
                                                  bool entering, const Species &sp)
{
    if (!entering && emask) elec_fringe(p, q + 47, q + 51, emask, -1.0, sp);
    if (flags & (entering ? F_EDGE_ENT : F_EDGE_EXIT)) {
        if (flags & F_BEXACT) {
            if (entering) quad_hard_edge(p, q[55], true);
            exact_bend_edge(p, q, entering);
            if (!entering) quad_hard_edge(p, q[55], false);
        } else {
            hwang_edge(p, q + (entering ? 10 : 16), q[entering ? 25 : 26], q[29], entering,
                       flags & F_EDGE_SIMPLE);
        }
    }
    if (entering && emask) elec_fringe(p, q + 47, q + 51, emask, 1.0, sp);
}

// The bend's multipole kicks (track_a_bend.f90:336-467, not exact): the magnetic kick
// coef r_step (1 + g x) (kx, ky) of the multipoles, then the electric one, step_len coef
// (1 + g x) charge / (ps beta p0c) E with ps = sqrt(1+pz)^2 - px^2 - py^2) / (1+pz), which
// changes the energy; factor = 2 coef (the table holds the half-step strengths).
__device__ __forceinline__ void bend_general_kicks(Particle &p, CP q, int32_t mmask, int32_t emask,
                                                   double factor, const Species &sp)
{
    const double g = q[1];
    double kx, ky;
    if (mmask) {
        multipole_sum(p.x, p.y, q + 31, q + 35, mmask, kx, ky);
        const double f = factor * fma(g, p.x, 1.0);
        p.px = fma(f, kx, p.px);
        p.py = fma(f, ky, p.py);
    }
    if (emask) {
        multipole_sum(p.x, p.y, q + 39, q + 43, emask, kx, ky);
        const double rel_p0 = p.rel_p, px0 = p.px, py0 = p.py;
        const double ps2 = rel_p0 * rel_p0 - px0 * px0 - py0 * py0;
        const double ps = sqrt(ps2) / rel_p0;
        const double f = factor * fma(g, p.x, 1.0) / (ps * p.beta);
        kx = kx * f;
        ky = ky * f;
        const double alpha = (kx * (2 * px0 + kx) + ky * (2 * py0 + ky)) / (rel_p0 * rel_p0);
        p.px = px0 + kx;
        p.py = py0 + ky;
        p.pz = p.pz + rel_p0 * sqrt_one(alpha);
        const double beta_new =
            (1 + p.pz) / sqrt(ps2 + p.px * p.px + p.py * p.py + sqr(sp.mc2 / sp.p0c));
        p.z = p.z * beta_new / p.beta;
        p.beta = beta_new;
        refresh_energy(p, sp);
    }
}

// track_a_bend.f90:500-659 sbend_body_with_k1_map: one step of a bend with a gradient k1 = b1/L
// (q[30]), g and g_tot = g + dg of the body block; then the low-energy z correction.
__device__ __noinline__ void sbend_k1_step(Particle &p, CP q)
{
    const double length = q[0], g = q[1], g_tot = q[2], k1 = q[30];
    const double rel_p = p.rel_p, rel_p2 = rel_p * rel_p;
    const double k_x = k1 + g * g_tot;
    const double x_c = (g * rel_p - g_tot) / k_x;
    const double om_x = sqrt(fabs(k_x) / rel_p);
    const double om_y = sqrt(fabs(k1) / rel_p);
    const double tau_x = -copysign(1.0, k_x);
    const double tau_y = copysign(1.0, k1);
    double s_x, c_x, z2, s_y, c_y;
    double arg = om_x * length;
    if (arg < 1e-6) {
        s_x = (1 + tau_x * arg * arg / 6) * length;
        c_x = 1 + tau_x * arg * arg / 2;
        z2 = g * length * length / (2 * rel_p);
    } else {
        if (k_x > 0) {
            s_x = sin(arg) / om_x;
            c_x = cos(arg);
        } else {
            s_x = sinh(arg) / om_x;
            c_x = cosh(arg);
        }
        z2 = tau_x * g * (1 - c_x) / (rel_p * om_x * om_x);
    }
    arg = om_y * length;
    if (arg < 1e-6) {
        s_y = (1 + tau_y * arg * arg / 6) * length;
        c_y = 1 + tau_y * arg * arg / 2;
    } else if (k1 < 0) {
        s_y = sin(arg) / om_y;
        c_y = cos(arg);
    } else {
        s_y = sinh(arg) / om_y;
        c_y = cosh(arg);
    }
    const double r0 = p.x - x_c, r1 = p.px, r2 = p.y, r3 = p.py;
    const double ox2 = om_x * om_x, oy2 = om_y * om_y;
    const double z0 = -g * x_c * length;
    const double z1 = -g * s_x;
    const double z11 = tau_x * ox2 * (length - c_x * s_x) / 4;
    const double z12 = -tau_x * ox2 * s_x * s_x / (2 * rel_p);
    const double z22 = -(length + c_x * s_x) / (4 * rel_p2);
    const double z33 = tau_y * oy2 * (length - c_y * s_y) / 4;
    const double z34 = -tau_y * oy2 * s_y * s_y / (2 * rel_p);
    const double z44 = -(length + c_y * s_y) / (4 * rel_p2);
    p.x = c_x * r0 + s_x * r1 / rel_p + x_c;
    p.px = tau_x * ox2 * rel_p * s_x * r0 + c_x * r1;
    p.y = c_y * r2 + s_y * r3 / rel_p;
    p.py = tau_y * oy2 * rel_p * s_y * r2 + c_y * r3;
    p.z = p.z + z0 + z1 * r0 + z2 * r1 + z11 * r0 * r0 + z12 * r0 * r1 + z22 * r1 * r1 +
          z33 * r2 * r2 + z34 * r2 * r3 + z44 * r3 * r3;
    p.z = p.z + length * p.lez;
}

// orbit_too_large inside the general bend, where pz may have changed (electric fields).
__device__ __forceinline__ bool bend_orbit_flagged(Particle &p)
{
    if (any_lane(orbit_bad(p))) {
        if (orbit_bad(p)) {
            orbit_too_large(p);
            return true;
        }
    }
    return false;
}

// bmad/low_level/track_a_bend.f90:18-333, the general sector bend: the frame change of
// offset_particle's bend branch (or the ref_tilt rotation alone), the entrance edge,
// orbit_too_large, half kicks, n_step body steps (the k1 map, a drift for g = dg = 0, else the
// semi-linear or exact body of K_BEND with g_tot = g + dg) each followed by a kick (half after
// the last), orbit_too_large, the exit edge, the frame change back. A particle the body loses,
// or orbit_too_large flags, leaves without the exit edge and frame change, as in the reference.
template <bool F>
__device__ __forceinline__ void bend_general(Particle &p, CP q, int32_t n_step, int32_t flags,
                                             const Species &sp)
{
    const int32_t mmask = (flags >> F_ORDER_SHIFT) & 0xF;
    const int32_t emask = (flags >> F_ELEC_SHIFT) & 0xF;
    if (flags & F_MISALIGN) {
        double s_body;
        offset_frame(p, q + 67, p.x - q[76], p.y - q[77], 0.0 - q[78], 0.0, 0.0, s_body);
        const double ds = 0.0 - s_body;
        if (p.state == ST_ALIVE && fabs(ds) > 1e-14) drift_edge(p, ds);
    } else if (flags & F_TILT) {
        tilt_rotate(p, q[91], q[92]);
    }
    if (flags & F_QEDGE_ENT) bend_general_edge(p, q, flags, emask, true, sp);
    if (bend_orbit_flagged(p)) return;
    if (mmask | emask) bend_general_kicks(p, q, mmask, emask, 1.0, sp);
    for (int i = 1; i <= n_step; i++) {
        if (flags & F_BK1) {
            sbend_k1_step(p, q);
        } else if (flags & F_BEND_DRIFT) {
            if (q[0] != 0) drift_exact(p, q[0]);
        } else {
            const int32_t st0 = p.state;
            const double g = q[1];
            bool semi = false;
            if (flags & F_BSEMI) {
                semi = (fabs(p.x * g) < 1e-9) & (fabs(p.px) < 1e-9) & (fabs(p.py) < 1e-9) &
                       (fabs(p.pz) < 1e-9);
                if (semi) sbend_semilinear(p, q, sp);
            }
            if (!semi) {
                if (F && (flags & F_BEND_SERIES)) {
                    const bool done = sbend_momentum_form(p, q);
                    if (!done) sbend_exact_fast(p, q);
                } else if (F) {
                    sbend_exact_fast(p, q);
                } else {
                    sbend_exact_ref(p, q);
                }
            }
            if (p.state != st0) return;
        }
        if (mmask | emask) bend_general_kicks(p, q, mmask, emask, i == n_step ? 1.0 : 2.0, sp);
    }
    if (bend_orbit_flagged(p)) return;
    if (flags & F_QEDGE_EXIT) bend_general_edge(p, q, flags, emask, false, sp);
    if (flags & F_MISALIGN) {
        double s_lab;
        offset_frame(p, q + 79, p.x, p.y, 0.0, q[88], q[89], s_lab);
        s_lab = s_lab + q[90];
        const double ds = q[93] - s_lab;
        if (p.state == ST_ALIVE && fabs(ds) > 1e-14) drift_edge(p, ds);
    } else if (flags & F_TILT) {
        tilt_rotate(p, q[91], -q[92]);
    }
}

// Thin multipole (multipole$, ab_multipole$; track1_bmad.f90:142-168 with multipole_kicks,
// multipole_mod.f90:32-51, 203-286): the kicks of orders 0 .. n-1 summed over (x + i y)^k as
// a/b strengths (the order tilts folded in on the host), and for a multipole$ the dipole order in
// the reference-orbit form (:250-254): px += k0l cos t0 pz, py += k0l sin t0 pz, z -= k0l (cos t0
// x + sin t0 y). The loop over the orders is wave-uniform.
__device__ __forceinline__ void thin_multipole(Particle &p, CP m, int32_t n, int32_t flags)
{
    const double x = p.x, y = p.y;
    double kx = 0, ky = 0, zr = 1, zi = 0;
    for (int32_t k = 0; k < n; k++) {
        k = __builtin_amdgcn_readfirstlane(k);
        const double a = m[2 + k], b = m[2 + n + k];
        kx += -b * zr + a * zi;
        ky += a * zr + b * zi;
        const double nr = zr * x - zi * y;
        zi = zr * y + zi * x;
        zr = nr;
    }
    if (flags & F_THIN_REF0) {
        const double c0 = m[0], s0 = m[1];
        p.px = p.px + c0 * p.pz;
        p.py = p.py + s0 * p.pz;
        p.z = p.z - (c0 * x + s0 * y);
    }
    p.px = p.px + kx;
    p.py = p.py + ky;
}

// ---- radiation ----------------------------------------------------------------------------

// Parameter count of an element kind (kParams; K_THIN, which has no radiation block, at its
// minimum).
__device__ __forceinline__ constexpr int param_len(int kind)
{
    return kind == K_DRIFT ? 1 : kind == K_QUAD ? 5 : kind == K_MULT ? 9 : kind == K_BEND ? 29
         : kind == K_RF ? 7 : kind == K_MARKER ? 0 : kind == K_QMULT ? 32 : kind == K_BMULT ? 94 : 2;
}
static_assert(param_len(K_BMULT) == kParams[K_BMULT] && param_len(K_QMULT) == kParams[K_QMULT] &&
              param_len(K_BEND) == kParams[K_BEND] && param_len(K_RF) == kParams[K_RF] &&
              param_len(K_MULT) == kParams[K_MULT] && param_len(K_QUAD) == kParams[K_QUAD] &&
              param_len(K_DRIFT) == kParams[K_DRIFT] && param_len(K_MARKER) == kParams[K_MARKER],
              "param_len disagrees with kParams");

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC11): 10 rounds of two 32x32 -> 64-bit products on a 128-bit counter under a 64-bit key.
struct Ctr4 {
    uint32_t c0, c1, c2, c3;
};
__device__ __forceinline__ Ctr4 philox4x32_10(Ctr4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.c2;
        c = Ctr4{(uint32_t)(p1 >> 32) ^ c.c1 ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.c3 ^ k1,
                 (uint32_t)p0};
    }
    return c;
}

// A uniform double in (0, 1] from 53 bits of two words (the sum rounds to nearest even).
__device__ __forceinline__ double uniform53(uint32_t hi, uint32_t lo)
{
    const uint64_t m = ((uint64_t)hi << 21) | (lo >> 11);
    return ((double)m + 0.5) * 0x1p-53;
}

// Six standard normal deviates for one particle at one edge of one element (ran_gauss's role in
// track1_radiation, radiation_mod.f90:116): three Philox blocks with the counter (particle,
// turn, element, 4 edge + block), each giving two deviates by the Box-Muller transform
// (g = sqrt(-2 ln u1) (cos, sin)(2 pi u2)).
__device__ __forceinline__ void gauss6(double *g, uint32_t id, int32_t turn, int32_t ix1, int32_t edge,
                                    uint32_t k0, uint32_t k1)
{
    for (int b = 0; b < 3; b++) {
        const Ctr4 r = philox4x32_10(Ctr4{id, (uint32_t)turn, (uint32_t)ix1, (uint32_t)(4 * edge + b)},
                                     k0, k1);
        const double u1 = uniform53(r.c0, r.c1), u2 = uniform53(r.c2, r.c3);
        const double rad = sqrt(-2 * log(u1));
        double sn, cs;
        sincos(kTwoPi * u2, &sn, &cs);
        g[2 * b] = rad * cs;
        g[2 * b + 1] = rad * sn;
    }
}

// The phase-space coordinates of a particle, passed to and returned from the radiation kick
// by value.
struct Vec6 {
    double x, px, y, py, z, pz;
};

// track1_radiation (radiation_mod.f90:61-147) at one edge (0 entrance, 1 exit) with the
// element's map m (radiation block layout in bmx_table.hpp): damping
//   vec += rr damp_dmat (vec - ref_orb) [+ rr xfer_damp_vec unless zero_average],
// then fluctuations vec += scale stoc_mat g, with rr = scale = synch_rad_scale (time_dir = 1).
// The caller tests pz < -1 (lost_pz) and refreshes the functions of the energy; beta is left as
// it was, as the reference leaves it. The settings come as values: flags and scale of RadCfg,
// key the fluctuation key; id is the particle's index.
__device__ __forceinline__ Vec6 rad_kick(Vec6 in, CP mc, uint32_t flags, double scale, uint32_t key0,
                                      uint32_t key1, uint32_t id, int32_t edge, int32_t ix1,
                                      int32_t turn)
{
    // The 69 map values are read with vector loads from one address on every lane (a pointer
    // held in VGPRs), so they take vector registers as they are used rather than about 140
    // scalar registers at once, which the element loop cannot spare.
    const double *m = (const double *)mc;
    asm volatile("" : "+v"(m));
    double v[6] = {in.x, in.px, in.y, in.py, in.z, in.pz};
    if (flags & RAD_DAMP) {
        const double rr = scale;
        double d[6], w[6];
        for (int j = 0; j < 6; j++) d[j] = v[j] - m[j];
        for (int i = 0; i < 6; i++) {
            double s = 0;
            for (int j = 0; j < 6; j++) s = s + m[6 + 6 * i + j] * d[j];
            w[i] = v[i] + rr * s;
        }
        for (int i = 0; i < 6; i++) v[i] = w[i];
        if (!(flags & RAD_ZERO_AVERAGE))
            for (int i = 0; i < 6; i++) v[i] = v[i] + rr * m[42 + i];
    }
    if (flags & RAD_FLUCT) {
        double g[6];
        gauss6(g, id, turn, ix1, edge, key0, key1);
        const double *L = m + 48;
        int k = 0;
        for (int i = 0; i < 6; i++) {
            double s = 0;
            for (int j = 0; j <= i; j++) s = s + L[k++] * g[j];
            v[i] = v[i] + scale * s;
        }
    }
    return Vec6{v[0], v[1], v[2], v[3], v[4], v[5]};
}

// The kick at one edge applied to p: the lost_pz test (radiation_mod.f90:122-125) and the
// functions of the energy refreshed.
__device__ __forceinline__ void rad_apply(Particle &p, CP m, const Species &sp, int32_t edge,
                                          int32_t ix1, int32_t turn)
{
    const Vec6 o = rad_kick(Vec6{p.x, p.px, p.y, p.py, p.z, p.pz}, m, sp.rad.flags, sp.rad.scale,
                            sp.rad.key0, sp.rad.key1, p.id, edge, ix1, turn);
    p.x = o.x;
    p.px = o.px;
    p.y = o.y;
    p.py = o.py;
    p.z = o.z;
    p.pz = o.pz;
    if (p.pz < -1) p.state = ST_LOST_PZ;
    refresh_energy(p, sp);
}

template <bool F, int MAPS = MAPS_ALL>
__device__ __forceinline__ bool track_element(Particle &p, EleRef h, CP prm, const Species &sp,
                                              bool from_prev, int32_t ix1, int32_t turn)
{
    const int32_t kind = h.kind, n = h.n, off = h.off;
    // The common element path (MAPS_NO_RF) never meets offsets or momentum windows: the loops
    // send those elements to the rare path, so here their bits are known to be clear and every
    // test and table offset that depends on them folds away.
    const int32_t flags = MAPS != MAPS_NO_RF ? h.flags : (h.flags & ~(F_OFFSET | F_APE_MOMENTUM | F_RAD));
    CP ob = prm + off;
    CP ap = (flags & F_OFFSET) ? ob + kOffsetLen : ob;
    CP q = (flags & (F_APE_ENT | F_APE_EXIT | F_APE_ENT_SAME)) ? ap + 4 : ap;
    if (flags & F_APE_MOMENTUM) q += 8;
    const int32_t aflags = flags;
    if ((flags & F_APE_ENT) || (!from_prev && (flags & F_APE_ENT_SAME))) {
        const bool out = aperture_fails(p, aflags, ap);
        if (any_lane(out)) {
            if (out) {
                aperture(p, aflags, ap);
                p.loss_ele = uniform_here(ix1);
                p.loss_turn = uniform_here(turn);
                return true;
            }
        }
    }
    constexpr bool kOffsets = MAPS != MAPS_NO_RF;
    // the radiation block follows the parameters (track1.f90:199-202: the upstream half's kick
    // after the entrance aperture test, in laboratory coordinates; a particle it loses is lost
    // at the entrance). Only the radiation map sets carry the kicks (a lattice with radiation
    // runs on them), so the kernels of the other sets keep their register budget.
    constexpr bool kRad = maps_rad(MAPS);
    CP rb = q + param_len(kind);
    if constexpr (kOffsets) {
        if (kRad && (flags & F_RAD)) {
            rad_apply(p, rb, sp, 0, ix1, turn);
            if (p.state != ST_ALIVE) {
                p.loss_ele = uniform_here(ix1);
                p.loss_turn = uniform_here(turn);
                return true;
            }
        }
        if (flags & F_OFFSET) offset_set(p, ob, flags);
    }
    bool unset = true;
    switch (kind) {
    case K_DRIFT: drift<F>(p, q[0]); break;
    case K_QUAD: quadrupole<F>(p, q, n, sp); break;
    case K_MULT: thick_multipole<F>(p, q, n, flags); break;
    case K_BEND:
        if constexpr (!maps_rare(MAPS) || maps_rad(MAPS)) {
            if ((flags & (F_EDGE_ENT | F_EDGE_EXIT | F_EDGE_SIMPLE | F_BEND_DRIFT)) ==
                (F_EDGE_ENT | F_EDGE_EXIT | F_EDGE_SIMPLE))
                sbend<F, true>(p, flags, q, sp);
            else
                sbend<F, false>(p, flags, q, sp);
        }
        break;
    case K_RF:
        if constexpr (MAPS != MAPS_NO_RF) rfcavity<F>(p, q, n, sp);
        break;
    case K_QMULT:
        if constexpr (maps_q(MAPS)) unset = quad_multipole<F>(p, q, n, flags, sp);
        break;
    case K_BMULT:
        if constexpr (maps_q(MAPS)) bend_general<F>(p, q, n, flags, sp);
        break;
    case K_THIN:
        if constexpr (MAPS != MAPS_NO_RF) thin_multipole(p, q, n, flags);
        break;
    default: break;
    }
    if constexpr (kOffsets) {
        // the rf cavity returns before its exit offset when it loses the particle (pz = -1),
        // the quadrupole when an electric edge loses it
        if (flags & F_OFFSET) {
            if ((kind != K_RF || p.state == ST_ALIVE) && unset) offset_unset(p, ob, flags);
        }
    }
    if constexpr (kRad) {
        if (flags & F_RAD) {
            // track1.f90:291-326: orbit_too_large, the state test, the downstream half's kick,
            // then check_aperture_limit, whose own orbit_too_large (check_aperture_limit.f90:61)
            // now tests the kicked coordinates and overrides the kick's lost_pz
            bool done = p.state != ST_ALIVE;
            if (orbit_bad(p)) {
                orbit_too_large(p);
                done = true;
            }
            if (!done) {
                rad_apply(p, rb + kRadEdge, sp, 1, ix1, turn);
                if (orbit_bad(p)) orbit_too_large(p);
                else if ((flags & F_APE_EXIT) && aperture_fails(p, aflags, ap)) aperture(p, aflags, ap);
            }
            if (p.state != ST_ALIVE) {
                p.loss_ele = uniform_here(ix1);
                p.loss_turn = uniform_here(turn);
            }
            return false;
        }
    }
    // orbit_too_large (overwrites any state), else the exit aperture for live particles; both
    // tests first without branches, the state codes and the loss record only on the (rare)
    // lanes that are lost here or were lost inside the map
    // The gate leaves out what cannot newly hold here: rel_p < 0 outside rf cavities (pz only
    // changes in them, and a particle enters with rel_p >= 0), px^2 + py^2 > (1 + pz)^2 after a
    // drift or marker (neither changes px, py or pz, and the particle entered without it), and
    // |x| or |y| above the limit when the exit aperture fails for those anyway
    // (F_APE_EXIT_COVERS).
    LaneMask hit_lanes = lanes(p.state != ST_ALIVE);
    if (kind != K_DRIFT && kind != K_MARKER) {
        const double f_unstable = p.px * p.px + p.py * p.py - p.rel_p * p.rel_p;
        hit_lanes |= lanes(f_unstable > 0);
    }
    if constexpr (MAPS != MAPS_NO_RF) hit_lanes |= lanes(p.rel_p < 0);
    if (!(flags & F_APE_EXIT_COVERS))
        hit_lanes |= lanes(fabs(p.x) > kMaxAperture) | lanes(fabs(p.y) > kMaxAperture);
    if (flags & F_APE_EXIT) hit_lanes |= aperture_fail_lanes(p, aflags, ap);
    if (hit_lanes) {
        const bool bad = orbit_bad(p);
        const bool alive = p.state == ST_ALIVE;
        const bool ape = (flags & F_APE_EXIT) ? aperture_fails(p, aflags, ap) : false;
        if (bad | ape | !alive) {
            if (bad) orbit_too_large(p);
            else if (alive) aperture(p, aflags, ap);
            p.loss_ele = uniform_here(ix1);
            p.loss_turn = uniform_here(turn);
        }
    }
    return false;
}

}  // namespace bmx
