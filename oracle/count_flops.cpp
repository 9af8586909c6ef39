// Algorithmic FP64 operation counter for the CPU restatement (TEST INFRASTRUCTURE).
//
// bmad_oracle.c is compiled a second time as C++ with `double` replaced by `cdouble`, a double
// that carries a "depends on the particle" flag. An arithmetic operation is counted only when
// an operand depends on the particle, so per-element constants that the reference recomputes
// per particle are excluded automatically. Convention: add/sub/mul/div = 1, sqrt = 1, each libm
// transcendental (sin, cos, asin, atan2, sinh, cosh, fmod) = 1; sign changes, fabs, comparisons,
// multiplications by an exact constant 1 and operations with an exact constant 0 are free, and so
// are validity tests that repeat one made on the same coordinates (BMO_REPEATED_TEST). The rf phase is computed in quad
// precision by the reference; its 4 operations are counted at the conversion back to double.
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/bmx.h"

namespace cnt {
thread_local long long g_ops = 0;

struct cdouble {
    double v;
    bool dyn;
    cdouble() = default;
    cdouble(double x) : v(x), dyn(false) {}
    cdouble(int x) : v(x), dyn(false) {}
    cdouble(long x) : v((double)x), dyn(false) {}
    cdouble(unsigned x) : v(x), dyn(false) {}
    cdouble(__float128 x) : v((double)x), dyn(true) { g_ops += 4; }
    cdouble(double x, bool d) : v(x), dyn(d) {}
    explicit operator __float128() const { return (__float128)v; }
    explicit operator double() const { return v; }
};

inline bool one(const cdouble &a) { return !a.dyn && a.v == 1.0; }
// An exact constant zero: an operation with it as an operand does no work (x + 0 = x, x * 0 = 0,
// 0 / x = 0), so it is not counted and a product or quotient with it is a constant, which makes
// every operation on that term free as well (the general Hwang edge kick with e = 0, for example,
// is the simple one).
inline bool zero(const cdouble &a) { return !a.dyn && a.v == 0.0; }
inline cdouble op(double r, const cdouble &a, const cdouble &b)
{
    bool d = a.dyn || b.dyn;
    if (d) g_ops++;
    return cdouble(r, d);
}
inline cdouble operator+(const cdouble &a, const cdouble &b)
{
    if (zero(a)) return cdouble(a.v + b.v, b.dyn);
    if (zero(b)) return cdouble(a.v + b.v, a.dyn);
    return op(a.v + b.v, a, b);
}
inline cdouble operator-(const cdouble &a, const cdouble &b)
{
    if (zero(a)) return cdouble(a.v - b.v, b.dyn);  // a sign change
    if (zero(b)) return cdouble(a.v - b.v, a.dyn);
    return op(a.v - b.v, a, b);
}
inline cdouble operator*(const cdouble &a, const cdouble &b)
{
    if (zero(a) || zero(b)) return cdouble(a.v * b.v, false);
    if (one(a)) return b;
    if (one(b)) return a;
    return op(a.v * b.v, a, b);
}
inline cdouble operator/(const cdouble &a, const cdouble &b)
{
    if (zero(a)) return cdouble(a.v / b.v, false);
    if (one(b)) return a;
    return op(a.v / b.v, a, b);
}
inline cdouble operator-(const cdouble &a) { return cdouble(-a.v, a.dyn); }
inline cdouble &operator+=(cdouble &a, const cdouble &b) { return a = a + b; }
inline cdouble &operator-=(cdouble &a, const cdouble &b) { return a = a - b; }
inline cdouble &operator*=(cdouble &a, const cdouble &b) { return a = a * b; }
inline bool operator<(const cdouble &a, const cdouble &b) { return a.v < b.v; }
inline bool operator>(const cdouble &a, const cdouble &b) { return a.v > b.v; }
inline bool operator<=(const cdouble &a, const cdouble &b) { return a.v <= b.v; }
inline bool operator>=(const cdouble &a, const cdouble &b) { return a.v >= b.v; }
inline bool operator==(const cdouble &a, const cdouble &b) { return a.v == b.v; }
inline bool operator!=(const cdouble &a, const cdouble &b) { return a.v != b.v; }

inline cdouble fn(double r, const cdouble &a)
{
    if (a.dyn) g_ops++;
    return cdouble(r, a.dyn);
}
inline cdouble sqrt(const cdouble &a) { return fn(::sqrt(a.v), a); }
inline cdouble sin(const cdouble &a) { return fn(::sin(a.v), a); }
inline cdouble cos(const cdouble &a) { return fn(::cos(a.v), a); }
inline cdouble asin(const cdouble &a) { return fn(::asin(a.v), a); }
inline cdouble sinh(const cdouble &a) { return fn(::sinh(a.v), a); }
inline cdouble cosh(const cdouble &a) { return fn(::cosh(a.v), a); }
inline cdouble exp(const cdouble &a) { return fn(::exp(a.v), a); }
inline cdouble log(const cdouble &a) { return fn(::log(a.v), a); }
inline cdouble atan(const cdouble &a) { return fn(::atan(a.v), a); }
inline cdouble tan(const cdouble &a) { return fn(::tan(a.v), a); }
inline cdouble copysign(double m, const cdouble &s) { return cdouble(::copysign(m, s.v), false); }
inline cdouble fabs(const cdouble &a) { return cdouble(::fabs(a.v), a.dyn); }
inline cdouble atan2(const cdouble &a, const cdouble &b) { return op(::atan2(a.v, b.v), a, b); }
inline cdouble fmod(const cdouble &a, const cdouble &b) { return op(::fmod(a.v, b.v), a, b); }
inline long lround(const cdouble &a) { return ::lround(a.v); }

// Runs f without counting its operations: validity tests that repeat one made just before on
// the same coordinates, which the GPU kernels do not repeat (BMO_REPEATED_TEST in the source).
template <class Fn>
inline auto uncounted(Fn f)
{
    const long long saved = g_ops;
    auto r = f();
    g_ops = saved;
    return r;
}
}  // namespace cnt

#define BMO_REPEATED_TEST(expr) (cnt::uncounted([&]() { return (expr); }))

using cnt::cdouble;
using cnt::sqrt;
using cnt::sin;
using cnt::cos;
using cnt::asin;
using cnt::sinh;
using cnt::cosh;
using cnt::log;
using cnt::fabs;
using cnt::atan2;
using cnt::fmod;
using cnt::lround;

#define BMAD_ORACLE_H 1
#define double cdouble
#include "bmad_oracle.c"
#undef double

extern "C" {

// One particle through one element; returns the operation count and advances the particle.
double bmo_count_track1(const bmx_ele *ele, double mc2, int32_t charge, double p0c, double *vec,
                        double *beta, int32_t *state)
{
    cdouble v[6];
    for (int i = 0; i < 6; i++) v[i] = cdouble(vec[i], true);
    cdouble b(*beta, true);
    cnt::g_ops = 0;
    int32_t err = bmo_track1(ele, cdouble(mc2), charge, cdouble(p0c), v, &b, state);
    for (int i = 0; i < 6; i++) vec[i] = v[i].v;
    *beta = b.v;
    return err ? -1.0 : (double)cnt::g_ops;
}

}  // extern "C"
