// axs_atan2.hpp -- atan2(y, x) evaluated in double-double arithmetic and
// rounded once to double: the correctly rounded angle except for inputs
// whose exact angle lies within about 2^-100 (relative) of a rounding
// boundary.  The same source compiles for the host (g++) and the device
// (hipcc); both evaluate the same sequence of IEEE additions,
// multiplications, divisions and fused multiply-adds (no contraction:
// -ffp-contract=off), so the two builds give the same bits.
//
// The reference's axion field is libm's atan2(phi1, phi2) (s_run.cpp:92-94),
// and the string finder compares sums of its scaled values against a
// threshold (string_finding.cpp:17-25, 49-78): a last-ulp difference between
// two atan2 implementations can flip a flag.  A correctly rounded result is
// the value every accurate libm returns, so the device decides the margin
// cases with this function.
//
// Method (double-double after Dekker 1971, with the TwoSum / FMA TwoProduct
// error-free transforms and the "accurate" double-double addition of Hida,
// Li and Bailey's QD library; table-driven reduction of the arctangent):
//   1. C99 Annex F special values (zeros of either sign, infinities, NaN).
//   2. u = min(|y|, |x|) / max(|y|, |x|) in [0, 1] as a double-double (the
//      quotient and its exact FMA remainder), after scaling both inputs by
//      the same power of two; a ratio below 2^-900 returns the rounded
//      quotient (atan(u) = u to far below an ulp there).
//   3. atan(u) = atan(c) + atan(t), c = k/256 nearest u, t = (u - c) /
//      (1 + u c), |t| <= 2^-9; atan(c) from a 257-entry double-double
//      table (mpmath, axs_atan2_table.hpp), atan(t) = t + t z Q(z), z = t^2,
//      Q = -1/3 + z/5 + z^2 R(z) with -1/3 and z/5 in double-double and
//      R(z) = -1/7 + z/9 - z^2/11 + ... + z^5/17 in double (the omitted
//      terms are below 2^-112 of t).
//   4. The octant: pi/2 - atan(u) where |y| > |x|, pi - that where x < 0,
//      then the sign of y; no step cancels more than one bit.
#pragma once

#include <math.h>

#include "axs_atan2_table.hpp"

#if defined(__HIPCC__)
#define AXS_HD __host__ __device__ inline
#else
#define AXS_HD inline
#endif

namespace axs {
namespace ddm {

struct dd {
    double hi, lo;
};

// s + e = a + b exactly (any magnitudes)
AXS_HD dd two_sum(double a, double b)
{
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return dd{s, e};
}

// s + e = a + b exactly, |a| >= |b| (or a == 0)
AXS_HD dd fast_two_sum(double a, double b)
{
    double s = a + b;
    double e = b - (s - a);
    return dd{s, e};
}

// p + e = a * b exactly (no overflow or underflow)
AXS_HD dd two_prod(double a, double b)
{
    double p = a * b;
    double e = fma(a, b, -p);
    return dd{p, e};
}

AXS_HD dd add(dd a, dd b)
{
    dd s = two_sum(a.hi, b.hi);
    dd t = two_sum(a.lo, b.lo);
    s.lo = s.lo + t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo = s.lo + t.lo;
    return fast_two_sum(s.hi, s.lo);
}

AXS_HD dd neg(dd a) { return dd{-a.hi, -a.lo}; }

AXS_HD dd mul(dd a, dd b)
{
    dd p = two_prod(a.hi, b.hi);
    double c = a.hi * b.lo;
    c = c + a.lo * b.hi;
    p.lo = p.lo + c;
    return fast_two_sum(p.hi, p.lo);
}

AXS_HD dd mul_d(dd a, double b)
{
    dd p = two_prod(a.hi, b);
    p.lo = p.lo + a.lo * b;
    return fast_two_sum(p.hi, p.lo);
}

// a / b for b in [1, 2]: three quotient digits
AXS_HD dd div(dd a, dd b)
{
    double q1 = a.hi / b.hi;
    dd r = add(a, neg(mul_d(b, q1)));
    double q2 = r.hi / b.hi;
    r = add(r, neg(mul_d(b, q2)));
    double q3 = r.hi / b.hi;
    dd q = fast_two_sum(q1, q2);
    return add(q, dd{q3, 0.0});
}

constexpr dd kPi{0x1.921fb54442d18p+1, 0x1.1a62633145c07p-53};
constexpr dd kPio2{0x1.921fb54442d18p+0, 0x1.1a62633145c07p-54};
constexpr dd kMinusThird{-0x1.5555555555555p-2, -0x1.5555555555555p-56};
constexpr dd kFifth{0x1.999999999999ap-3, -0x1.999999999999ap-57};

// atan(u), u = (uh, ul) in [0, 1], uh >= 2^-901
AXS_HD dd atan_unit(double uh, double ul)
{
    const int k = (int)(uh * 256.0 + 0.5);
    const double c = (double)k * 0x1p-8;
    // n = u - c: uh - c is exact (Sterbenz for k >= 1; c = 0 for k = 0)
    dd n = two_sum(uh - c, ul);
    // d = 1 + u c
    dd p = two_prod(uh, c);
    dd s = two_sum(1.0, p.hi);
    s.lo = s.lo + (p.lo + ul * c);
    dd d = fast_two_sum(s.hi, s.lo);
    dd t = div(n, d);
    dd z = mul(t, t);
    const double zh = z.hi;
    double r = 1.0 / 17.0;
    r = -1.0 / 15.0 + zh * r;
    r = 1.0 / 13.0 + zh * r;
    r = -1.0 / 11.0 + zh * r;
    r = 1.0 / 9.0 + zh * r;
    r = -1.0 / 7.0 + zh * r;
    const double z2r = (zh * zh) * r;
    dd q = add(kMinusThird, mul(z, kFifth));
    q = add(q, dd{z2r, 0.0});
    dd at = add(t, mul(mul(t, z), q));
    return add(dd{kAtanTable[k][0], kAtanTable[k][1]}, at);
}

}  // namespace ddm

// atan2(y, x), correctly rounded but for the rare boundary cases above
AXS_HD double atan2_dd(double y, double x)
{
    using namespace ddm;
    if (isnan(x) || isnan(y)) return x + y;
    const double ax = fabs(x), ay = fabs(y);
    const bool ny = signbit(y), nx = signbit(x);
    double r;
    if (ay == 0.0) {                       // atan2(+-0, x): +-0 or +-pi
        r = nx ? kPi.hi : 0.0;
        return ny ? -r : r;
    }
    if (ax == 0.0) return ny ? -kPio2.hi : kPio2.hi;
    if (isinf(ax) || isinf(ay)) {
        if (isinf(ax) && isinf(ay))
            r = nx ? 0x1.2d97c7f3321d2p+1 : 0x1.921fb54442d18p-1;    // 3pi/4, pi/4
        else if (isinf(ax))
            r = nx ? kPi.hi : 0.0;
        else
            r = kPio2.hi;
        return ny ? -r : r;
    }
    const bool swap = ay > ax;
    const double num = swap ? ax : ay, den = swap ? ay : ax;
    int e;
    (void)frexp(den, &e);
    const double ns = ldexp(num, -e), ds = ldexp(den, -e);    // ds in [0.5, 1)
    dd a;
    if (ns < 0x1p-900) {
        // atan(u) = u (1 - u^2/3 + ...): the rounded quotient is the rounded
        // angle (a quotient of two doubles is never a rounding midpoint)
        a = dd{num / den, 0.0};
    } else {
        const double q = ns / ds;
        const double rem = fma(-q, ds, ns);                  // exact
        a = atan_unit(q, rem / ds);
    }
    if (swap) a = add(kPio2, neg(a));
    if (nx) a = add(kPi, neg(a));
    r = a.hi + a.lo;
    return ny ? -r : r;
}

// 1/d to within a few units in the last place before the Newton steps of
// the callers: the hardware reciprocal on the device, a division on the host
AXS_HD double recip_seed(double d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcp(d);
#else
    return 1.0 / d;
#endif
}

// |atan2(y, x)|, the magnitude the background sums take (s_run.cpp:145),
// within 4.5e-16 of the exact angle in [0, pi]: the same table reduction as
// atan2_dd in plain double arithmetic, for num = min(|x|, |y|) and den =
// max(|x|, |y|), den in [2^-1000, 2^1000] (abs_atan2_fast scales others).
//   u + ul = num / den from a reciprocal q refined by one Newton step (to
//   about 2^-52) and ul = (num - den u) q from the exact remainder (the
//   pair then errs by about 2^-104); c = k/256 nearest u, t = (u - c + ul)
//   / (1 + u c) (u - c exact, |t| <= 2^-9, the reciprocal of 1 + u c
//   refined once); atan(u) = atan(c) + t - t^3/3 + t^5/5 - t^7/7 (the next term is
//   below 2^-84) with atan(c) the table's double-double entry (`tab`: a
//   kernel passes its copy in LDS, the gather being in the dependency
//   chain); then the octant, pi/2 - that where |y| > |x| (sw) and pi - that
//   where x < 0 (nx), pi/2 + that where both, added to the constant's
//   leading part and then its trailing part.
AXS_HD double abs_atan2_core(double num, double den, bool sw, bool nx, const double (*tab)[2])
{
    using namespace ddm;
    double q = recip_seed(den);
    q = fma(q, fma(-den, q, 1.0), q);
    const double u = num * q;
    const double ul = fma(-den, u, num) * q;        // quotient = u + ul
    const int k = (int)fma(u, 256.0, 0.5);
    const double c = (double)k * 0x1p-8;
    const double dn = fma(u, c, 1.0);
    double p = recip_seed(dn);
    p = fma(p, fma(-dn, p, 1.0), p);
    const double t = ((u - c) + ul) * p;
    const double z = t * t;
    const double s = fma(z, fma(z, -1.0 / 7.0, 1.0 / 5.0), -1.0 / 3.0);
    const double at = tab[k][0] + (tab[k][1] + fma(t * z, s, t));
    const double bh = sw ? kPio2.hi : (nx ? kPi.hi : 0.0);
    const double bl = sw ? kPio2.lo : (nx ? kPi.lo : 0.0);
    return (bh + ((sw != nx) ? -at : at)) + bl;
}

// |atan2(y, x)| where x or y is NaN, den = max(|x|, |y|) is infinite or
// zero (num = min(|x|, |y|)): the C99 Annex F values
AXS_HD double abs_atan2_special(double y, double x, double num, double den, bool sw, bool nx)
{
    using namespace ddm;
    if (isnan(x) || isnan(y)) return x + y;
    if (isinf(den)) {
        if (isinf(num)) return nx ? 0x1.2d97c7f3321d2p+1 : 0x1.921fb54442d18p-1;   // 3pi/4, pi/4
        return sw ? kPio2.hi : (nx ? kPi.hi : 0.0);
    }
    return nx ? kPi.hi : 0.0;     // both zero
}

// |atan2(y, x)| (see abs_atan2_core).  Zeros, infinities and NaN give the
// C99 Annex F values; finite magnitudes outside [2^-1000, 2^1000] are
// scaled by a common power of two first (the angle does not change).
AXS_HD double abs_atan2_fast(double y, double x, const double (*tab)[2] = kAtanTable)
{
    using namespace ddm;
    const double ax = fabs(x), ay = fabs(y);
    const bool sw = ay > ax, nx = signbit(x);
    double num = sw ? ax : ay, den = sw ? ay : ax;
    if (!(ax + ay <= 0x1p+1000) || !(den >= 0x1p-1000)) {
        if (isnan(x) || isnan(y) || isinf(den) || den == 0.0)
            return abs_atan2_special(y, x, num, den, sw, nx);
        int e;
        (void)frexp(den, &e);
        num = ldexp(num, -e);
        den = ldexp(den, -e);
    }
    return abs_atan2_core(num, den, sw, nx, tab);
}

// abs_atan2_fast for arguments that are float values: their nonzero finite
// magnitudes lie in [2^-149, 2^128], inside the core's range, so only
// zeros, infinities and NaN (one class test and one comparison) leave the
// core, for the special values.
AXS_HD double abs_atan2_f32in(double y, double x, const double (*tab)[2] = kAtanTable)
{
    const double ax = fabs(x), ay = fabs(y);
    const bool sw = ay > ax, nx = signbit(x);
    const double num = sw ? ax : ay, den = sw ? ay : ax;
#if defined(__HIP_DEVICE_COMPILE__)
    // den +-0 or +inf (class bits 0x20, 0x40, 0x200), or NaN in either
    const bool special = __builtin_amdgcn_class(den, 0x260) || !(num <= den);
#else
    const bool special = !(den > 0.0) || isinf(den) || !(num <= den);
#endif
    if (special) return abs_atan2_special(y, x, num, den, sw, nx);
    return abs_atan2_core(num, den, sw, nx, tab);
}

}  // namespace axs
