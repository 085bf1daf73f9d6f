// Portable double-precision elementary functions, identical bit for bit on the
// host (gcc/clang, x86-64) and on gfx950.
//
// Why: a free-wake lifting-line run is chaotic (a 1e-12 change of the wake
// geometry grows to metres within a few hundred steps), so trajectory parity
// with the f64 CPU restatement at 1e-10 over thousands of steps holds only if
// every operation rounds identically on both sides.  Basic IEEE operations and
// sqrt already do; the transcendental functions of two different libms do not.
// These functions use only +, -, *, /, sqrt, fma and rint, in a fixed order,
// so both compilers produce the same results.  Accuracy is <= 1 ulp for exp,
// sin, cos, acos and log on the arguments the lifting-line path produces, and
// <= 2 ulp for pow and atan; tests/test_library_host.py measures them against
// the system libm.
//
// Methods (all textbook, coefficients are the exact Taylor coefficients
// rounded to double):
//   exp  — Cody & Waite (1980) additive range reduction x = k ln2 + r with a
//          two-part ln2, Taylor polynomial of degree 13 on |r| <= ln2/2,
//          scaling by 2^k through the exponent field.
//   sin/cos — Cody & Waite reduction by pi/2 with a three-part pi/2 (exact for
//          |x| < 2^20), Taylor polynomials of degree 19 (sin) / 20 (cos) on
//          |r| <= pi/4, quadrant selection.
//   acos — asin Taylor series in z = x^2 on |x| <= 1/2, and the half-angle
//          identities acos(x) = 2 asin(sqrt((1-x)/2)) (x > 1/2),
//          acos(x) = pi - 2 asin(sqrt((1+x)/2)) (x < -1/2).
//   log  — x = 2^k m, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//          s = (m-1)/(m+1), odd series in s, carried in double-double.
//   pow  — x^y = exp(y log x) with log in double-double and the low part of
//          y*log(x) applied as exp(h + l) ~ exp(h) (1 + l).
//   atan — reduction to |u| <= tan(pi/8) (complement and pi/4 shift), odd
//          Taylor series (wind profiles only).
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define SB_HD __host__ __device__ inline
#else
#define SB_HD inline
#endif

namespace sbm {

SB_HD double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
SB_HD double rint_(double x) { return __builtin_rint(x); }
SB_HD double sqrt_(double x) { return __builtin_sqrt(x); }
SB_HD double fabs_(double x) { return __builtin_fabs(x); }

SB_HD uint64_t bits(double x) {
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}
SB_HD double from_bits(uint64_t u) {
    double x;
    memcpy(&x, &u, 8);
    return x;
}
// 2^k for -1022 <= k <= 1023
SB_HD double pow2i(int k) { return from_bits((uint64_t)(k + 1023) << 52); }

constexpr double LN2_HI = 0x1.62e42fefa39efp-1;
constexpr double LN2_LO = 0x1.abc9e3b39803fp-56;
constexpr double INV_LN2 = 0x1.71547652b82fep+0;
constexpr double PIO2_1 = 0x1.921fb54442d18p+0;
constexpr double PIO2_2 = 0x1.1a62633145c07p-54;
constexpr double PIO2_3 = -0x1.f1976b7ed8fbcp-110;
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
constexpr double PI_HI = 0x1.921fb54442d18p+1;
constexpr double PI_LO = 0x1.1a62633145c07p-53;
constexpr double PIO2_HI = 0x1.921fb54442d18p+0;
constexpr double PIO2_LO = 0x1.1a62633145c07p-54;

// e^r - 1 - r for |r| <= ln2/2: r^2 * (1/2 + r/6 + ... + r^11/13!)
SB_HD double expm1_tail(double r) {
    double p = 0x1.6124613a86d09p-33;  // 1/13!
    p = fma_(p, r, 0x1.1eed8eff8d898p-29);
    p = fma_(p, r, 0x1.ae64567f544e4p-26);
    p = fma_(p, r, 0x1.27e4fb7789f5cp-22);
    p = fma_(p, r, 0x1.71de3a556c734p-19);
    p = fma_(p, r, 0x1.a01a01a01a01ap-16);
    p = fma_(p, r, 0x1.a01a01a01a01ap-13);
    p = fma_(p, r, 0x1.6c16c16c16c17p-10);
    p = fma_(p, r, 0x1.1111111111111p-7);
    p = fma_(p, r, 0x1.5555555555555p-5);
    p = fma_(p, r, 0x1.5555555555555p-3);
    p = fma_(p, r, 0x1.0000000000000p-1);
    return (r * r) * p;
}

SB_HD double exp(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return from_bits(0x7ff0000000000000ull);
    if (x < -745.1332191019412) return 0.0;
    double k = rint_(x * INV_LN2);
    double r = fma_(-k, LN2_HI, x);
    r = fma_(-k, LN2_LO, r);
    double e = 1.0 + (r + expm1_tail(r));
    int ki = (int)k;
    if (ki > 1023) return (e * 2.0) * pow2i(ki - 1);
    if (ki < -1021) return (e * pow2i(ki + 600)) * pow2i(-600);
    return e * pow2i(ki);
}

// sin(r), cos(r) for |r| <= pi/4
SB_HD double sin_kernel(double r) {
    double z = r * r;
    double p = -0x1.2f49b46814157p-57;  // -1/19!
    p = fma_(p, z, 0x1.952c77030ad4ap-49);
    p = fma_(p, z, -0x1.ae7f3e733b81fp-41);
    p = fma_(p, z, 0x1.6124613a86d09p-33);
    p = fma_(p, z, -0x1.ae64567f544e4p-26);
    p = fma_(p, z, 0x1.71de3a556c734p-19);
    p = fma_(p, z, -0x1.a01a01a01a01ap-13);
    p = fma_(p, z, 0x1.1111111111111p-7);
    p = fma_(p, z, -0x1.5555555555555p-3);
    return fma_(r * z, p, r);
}
SB_HD double cos_kernel(double r) {
    double z = r * r;
    double p = 0x1.e542ba4020225p-62;  // 1/20!
    p = fma_(p, z, -0x1.6827863b97d97p-53);
    p = fma_(p, z, 0x1.ae7f3e733b81fp-45);
    p = fma_(p, z, -0x1.93974a8c07c9dp-37);
    p = fma_(p, z, 0x1.1eed8eff8d898p-29);
    p = fma_(p, z, -0x1.27e4fb7789f5cp-22);
    p = fma_(p, z, 0x1.a01a01a01a01ap-16);
    p = fma_(p, z, -0x1.6c16c16c16c17p-10);
    p = fma_(p, z, 0x1.5555555555555p-5);
    p = fma_(p, z, -0x1.0000000000000p-1);
    return fma_(z, p, 1.0);
}
// x = q*pi/2 + r, |r| <= ~pi/4; returns q mod 4
SB_HD int reduce_pio2(double x, double* r) {
    double q = rint_(x * TWO_OVER_PI);
    double y = fma_(-q, PIO2_1, x);
    y = fma_(-q, PIO2_2, y);
    y = fma_(-q, PIO2_3, y);
    *r = y;
    long long qi = (long long)q;
    return (int)(qi & 3);
}
SB_HD double sin(double x) {
    if (fabs_(x) < 0x1p-26) return x;
    double r;
    int q = reduce_pio2(x, &r);
    switch (q) {
        case 0: return sin_kernel(r);
        case 1: return cos_kernel(r);
        case 2: return -sin_kernel(r);
        default: return -cos_kernel(r);
    }
}
SB_HD double cos(double x) {
    if (fabs_(x) < 0x1p-27) return 1.0;
    double r;
    int q = reduce_pio2(x, &r);
    switch (q) {
        case 0: return cos_kernel(r);
        case 1: return -sin_kernel(r);
        case 2: return -cos_kernel(r);
        default: return sin_kernel(r);
    }
}

// asin(y) - y for 0 <= z = y^2 <= 1/4:  y * z * P(z)
SB_HD double asin_tail(double y, double z) {
    double p = 0x1.1052bc5fa960ap-9;
    p = fma_(p, z, 0x1.208d3570ae5a6p-9);
    p = fma_(p, z, 0x1.3275586c5f2f0p-9);
    p = fma_(p, z, 0x1.464c0950f7d47p-9);
    p = fma_(p, z, 0x1.5c5f56efaaaabp-9);
    p = fma_(p, z, 0x1.750de64d7d05fp-9);
    p = fma_(p, z, 0x1.90cb77f60c7cep-9);
    p = fma_(p, z, 0x1.b026f57b13b14p-9);
    p = fma_(p, z, 0x1.d3d2a8e0dd67dp-9);
    p = fma_(p, z, 0x1.fcaf8fb6db6dbp-9);
    p = fma_(p, z, 0x1.15ee9d45d1746p-8);
    p = fma_(p, z, 0x1.31683bdef7bdfp-8);
    p = fma_(p, z, 0x1.51ba308d3dcb1p-8);
    p = fma_(p, z, 0x1.782dda12f684cp-8);
    p = fma_(p, z, 0x1.a6863d70a3d71p-8);
    p = fma_(p, z, 0x1.df3bd37a6f4dfp-8);
    p = fma_(p, z, 0x1.12ef3cf3cf3cfp-7);
    p = fma_(p, z, 0x1.3fde50d79435ep-7);
    p = fma_(p, z, 0x1.7a87878787878p-7);
    p = fma_(p, z, 0x1.c99999999999ap-7);
    p = fma_(p, z, 0x1.1c4ec4ec4ec4fp-6);
    p = fma_(p, z, 0x1.6e8ba2e8ba2e9p-6);
    p = fma_(p, z, 0x1.f1c71c71c71c7p-6);
    p = fma_(p, z, 0x1.6db6db6db6db7p-5);
    p = fma_(p, z, 0x1.3333333333333p-4);
    p = fma_(p, z, 0x1.5555555555555p-3);
    return (y * z) * p;
}
SB_HD double acos(double x) {
    if (x != x || x > 1.0 || x < -1.0) return from_bits(0x7ff8000000000000ull);
    if (fabs_(x) <= 0.5) {
        double z = x * x;
        double t = asin_tail(x, z);
        return PIO2_HI - (x - (PIO2_LO - t));
    }
    if (x > 0.0) {
        double z = (1.0 - x) * 0.5;
        double s = sqrt_(z);
        double w = s + asin_tail(s, z);
        return 2.0 * w;
    }
    double z = (1.0 + x) * 0.5;
    double s = sqrt_(z);
    double w = s + asin_tail(s, z);
    return 2.0 * (PIO2_HI - (w - PIO2_LO));
}

// log(x) = hi + lo for finite x > 0
SB_HD void log_dd(double x, double* hi, double* lo) {
    int k = 0;
    uint64_t u = bits(x);
    if ((u >> 52) == 0) {  // subnormal
        x *= 0x1p54;
        k -= 54;
        u = bits(x);
    }
    k += (int)(u >> 52) - 1023;
    double m = from_bits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) {
        m *= 0.5;
        k += 1;
    }
    double f = m - 1.0;  // exact (Sterbenz)
    double d = 2.0 + f;
    double s = f / d;
    double s_lo = fma_(-s, d, f) / d;
    double z = s * s;
    double q = 0x1.47ae147ae147bp-5;  // 1/25
    q = fma_(q, z, 0x1.642c8590b2164p-5);
    q = fma_(q, z, 0x1.8618618618618p-5);
    q = fma_(q, z, 0x1.af286bca1af28p-5);
    q = fma_(q, z, 0x1.e1e1e1e1e1e1ep-5);
    q = fma_(q, z, 0x1.1111111111111p-4);
    q = fma_(q, z, 0x1.3b13b13b13b14p-4);
    q = fma_(q, z, 0x1.745d1745d1746p-4);
    q = fma_(q, z, 0x1.c71c71c71c71cp-4);
    q = fma_(q, z, 0x1.2492492492492p-3);
    q = fma_(q, z, 0x1.999999999999ap-3);
    q = fma_(q, z, 0x1.5555555555555p-2);
    double t = (2.0 * s) * z * q + 2.0 * s_lo;  // log m = 2s + t
    double kd = (double)k;
    // (k*LN2_HI) + 2s as a two-sum
    double a = kd * LN2_HI;  // exact: LN2_HI * k for |k| < 2^11 needs 64 bits; use fma error
    double a_lo = fma_(kd, LN2_HI, -a);
    double b = 2.0 * s;
    double sum = a + b;
    double bb = sum - a;
    double err = (a - (sum - bb)) + (b - bb);
    *hi = sum;
    *lo = err + a_lo + (t + kd * LN2_LO);
    double h2 = *hi + *lo;
    *lo = *lo - (h2 - *hi);
    *hi = h2;
}
SB_HD double log(double x) {
    if (x != x || x < 0.0) return from_bits(0x7ff8000000000000ull);
    if (x == 0.0) return -from_bits(0x7ff0000000000000ull);
    if (bits(x) == 0x7ff0000000000000ull) return x;
    double h, l;
    log_dd(x, &h, &l);
    return h + l;
}
// atan — for |u| <= tan(pi/8) the Taylor series u + u^3 P(u^2) (odd powers to
//   u^45); on (tan(pi/8), 1] atan(t) = pi/4 + atan((t - 1) / (t + 1)); for t > 1
//   atan(t) = pi/2 - atan(1 / t); odd symmetry (signed zero kept).
constexpr double PIO4_HI = 0x1.921fb54442d18p-1;
constexpr double PIO4_LO = 0x1.1a62633145c07p-55;
SB_HD double atan_kernel(double u) {
    double z = u * u;
    double p = 0x1.6c16c16c16c17p-6;  // 1/45
    p = fma_(p, z, -0x1.7d05f417d05f4p-6);
    p = fma_(p, z, 0x1.8f9c18f9c18fap-6);
    p = fma_(p, z, -0x1.a41a41a41a41ap-6);
    p = fma_(p, z, 0x1.bacf914c1bad0p-6);
    p = fma_(p, z, -0x1.d41d41d41d41dp-6);
    p = fma_(p, z, 0x1.f07c1f07c1f08p-6);
    p = fma_(p, z, -0x1.0842108421084p-5);
    p = fma_(p, z, 0x1.1a7b9611a7b96p-5);
    p = fma_(p, z, -0x1.2f684bda12f68p-5);
    p = fma_(p, z, 0x1.47ae147ae147bp-5);
    p = fma_(p, z, -0x1.642c8590b2164p-5);
    p = fma_(p, z, 0x1.8618618618618p-5);
    p = fma_(p, z, -0x1.af286bca1af28p-5);
    p = fma_(p, z, 0x1.e1e1e1e1e1e1ep-5);
    p = fma_(p, z, -0x1.1111111111111p-4);
    p = fma_(p, z, 0x1.3b13b13b13b14p-4);
    p = fma_(p, z, -0x1.745d1745d1746p-4);
    p = fma_(p, z, 0x1.c71c71c71c71cp-4);
    p = fma_(p, z, -0x1.2492492492492p-3);
    p = fma_(p, z, 0x1.999999999999ap-3);
    p = fma_(p, z, -0x1.5555555555555p-2);  // -1/3
    return u + (u * z) * p;
}
SB_HD double atan_01(double t) {  // 0 <= t <= 1
    if (t <= 0x1.a827999fcef32p-2) return atan_kernel(t);  // tan(pi/8)
    double u = (t - 1.0) / (t + 1.0);
    return PIO4_HI + (atan_kernel(u) + PIO4_LO);
}
SB_HD double atan(double x) {
    if (x != x) return x;
    double t = fabs_(x);
    double r = t <= 1.0 ? atan_01(t) : PIO2_HI - (atan_01(1.0 / t) - PIO2_LO);
    return (bits(x) >> 63) ? -r : r;
}

// x^y for x >= 0 (all uses on the lifting-line path take |.| first)
SB_HD double pow(double x, double y) {
    if (y == 0.0 || x == 1.0) return 1.0;
    if (x != x || y != y) return from_bits(0x7ff8000000000000ull);
    if (x == 0.0) return y > 0.0 ? 0.0 : from_bits(0x7ff0000000000000ull);
    if (x < 0.0) return from_bits(0x7ff8000000000000ull);
    double lh, ll;
    log_dd(x, &lh, &ll);
    double ph = y * lh;
    double pl = fma_(y, lh, -ph) + y * ll;
    double e = exp(ph);
    if (e == 0.0 || bits(e) == 0x7ff0000000000000ull) return e;
    return fma_(e, pl, e);
}

}  // namespace sbm
