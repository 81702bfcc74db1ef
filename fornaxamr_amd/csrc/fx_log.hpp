// Correctly rounded natural logarithm in FP64 for host and device.
//
// The anomalous-resistivity mask of the CT step (src/prj_timeint.c:230-263)
// takes log() of rho / 1e13 and Ye / 0.25.  Neither the device's libm nor a
// host libm promises a correctly rounded log, so the two can differ in the
// last bit.  log_cr evaluates log(x) in double-double arithmetic (about 2^-104
// relative error) and rounds once, which gives the correctly rounded double
// except when log(x) lies within ~2^-104 relative of a rounding boundary.  A
// libm that rounds correctly on an input therefore returns the same bits.
//
// Method: x = 2^e m with m in [sqrt(1/2), sqrt(2)), s = (m - 1)/(m + 1) (|s| <=
// 0.1716), log(m) = 2 atanh(s) = 2 s sum_{n>=0} s^(2n) / (2n + 1), truncated
// after n = 20 (the remainder is below 2^-108 relative); the terms n >= 10 are
// summed in double (their sum is below 2^-55 of the total), the rest by a
// double-double Horner scheme; log(x) = e ln2 + log(m) with ln2 as a
// double-double.  Constants: 1/(2n+1) and ln2 rounded to double-double
// (leading double and the double nearest to the remainder).
//
// Only the IEEE operations +, -, *, / and fma are used, so the host and device
// builds of this code return identical bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace fx {

struct dd {
    double hi, lo;
};

__host__ __device__ inline dd dd_two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}

__host__ __device__ inline dd dd_fast_sum(double a, double b)   // |a| >= |b|
{
    const double s = a + b;
    return {s, b - (s - a)};
}

__host__ __device__ inline dd dd_two_prod(double a, double b)
{
    const double p = a * b;
    return {p, fma(a, b, -p)};
}

__host__ __device__ inline dd dd_add(dd a, dd b)
{
    dd s = dd_two_sum(a.hi, b.hi);
    const dd t = dd_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = dd_fast_sum(s.hi, s.lo);
    s.lo += t.lo;
    return dd_fast_sum(s.hi, s.lo);
}

__host__ __device__ inline dd dd_mul(dd a, dd b)
{
    dd p = dd_two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return dd_fast_sum(p.hi, p.lo);
}

// a / b for double a and double-double b: three correction steps
__host__ __device__ inline dd dd_div(double a, dd b)
{
    const double q1 = a / b.hi;
    dd r = dd_add({a, 0.0}, dd_mul({-q1, 0.0}, b));
    const double q2 = r.hi / b.hi;
    r = dd_add(r, dd_mul({-q2, 0.0}, b));
    const double q3 = r.hi / b.hi;
    return dd_add(dd_fast_sum(q1, q2), {q3, 0.0});
}

// log(x) for finite x >= 1, correctly rounded (see above).
__host__ __device__ inline double log_cr(double x)
{
    // 1/(2n+1), n = 0..9, as double-doubles
    constexpr double inv_hi[10] = {
        0x1.0000000000000p+0, 0x1.5555555555555p-2, 0x1.999999999999ap-3, 0x1.2492492492492p-3,
        0x1.c71c71c71c71cp-4, 0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4, 0x1.1111111111111p-4,
        0x1.e1e1e1e1e1e1ep-5, 0x1.af286bca1af28p-5};
    constexpr double inv_lo[10] = {
        0.0, 0x1.5555555555555p-56, -0x1.999999999999ap-57, 0x1.2492492492492p-57,
        0x1.c71c71c71c71cp-58, -0x1.745d1745d1746p-59, -0x1.3b13b13b13b14p-58,
        0x1.1111111111111p-60, 0x1.e1e1e1e1e1e1ep-61, 0x1.af286bca1af28p-59};
    // 1/(2n+1), n = 10..20, as doubles
    constexpr double inv_tail[11] = {
        0x1.8618618618618p-5, 0x1.642c8590b2164p-5, 0x1.47ae147ae147bp-5, 0x1.2f684bda12f68p-5,
        0x1.1a7b9611a7b96p-5, 0x1.0842108421084p-5, 0x1.f07c1f07c1f08p-6, 0x1.d41d41d41d41dp-6,
        0x1.bacf914c1bad0p-6, 0x1.a41a41a41a41ap-6, 0x1.8f9c18f9c18fap-6};
    constexpr double ln2_hi = 0x1.62e42fefa39efp-1, ln2_lo = 0x1.abc9e3b39803fp-56;
    constexpr double sqrt_half = 0x1.6a09e667f3bcdp-1;

    int e = 0;
    double m = frexp(x, &e);            // x = m 2^e, m in [0.5, 1): exact
    if (m < sqrt_half) {
        m *= 2.0;                       // exact
        e -= 1;
    }
    const double num = m - 1.0;         // exact (Sterbenz)
    const dd den = dd_two_sum(m, 1.0);
    const dd s = dd_div(num, den);
    const dd t = dd_mul(s, s);
    double tail = inv_tail[10];
    for (int n = 9; n >= 0; --n) tail = inv_tail[n] + t.hi * tail;
    dd S = dd_add({inv_hi[9], inv_lo[9]}, dd_mul(t, {tail, 0.0}));
    for (int n = 8; n >= 0; --n) S = dd_add({inv_hi[n], inv_lo[n]}, dd_mul(t, S));
    dd lm = dd_mul(s, S);
    lm.hi *= 2.0;                       // exact
    lm.lo *= 2.0;
    const double ke = (double)e;
    dd k = dd_two_prod(ke, ln2_hi);
    k.lo += ke * ln2_lo;
    const dd r = dd_add(k, lm);
    return r.hi;
}

}  // namespace fx
