// Device math for the w-phase term of the image chain.
#pragma once

#include <hip/hip_runtime.h>

#include "rick_internal.h"

namespace rick {

// sin and cos of a double, accurate to a few 1e-16 absolute for |x| < 2^40.
//
// Cody-Waite argument reduction (W. J. Cody and W. Waite, "Software Manual
// for the Elementary Functions", 1980) by pi/2 with a three-part split of
// pi/2 (P1 = RN(pi/2), P2 = RN(pi/2 - P1), P3 = RN(pi/2 - P1 - P2)), each
// product k*Pi applied inside one fused multiply-add, then the Taylor series
// of sin and cos on |r| <= pi/4 truncated after r^15 / r^16 (the first
// omitted terms are below 7e-17 relative).  Coefficients are the exact
// reciprocal factorials rounded to double.
__device__ __forceinline__ void rick_sincos(double x, double *s, double *c)
{
    constexpr double kTwoOverPi = 0.63661977236758134308;
    constexpr double P1 = 1.5707963267948966;        // RN(pi/2)
    constexpr double P2 = 6.123233995736766e-17;     // RN(pi/2 - P1)
    constexpr double P3 = -1.4973849048591698e-33;   // RN(pi/2 - P1 - P2)
    const double k = rint(x * kTwoOverPi);
    double r = __fma_rn(-k, P1, x);
    r = __fma_rn(-k, P2, r);
    r = __fma_rn(-k, P3, r);
    const double r2 = r * r;
    // sin r = r * (1 - r^2/3! + r^4/5! - ... - r^14/15!)
    double ps = -7.6471637318198164759e-13;                // -1/15!
    ps = __fma_rn(ps, r2, 1.6059043836821614599e-10);       //  1/13!
    ps = __fma_rn(ps, r2, -2.5052108385441718775e-08);      // -1/11!
    ps = __fma_rn(ps, r2, 2.7557319223985890653e-06);       //  1/9!
    ps = __fma_rn(ps, r2, -1.9841269841269841270e-04);      // -1/7!
    ps = __fma_rn(ps, r2, 8.3333333333333333333e-03);       //  1/5!
    ps = __fma_rn(ps, r2, -1.6666666666666666667e-01);      // -1/3!
    const double sr = __fma_rn(ps * r2, r, r);
    // cos r = 1 - r^2/2! + r^4/4! - ... + r^16/16!
    double pc = 4.7794773323873852974e-14;                  //  1/16!
    pc = __fma_rn(pc, r2, -1.1470745597729724714e-11);      // -1/14!
    pc = __fma_rn(pc, r2, 2.0876756987868098979e-09);       //  1/12!
    pc = __fma_rn(pc, r2, -2.7557319223985890653e-07);      // -1/10!
    pc = __fma_rn(pc, r2, 2.4801587301587301587e-05);       //  1/8!
    pc = __fma_rn(pc, r2, -1.3888888888888888889e-03);      // -1/6!
    pc = __fma_rn(pc, r2, 4.1666666666666666667e-02);       //  1/4!
    pc = __fma_rn(pc, r2, -5.0000000000000000000e-01);      // -1/2!
    const double cr = __fma_rn(pc, r2, 1.0);
    // Quadrant: x = k*pi/2 + r.
    const int q = (int)((long long)k & 3);
    const double ss = (q & 1) ? cr : sr;
    const double cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}

// Per-pixel n = sqrt(1 - l^2 - m^2) (phase_correction.cu:212, 222), no FMA:
// 1 - radius2 cancels, so contracting l*l + m*m would move n by an ulp and
// the phase by ~1e-11 rad at |w| ~ 1e4.
__device__ __forceinline__ double pixel_n(double l, double m)
{
#pragma clang fp contract(off)
    const double radius2 = (l * l + m * m);
    return sqrt(1 - radius2);
}

// One w-plane's contribution to a pixel, phase_correction.cu:222-223 and
// 233-235 with the same operation order: arg = 2*PI*wterm*(n - 1),
// (p + iq) * exp(i*arg) * dwnorm * n.
__device__ __forceinline__ void phase_add_f64(double p, double q, double wterm, double nf,
                                              double dwnorm, double &are, double &aim)
{
#pragma clang fp contract(off)
    const double arg = 2.0 * kRefPi * wterm * (nf - 1.0);
    double sn, cs;
    rick_sincos(arg, &sn, &cs);
    are += (p * cs - q * sn) * dwnorm * nf;
    aim += (p * sn + q * cs) * dwnorm * nf;
}

// FP32 plane data: the phase in turns, wterm * (n - 1) = arg / (2 pi), and
// its reduction to one turn stay FP64; the rotation is FP32 (the hardware
// sine/cosine of a fraction of a turn in [-1/2, 1/2], absolute error a few
// 1e-7), the accumulation FP64.  The pixel's factor dwnorm * n is applied
// once to the plane sum by the caller.
__device__ __forceinline__ void phase_add_f32(float p, float q, double wterm, double nm1,
                                              double &are, double &aim)
{
    const double turns = wterm * nm1;
    const float f = (float)(turns - rint(turns));
    const float sn = __builtin_amdgcn_sinf(f), cs = __builtin_amdgcn_cosf(f);
    are += (double)(p * cs - q * sn);
    aim += (double)(p * sn + q * cs);
}

}  // namespace rick
