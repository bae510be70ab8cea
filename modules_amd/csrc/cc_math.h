// cc_math.h -- device math for the CC kernels (gfx950).
//
// sincos_rev(u): sin(2*pi*u) and cos(2*pi*u) in FP64 for an argument already reduced to revolutions.
// Method (textbook, written from the mathematics, no vendor library code):
//   1. f = u - rint(u)            (exact for |u| < 2^52)          -> f in [-1/2, 1/2]
//   2. q = rint(4 f), g = f - q/4                                  -> g in [-1/8, 1/8] revolutions
//   3. x = 2 pi g (two-term 2*pi so the product keeps ~106 bits)   -> x in [-pi/4, pi/4]
//   4. Taylor series: sin x = x (1 - x^2/3! + ... + x^16/17!), cos x = 1 - x^2/2! + ... + x^16/16!
//      truncation error below 1e-19 relative on |x| <= pi/4 (next terms x^19/19!, x^18/18!)
//   5. quadrant q mod 4 rotates (sin, cos) by q quarter turns.
// Error: a few ulp against the correctly rounded value (measured in tests against libm).
#pragma once
#include <hip/hip_runtime.h>

namespace ccgpu {

// 2*pi split as hi + lo with hi exactly representable, lo the rounding remainder.
#define CC_TWO_PI_HI 6.283185307179586232
#define CC_TWO_PI_LO 2.4492935982947064e-16

__device__ __forceinline__ void sincos_rev(double u, double &s, double &c)
{
    double f = u - __builtin_rint(u);
    double qd = __builtin_rint(4.0 * f);
    double g = __builtin_fma(-0.25, qd, f);
    // x = 2 pi g with the low part of 2 pi folded in
    double x = __builtin_fma(g, CC_TWO_PI_HI, g * CC_TWO_PI_LO);
    double x2 = x * x;
    // sin: x * P(x^2), P = sum_{m=0..8} (-1)^m x^(2m)/(2m+1)!
    double ps = 2.8114572543455207632e-15;          //  1/17!
    ps = __builtin_fma(ps, x2, -7.6471637318198164759e-13); // -1/15!
    ps = __builtin_fma(ps, x2, 1.6059043836821614599e-10);  //  1/13!
    ps = __builtin_fma(ps, x2, -2.5052108385441718775e-08); // -1/11!
    ps = __builtin_fma(ps, x2, 2.7557319223985890653e-06);  //  1/9!
    ps = __builtin_fma(ps, x2, -1.9841269841269841270e-04); // -1/7!
    ps = __builtin_fma(ps, x2, 8.3333333333333333333e-03);  //  1/5!
    ps = __builtin_fma(ps, x2, -1.6666666666666666667e-01); // -1/3!
    double sx = __builtin_fma(ps * x2, x, x);
    // cos: Q(x^2) = sum_{m=0..8} (-1)^m x^(2m)/(2m)!
    double pc = 4.7794773323873852974e-14;           //  1/16!
    pc = __builtin_fma(pc, x2, -1.1470745597729724714e-11); // -1/14!
    pc = __builtin_fma(pc, x2, 2.0876756987868098979e-09);  //  1/12!
    pc = __builtin_fma(pc, x2, -2.7557319223985890653e-07); // -1/10!
    pc = __builtin_fma(pc, x2, 2.4801587301587301587e-05);  //  1/8!
    pc = __builtin_fma(pc, x2, -1.3888888888888888889e-03); // -1/6!
    pc = __builtin_fma(pc, x2, 4.1666666666666666667e-02);  //  1/4!
    pc = __builtin_fma(pc, x2, -0.5);                        // -1/2!
    double cx = __builtin_fma(pc, x2, 1.0);
    int q = (int)qd;
    double ss = (q & 1) ? cx : sx;
    double cc = (q & 1) ? sx : cx;
    s = (q & 2) ? -ss : ss;
    c = ((q + 1) & 2) ? -cc : cc;
}

// a * b + c for a constant c (wavefront-uniform, held in SGPRs), always in the three-operand encoding.  Left to
// itself the compiler picks the two-operand form, whose destination is the addend, and then copies the constant
// into the destination register first: one extra VALU move per use in the phase loops.
__device__ __forceinline__ double fma_sc(double a, double b, double c)
{
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
}

// Table of exp(2 pi i t / 64), t = 0..63, for sincos_rev_tab: threads 0..63 of the block each write one
// entry (sincos_rev, a few ulp); the caller synchronises before use.
__device__ __forceinline__ void sincos_table_fill(double2 *tab)
{
    if (threadIdx.x < 64) {
        double s, c;
        sincos_rev((double)threadIdx.x * 0.015625, s, c);
        tab[threadIdx.x] = make_double2(c, s);
    }
}

// sin and cos of 2 pi u through the 64-entry table: u = j/64 + f_lo with |f_lo| <= 1/128 (exact split), then
// exp(2 pi i u) = T[j mod 64] * exp(2 pi i f_lo) with short Taylor series on |x| = |2 pi f_lo| <= pi/64:
// sin to x^7 (next term below 5e-18), cos to x^8 (next term below 3e-20).  About 19 VALU operations and one
// LDS read, against about 36 operations for sincos_rev.
//
// The split takes j = rint(64 u) from the sum 64 u + 1.5 * 2^52 (valid for |u| < 2^45): its rounding to an integer
// is rint's (ties to even, the constant being even), j = sum - constant is exact, and the low word of the sum holds
// j mod 2^32 in two's complement, so j mod 64 needs no conversion.  f_lo = u - j/64 is exact (the same value as
// reducing u to [-1/2, 1/2] first and splitting that: the two j differ by a multiple of 64).
__device__ __forceinline__ void sincos_rev_tab(double u, const double2 *tab, double &s, double &c)
{
    constexpr double kRound = 6755399441055744.0;  // 1.5 * 2^52
    const double sum = fma_sc(u, 64.0, kRound);
    const double jd = sum - kRound;
    const double2 t = tab[__double2loint(sum) & 63];  // read first: the polynomials below cover its latency
    const double fl = __builtin_fma(-0.015625, jd, u);  // exact
    const double x = fl * CC_TWO_PI_HI;
    const double x2 = x * x;
    double ps = fma_sc(x2, -1.9841269841269841270e-04, 8.3333333333333333333e-03);
    ps = fma_sc(ps, x2, -1.6666666666666666667e-01);
    const double sl = __builtin_fma(ps * x2, x, x);
    double pc = fma_sc(x2, 2.4801587301587301587e-05, -1.3888888888888888889e-03);
    pc = fma_sc(pc, x2, 4.1666666666666666667e-02);
    pc = __builtin_fma(pc, x2, -0.5);
    const double cl = __builtin_fma(pc, x2, 1.0);
    c = __builtin_fma(t.x, cl, -t.y * sl);
    s = __builtin_fma(t.y, cl, t.x * sl);
}

// The 256-entry variant of the particle-row contractions: u = j/256 + f with |f| <= 1/512 revolutions (the same
// exact split), and 2 pi folded into the series in f: sin(2 pi f) = f (c1 + c3 f^2 + c5 f^4) (next term below
// 9e-18), cos(2 pi f) = 1 + c2 f^2 + c4 f^4 + c6 f^6 (next term below 2e-20), c_k = (+-) (2 pi)^k / k!.  About 16
// VALU operations and one LDS read.  The table (4 KB in LDS) is exp(2 pi i t / 256) correctly rounded, copied by
// the whole block from a constant array (cc_tab256.h).

__device__ __forceinline__ void sincos_rev_tab256(double u, const double2 *tab, double &s, double &c)
{
    constexpr double kRound = 6755399441055744.0;  // 1.5 * 2^52: j = rint(256 u) for |u| < 2^43
    const double sum = fma_sc(u, 256.0, kRound);
    const double jd = sum - kRound;
    const double2 t = tab[__double2loint(sum) & 255];  // read first: the polynomials below cover its latency
    const double f = __builtin_fma(-0.00390625, jd, u);  // exact
    const double f2 = f * f;
    double ps = fma_sc(f2, 81.605249276075054203, -41.341702240399760234);
    ps = fma_sc(ps, f2, 6.2831853071795864769);
    const double sl = ps * f;
    double pc = fma_sc(f2, -85.456817206693727736, 64.939394022668291491);
    pc = fma_sc(pc, f2, -19.739208802178717238);
    const double cl = __builtin_fma(pc, f2, 1.0);
    c = __builtin_fma(t.x, cl, -t.y * sl);
    s = __builtin_fma(t.y, cl, t.x * sl);
}

// FP32 counterparts for the mixed-precision path: the table as float2 (rounded from the FP64 table), and
// sin/cos of 2 pi f for an argument already reduced to f in [-1/2, 1/2] revolutions in FP64 (the
// per-particle range reduction) and rounded to FP32.  Same split as sincos_rev_tab; sin to x^5, cos to x^4
// (next terms below 7e-13 and 2e-11, under FP32 rounding).
__device__ __forceinline__ void sincos_table_fill32(float2 *tab32, const double2 *tab)
{
    if (threadIdx.x < 64) tab32[threadIdx.x] = make_float2((float)tab[threadIdx.x].x, (float)tab[threadIdx.x].y);
}

__device__ __forceinline__ void sincos_rev_tab32(float f, const float2 *tab, float &s, float &c)
{
    const float jd = __builtin_rintf(64.0f * f);
    const float2 t = tab[((int)jd) & 63];
    const float fl = __builtin_fmaf(-0.015625f, jd, f);  // exact (Sterbenz)
    const float x = fl * 6.28318530717958648f;
    const float x2 = x * x;
    const float sl = __builtin_fmaf(x * x2, __builtin_fmaf(x2, 8.33333333e-3f, -1.66666667e-1f), x);
    const float cl = __builtin_fmaf(x2, __builtin_fmaf(x2, 4.16666667e-2f, -0.5f), 1.0f);
    c = __builtin_fmaf(t.x, cl, -t.y * sl);
    s = __builtin_fmaf(t.y, cl, t.x * sl);
}

// Radian argument: u = theta / (2 pi) formed in double-double before the reduction, so the conversion
// adds no error beyond the ulp of theta itself.
__device__ __forceinline__ void sincos_rad(double theta, double &s, double &c)
{
    const double inv_hi = 0.15915494309189534561;   // 1/(2 pi) rounded
    const double inv_lo = -9.8393383375912435e-18;  // remainder
    double uh = theta * inv_hi;
    double ul = __builtin_fma(theta, inv_hi, -uh);
    ul = __builtin_fma(theta, inv_lo, ul);
    double r = __builtin_rint(uh);
    sincos_rev((uh - r) + ul, s, c);
}

}  // namespace ccgpu
