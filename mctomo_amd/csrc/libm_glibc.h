// libm_glibc.h -- exp and sincos that return the host C library's numbers.
//
// surfdisp96's period equations (surfmodes/surfdisp96.f:1036-1217: dltar1,
// var, dnka) call exp, sin and cos; the reference's build (and the oracle)
// link them from glibc, where a sin and a cos of the same argument become one
// sincos call.  Near a root the period equation is a small difference of large
// terms, so a last-bit change in one of these values can move a sign test of
// the root search and with it the root (within surfdisp96's own tolerance,
// 1e-6 relative).  These two functions restate, in double arithmetic with the
// same operations in the same order, the published methods glibc (2.28 and
// later) implements, so the device computes the same numbers:
//
//   libm_exp: the table-driven exp of ARM's Optimized Routines (Szabolcs Nagy,
//     2018): x = k ln2/128 + r with k = round(x 128/ln2) (a shift by 1.5*2^52),
//     r from a two-part ln2/128; exp(r) - 1 by a degree-5 polynomial; 2^(k/128)
//     from a 128-entry table of (tail, head) pairs with the exponent added to
//     the head's bits.  x86-64 glibc selects the build compiled with fused
//     multiply-adds on CPUs that have them, so the contractions are written
//     out as fma().  Valid for |x| < 512 (the only arguments here: -120..0).
//
//   libm_sincos: the IBM Accurate Mathematical Library's sin/cos (Ziv et al.)
//     as glibc's sincos runs it: |x| < 0.126 by an odd polynomial in x
//     (TAYLOR_SIN with its minimax coefficients), otherwise the argument
//     rounded to a multiple of 1/128 (by adding 1.5*2^45), sin and cos at that
//     point as double-doubles from a table and short polynomials for the
//     remainder; 0.855 <= |x| < 2.426 from pi/2 - |x| as a double-double;
//     larger |x| by a four-part pi/2 (Cody-Waite) to a double-double in
//     [-pi/4, pi/4] and the quadrant.  glibc's sincos has no fused-multiply-add
//     build, so every operation rounds on its own: this file must be compiled
//     with -ffp-contract=off.  Valid for |x| < 105414350.
//
// The tables (libm_glibc_tables.h) are computed from their definitions by
// gen_libm_tables.py.  tests/test_libm_replica.py compiles this header for the
// host and compares both functions with the host's glibc bit for bit on random
// arguments over every branch.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "libm_glibc_tables.h"

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#define __device__
#endif
#ifndef __forceinline__
#define __forceinline__ inline __attribute__((always_inline))
#endif
#endif

namespace mct_libm {

__host__ __device__ __forceinline__ uint64_t lm_bits(double x)
{
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}
__host__ __device__ __forceinline__ double lm_dbl(uint64_t u)
{
    double x;
    memcpy(&x, &u, 8);
    return x;
}

// exp(x) for |x| < 512.
__host__ __device__ __forceinline__ double libm_exp(double x)
{
    const double inv_ln2_n = 0x1.71547652b82fep0 * 128, shift = 0x1.8p52;
    const double neg_ln2_hi_n = -0x1.62e42fefa0000p-8, neg_ln2_lo_n = -0x1.cf79abc9e3b3ap-47;
    const double c2 = 0x1.ffffffffffdbdp-2, c3 = 0x1.555555555543cp-3;
    const double c4 = 0x1.55555cf172b91p-5, c5 = 0x1.1111167a4d017p-7;
    const uint32_t abstop = (uint32_t)(lm_bits(x) >> 52) & 0x7ff;
    if (abstop < 0x3c9)   // |x| < 2^-54
        return 1.0 + x;
    double kd = fma(inv_ln2_n, x, shift);
    const uint64_t ki = lm_bits(kd);
    kd -= shift;
    const double r = fma(kd, neg_ln2_lo_n, fma(kd, neg_ln2_hi_n, x));
    const int idx = (int)(ki & 127);
    const uint64_t top = ki << 45;
    const double tail = LIBM_EXP_TAB[2 * idx];
    const uint64_t sbits = lm_bits(LIBM_EXP_TAB[2 * idx + 1]) + (top - ((uint64_t)idx << 45));
    const double r2 = r * r;
    const double tmp = fma(r2 * r2, fma(r, c5, c4), fma(r2, fma(r, c3, c2), tail + r));
    const double scale = lm_dbl(sbits);
    return fma(scale, tmp, scale);
}

// ---- sincos -----------------------------------------------------------------------
struct lm_sc {
    static constexpr double big = 0x1.8p45;
    static constexpr double sn3 = -1.66666666666664880952546298448555E-01;
    static constexpr double sn5 = 8.33333214285722277379541354343671E-03;
    static constexpr double cs2 = 4.99999999999999999999950396842453E-01;
    static constexpr double cs4 = -4.16666666666664434524222570944589E-02;
    static constexpr double cs6 = 1.38888874007937613028114285595617E-03;
    static constexpr double s1 = -0x1.5555555555555p-3, s2 = 8.3333333333323288e-03;
    static constexpr double s3 = -1.9841269834414642e-04, s4 = 2.755729806860771e-06;
    static constexpr double s5 = -2.5022014848318398e-08;
    static constexpr double hp0 = 0x1.921fb54442d18p+0, hp1 = 0x1.1a62633145c07p-54;
    static constexpr double hpinv = 0x1.45f306dc9c883p-1, toint = 0x1.8p52;
    static constexpr double mp1 = 0x1.921fb58p0, mp2 = -0x1.dde973cp-27;
    static constexpr double pp3 = -0x1.cb3b398p-55, pp4 = -0x1.d747f23e32ed7p-83;
};

__host__ __device__ __forceinline__ double lm_copysign(double a, double s)
{
    return lm_dbl((lm_bits(a) & 0x7fffffffffffffffull) | (lm_bits(s) & 0x8000000000000000ull));
}
__host__ __device__ __forceinline__ double lm_fabs(double a) { return lm_dbl(lm_bits(a) & 0x7fffffffffffffffull); }

// sin(a + da), |a| < 0.855
__host__ __device__ __forceinline__ double lm_do_sin(double x, double dx)
{
    const double xold = x;
    if (lm_fabs(x) < 0.126) {
        const double xx = x * x;
        const double p = ((((lm_sc::s5 * xx + lm_sc::s4) * xx + lm_sc::s3) * xx + lm_sc::s2) * xx) + lm_sc::s1;
        const double t = ((p * x - 0.5 * dx) * xx + dx);
        return x + t;
    }
    if (x <= 0) dx = -dx;
    const double u = lm_sc::big + lm_fabs(x);
    x = lm_fabs(x) - (u - lm_sc::big);
    const double xx = x * x;
    const double s = x + (dx + x * xx * (lm_sc::sn3 + xx * lm_sc::sn5));
    const double c = x * dx + xx * (lm_sc::cs2 + xx * (lm_sc::cs4 + xx * lm_sc::cs6));
    const int k = (int)(uint32_t)lm_bits(u) << 2;
    const double sn = LIBM_SC_TAB[k], ssn = LIBM_SC_TAB[k + 1], cs = LIBM_SC_TAB[k + 2], ccs = LIBM_SC_TAB[k + 3];
    const double cor = (ssn + s * ccs - sn * c) + cs * s;
    return lm_copysign(sn + cor, xold);
}

// cos(a + da), |a| < 0.855
__host__ __device__ __forceinline__ double lm_do_cos(double x, double dx)
{
    if (x < 0) dx = -dx;
    const double u = lm_sc::big + lm_fabs(x);
    x = lm_fabs(x) - (u - lm_sc::big) + dx;
    const double xx = x * x;
    const double s = x + x * xx * (lm_sc::sn3 + xx * lm_sc::sn5);
    const double c = xx * (lm_sc::cs2 + xx * (lm_sc::cs4 + xx * lm_sc::cs6));
    const int k = (int)(uint32_t)lm_bits(u) << 2;
    const double sn = LIBM_SC_TAB[k], ssn = LIBM_SC_TAB[k + 1], cs = LIBM_SC_TAB[k + 2], ccs = LIBM_SC_TAB[k + 3];
    const double cor = (ccs - s * ssn - cs * c) - sn * s;
    return cs + cor;
}

// sin(x) and cos(x) for |x| < 105414350 (larger |x| is not reduced: NaN).
__host__ __device__ __forceinline__ void libm_sincos(double x, double *sinx, double *cosx)
{
    const uint32_t k = (uint32_t)(lm_bits(x) >> 32) & 0x7fffffffu;
    if (k < 0x400368fdu) {
        if (k < 0x3e400000u) { *sinx = x; *cosx = 1.0; return; }
        if (k < 0x3feb6000u) { *sinx = lm_do_sin(x, 0); *cosx = lm_do_cos(x, 0); return; }
        const double y = lm_sc::hp0 - lm_fabs(x), a = y + lm_sc::hp1, da = (y - a) + lm_sc::hp1;
        *sinx = lm_copysign(lm_do_cos(a, da), x);
        *cosx = lm_do_sin(a, da);
        return;
    }
    if (k >= 0x419921fbu) { *sinx = *cosx = __builtin_nan(""); return; }
    // pi/2 in four parts; the quadrant from the low bits of the shifted multiple
    const double t = (x * lm_sc::hpinv + lm_sc::toint);
    const double xn = t - lm_sc::toint;
    const double y = (x - xn * lm_sc::mp1) - xn * lm_sc::mp2;
    const unsigned n = (unsigned)lm_bits(t) & 3u;
    double t1 = xn * lm_sc::pp3, t2 = y - t1;
    double db = (y - t2) - t1;
    t1 = xn * lm_sc::pp4;
    double a = t2 - t1;
    db += (t2 - a) - t1;
    if (n == 1 || n == 2) { a = -a; db = -db; }
    const double s = lm_do_sin(a, db), c = lm_do_cos(a, db);
    if (n & 1) { *cosx = s; *sinx = (n & 2) ? -c : c; }
    else { *sinx = s; *cosx = (n & 2) ? -c : c; }
}

}   // namespace mct_libm
