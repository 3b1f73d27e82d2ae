// Device transcendentals that reproduce the host C library the reference's
// CPU simulator links (glibc 2.35 on x86-64 with FMA) bit for bit, so that a
// rule's priority -- and therefore every argmax and reward -- is identical on
// the GPU and the CPU.
//
//   sigp_exp   table-driven exp (ARM optimized-routines method, 128-entry
//              2^(k/128) table from gen_exp_table.py, degree-5 polynomial),
//              with the fused multiply-adds placed where glibc's FMA build
//              places them; its subnormal-range rescaling is unfused.
//   sigp_erf,  Sun fdlibm s_erf.c rational approximations (5 intervals),
//   sigp_erfc  evaluated in glibc's split even/odd polynomial order, calling
//              sigp_exp for the exp(-x^2) factor.
//   sigp_pow_glibc  pow by the same library's method (sigp_pow.h): 128-entry
//              log table (gen_pow_log_table.py), double-double log, exp of
//              the pair.  Used by the rmod operator only.
//
// Compiled with -ffp-contract=off: every fma below is explicit.
//
// Licences of the methods and coefficients restated here:
//   fdlibm (erf, erfc): Copyright (C) 1993 by Sun Microsystems, Inc. All
//     rights reserved.  Developed at SunPro, a Sun Microsystems, Inc.
//     business.  Permission to use, copy, modify, and distribute this
//     software is freely granted, provided that this notice is preserved.
//   optimized-routines (exp, pow): Copyright (c) 2018 Arm Limited,
//     SPDX-License-Identifier: MIT OR Apache-2.0 WITH LLVM-exception.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include "sigp_exp_table.h"
#include "sigp_pow.h"
#include "sigp_pow_log_table.h"

__device__ __constant__ static const uint64_t sigp_exp_tab[256] = SIGP_EXP_TAB_INIT;

__device__ __forceinline__ uint64_t sigp_bits(double x) { return (uint64_t)__double_as_longlong(x); }
__device__ __forceinline__ double sigp_dbl(uint64_t u) { return __longlong_as_double((long long)u); }

__device__ __forceinline__ static double sigp_exp(double x) {
    const double inv_ln2_n = 0x1.71547652b82fep0 * 128.0;
    const double shift = 0x1.8p52;
    const double neg_ln2_hi_n = -0x1.62e42fefa0000p-8;
    const double neg_ln2_lo_n = -0x1.cf79abc9e3b3ap-47;
    const double c2 = 0x1.ffffffffffdbdp-2, c3 = 0x1.555555555543cp-3;
    const double c4 = 0x1.55555cf172b91p-5, c5 = 0x1.1111167a4d017p-7;
    const uint32_t abstop = (uint32_t)(sigp_bits(x) >> 52) & 0x7ffu;
    bool large = false;
    if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {
        if (abstop - 0x3c9u >= 0x80000000u) return 1.0 + x;   // |x| < 2^-54
        if (abstop >= 0x409u) {                                 // |x| >= 1024
            if (sigp_bits(x) == 0xfff0000000000000ull) return 0.0;
            if (abstop >= 0x7ffu) return 1.0 + x;
            return (sigp_bits(x) >> 63) ? 0.0 : __longlong_as_double(0x7ff0000000000000ll);
        }
        large = true;                                            // 512 <= |x| < 1024
    }
    double kd = __fma_rn(inv_ln2_n, x, shift);
    const uint64_t ki = sigp_bits(kd);
    kd -= shift;
    const double r = __fma_rn(kd, neg_ln2_lo_n, __fma_rn(kd, neg_ln2_hi_n, x));
    const uint32_t idx = 2u * (uint32_t)(ki & 127u);
    const uint64_t top = ki << 45;
    const double tail = sigp_dbl(sigp_exp_tab[idx]);
    uint64_t sbits = sigp_exp_tab[idx + 1] + top;
    const double r2 = r * r;
    const double tmp = __fma_rn(r2 * r2, __fma_rn(r, c5, c4), __fma_rn(r2, __fma_rn(r, c3, c2), tail + r));
    if (large) {
        if ((ki & 0x80000000ull) == 0) {                         // result may overflow
            sbits -= 1009ull << 52;
            const double scale = sigp_dbl(sbits);
            return 0x1p1009 * __fma_rn(scale, tmp, scale);
        }
        sbits += 1022ull << 52;                                  // result may be subnormal
        const double scale = sigp_dbl(sbits);
        const double st = scale * tmp;
        double y = scale + st;
        if (y < 1.0) {
            double lo = scale - y + st;
            const double hi = 1.0 + y;
            lo = 1.0 - hi + y + lo;
            y = (hi + lo) - 1.0;
            if (y == 0.0) y = 0.0;
        }
        return 0x1p-1022 * y;
    }
    const double scale = sigp_dbl(sbits);
    return __fma_rn(scale, tmp, scale);
}

// fdlibm s_erf.c coefficients (Sun Microsystems, 1993).
#define SIGP_ERX 8.45062911510467529297e-01
#define SIGP_EFX 1.28379167095512586316e-01

__device__ __forceinline__ void sigp_erf_small(double z, double *r, double *s) {
    const double pp0 = 1.28379167095512558561e-01, pp1 = -3.25042107247001499370e-01,
                 pp2 = -2.84817495755985104766e-02, pp3 = -5.77027029648944159157e-03,
                 pp4 = -2.37630166566501626084e-05;
    const double qq1 = 3.97917223959155352819e-01, qq2 = 6.50222499887672944485e-02,
                 qq3 = 5.08130628187576562776e-03, qq4 = 1.32494738004321644526e-04,
                 qq5 = -3.96022827877536812320e-06;
    const double r1 = pp0 + z * pp1, z2 = z * z;
    const double r2 = pp2 + z * pp3, z4 = z2 * z2;
    const double s1 = 1.0 + z * qq1;
    const double s2 = qq2 + z * qq3;
    const double s3 = qq4 + z * qq5;
    *r = r1 + z2 * r2 + z4 * pp4;
    *s = s1 + z2 * s2 + z4 * s3;
}

__device__ __forceinline__ void sigp_erf_mid(double s, double *p, double *q) {
    const double pa0 = -2.36211856075265944077e-03, pa1 = 4.14856118683748331666e-01,
                 pa2 = -3.72207876035701323847e-01, pa3 = 3.18346619901161753674e-01,
                 pa4 = -1.10894694282396677476e-01, pa5 = 3.54783043256182359371e-02,
                 pa6 = -2.16637559486879084300e-03;
    const double qa1 = 1.06420880400844228286e-01, qa2 = 5.40397917702171048937e-01,
                 qa3 = 7.18286544141962662868e-02, qa4 = 1.26171219808761642112e-01,
                 qa5 = 1.36370839120290507362e-02, qa6 = 1.19844998467991074170e-02;
    const double p1 = pa0 + s * pa1, s2 = s * s;
    const double q1 = 1.0 + s * qa1, s4 = s2 * s2;
    const double p2 = pa2 + s * pa3, s6 = s4 * s2;
    const double q2 = qa2 + s * qa3;
    const double p3 = pa4 + s * pa5;
    const double q3 = qa4 + s * qa5;
    *p = p1 + s2 * p2 + s4 * p3 + s6 * pa6;
    *q = q1 + s2 * q2 + s4 * q3 + s6 * qa6;
}

// Tail rational R/S in s = 1/x^2; `near` selects the [1.25, 1/0.35) interval.
__device__ __forceinline__ void sigp_erf_tail(double s, bool near, double *R, double *S) {
    if (near) {
        const double ra0 = -9.86494403484714822705e-03, ra1 = -6.93858572707181764372e-01,
                     ra2 = -1.05586262253232909814e+01, ra3 = -6.23753324503260060396e+01,
                     ra4 = -1.62396669462573470355e+02, ra5 = -1.84605092906711035994e+02,
                     ra6 = -8.12874355063065934246e+01, ra7 = -9.81432934416914548592e+00;
        const double sa1 = 1.96512716674392571292e+01, sa2 = 1.37657754143519042600e+02,
                     sa3 = 4.34565877475229228821e+02, sa4 = 6.45387271733267880336e+02,
                     sa5 = 4.29008140027567833386e+02, sa6 = 1.08635005541779435134e+02,
                     sa7 = 6.57024977031928170135e+00, sa8 = -6.04244152148580987438e-02;
        const double r1 = ra0 + s * ra1, s2 = s * s;
        const double t1 = 1.0 + s * sa1, s4 = s2 * s2;
        const double r2 = ra2 + s * ra3, s6 = s4 * s2;
        const double t2 = sa2 + s * sa3, s8 = s4 * s4;
        const double r3 = ra4 + s * ra5;
        const double t3 = sa4 + s * sa5;
        const double r4 = ra6 + s * ra7;
        const double t4 = sa6 + s * sa7;
        *R = r1 + s2 * r2 + s4 * r3 + s6 * r4;
        *S = t1 + s2 * t2 + s4 * t3 + s6 * t4 + s8 * sa8;
    } else {
        const double rb0 = -9.86494292470009928597e-03, rb1 = -7.99283237680523006574e-01,
                     rb2 = -1.77579549177547519889e+01, rb3 = -1.60636384855821916062e+02,
                     rb4 = -6.37566443368389627722e+02, rb5 = -1.02509513161107724954e+03,
                     rb6 = -4.83519191608651397019e+02;
        const double sb1 = 3.03380607434824582924e+01, sb2 = 3.25792512996573918826e+02,
                     sb3 = 1.53672958608443695994e+03, sb4 = 3.19985821950859553908e+03,
                     sb5 = 2.55305040643316442583e+03, sb6 = 4.74528541206955367215e+02,
                     sb7 = -2.24409524465858183362e+01;
        const double r1 = rb0 + s * rb1, s2 = s * s;
        const double t1 = 1.0 + s * sb1, s4 = s2 * s2;
        const double r2 = rb2 + s * rb3, s6 = s4 * s2;
        const double t2 = sb2 + s * sb3;
        const double r3 = rb4 + s * rb5;
        const double t3 = sb4 + s * sb5;
        const double t4 = sb6 + s * sb7;
        *R = r1 + s2 * r2 + s4 * r3 + s6 * rb6;
        *S = t1 + s2 * t2 + s4 * t3 + s6 * t4;
    }
}

// exp(-x^2 - 0.5625) * exp((z - x)(z + x) + R/S) with z = x truncated to its
// high word, for x = |argument| >= 1.25.
__device__ __forceinline__ double sigp_erf_expterm(double ax, double R, double S) {
    const double z = sigp_dbl(sigp_bits(ax) & 0xffffffff00000000ull);
    return sigp_exp(-z * z - 0.5625) * sigp_exp((z - ax) * (z + ax) + R / S);
}

__device__ __forceinline__ static double sigp_erf(double x) {
    const int32_t hx = (int32_t)(sigp_bits(x) >> 32);
    const int32_t ix = hx & 0x7fffffff;
    if (ix >= 0x7ff00000) return (double)(1 - (int32_t)(((uint32_t)hx >> 31) << 1)) + 1.0 / x;
    if (ix < 0x3feb0000) {                                  // |x| < 0.84375
        if (ix < 0x3e300000) {                              // |x| < 2^-28
            if (ix < 0x00800000) return 0.0625 * (16.0 * x + (16.0 * SIGP_EFX) * x);
            return x + SIGP_EFX * x;
        }
        double r, s;
        sigp_erf_small(x * x, &r, &s);
        const double y = r / s;
        return x + x * y;
    }
    if (ix < 0x3ff40000) {                                  // 0.84375 <= |x| < 1.25
        double p, q;
        sigp_erf_mid(fabs(x) - 1.0, &p, &q);
        return hx >= 0 ? SIGP_ERX + p / q : -SIGP_ERX - p / q;
    }
    if (ix >= 0x40180000) return hx >= 0 ? 1.0 - 1e-300 : 1e-300 - 1.0;   // |x| >= 6
    const double ax = fabs(x);
    double R, S;
    sigp_erf_tail(1.0 / (ax * ax), ix < 0x4006DB6E, &R, &S);
    const double r = sigp_erf_expterm(ax, R, S);
    return hx >= 0 ? 1.0 - r / ax : r / ax - 1.0;
}

__device__ __forceinline__ static double sigp_erfc(double x) {
    const int32_t hx = (int32_t)(sigp_bits(x) >> 32);
    const int32_t ix = hx & 0x7fffffff;
    if (ix >= 0x7ff00000) return (double)(((uint32_t)hx >> 31) << 1) + 1.0 / x;
    if (ix < 0x3feb0000) {                                  // |x| < 0.84375
        if (ix < 0x3c700000) return 1.0 - x;                // |x| < 2^-56
        double r, s;
        sigp_erf_small(x * x, &r, &s);
        const double y = r / s;
        if (hx < 0x3fd00000) return 1.0 - (x + x * y);      // x < 1/4
        double t = x * y;
        t += (x - 0.5);
        return 0.5 - t;
    }
    if (ix < 0x3ff40000) {                                  // 0.84375 <= |x| < 1.25
        double p, q;
        sigp_erf_mid(fabs(x) - 1.0, &p, &q);
        if (hx >= 0) { const double z = 1.0 - SIGP_ERX; return z - p / q; }
        const double z = SIGP_ERX + p / q;
        return 1.0 + z;
    }
    if (ix < 0x403c0000) {                                  // |x| < 28
        const bool near = ix < 0x4006DB6D;
        if (!near && hx < 0 && ix >= 0x40180000) return 2.0 - 1e-300;   // x < -6
        const double ax = fabs(x);
        double R, S;
        sigp_erf_tail(1.0 / (ax * ax), near, &R, &S);
        const double r = sigp_erf_expterm(ax, R, S);
        return hx > 0 ? r / ax : 2.0 - r / ax;
    }
    return hx > 0 ? 1e-300 * 1e-300 : 2.0 - 1e-300;
}

// ---- pow for the rmod operator ---------------------------------------------
// glibc's pow, restated in sigp_pow.h (shared with the host build that the
// CPU tests sweep against libm), over this file's 2^(k/128) exp table.
__device__ __constant__ static const double sigp_pow_log_tab[384] = SIGP_POW_LOG_TAB_INIT;

__device__ __forceinline__ static double sigp_pow_glibc(double x, double y) {
    return sigp_pow::pow(x, y, sigp_pow_log_tab, sigp_exp_tab);
}
