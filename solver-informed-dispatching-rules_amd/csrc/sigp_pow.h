// pow(x, y) as glibc 2.35 computes it on x86-64 with FMA, bit for bit, for
// host and device code.  The reference's CPU simulator evaluates the rmod
// operator as pow(max(a, 0), 1 + clip(b, 0, 1)) (engine/sim_core.hpp:29-35)
// with the C library's pow; a priority that differs in its last bit can
// change an argmax, so the device must round exactly as that pow does, not
// correctly (glibc's pow is not correctly rounded).
//
// Method (ARM optimized-routines math/pow.c, used by glibc since 2.28):
//   log:  x = 2^k z with z in [0x1.69555p-1, 0x1.69555p0); a 128-entry table
//         (sigp_pow_log_table.h, built by gen_pow_log_table.py from the
//         construction the method describes) gives 1/c and log c for z's
//         subinterval; r = z/c - 1 is exact by one fused multiply-add and
//         log(x) = k ln2 + log c + log1p(r) is kept as hi + lo, log1p(r) by
//         a degree-8 polynomial (coefficients scaled by the powers of -1/2
//         the evaluation order uses);
//   exp:  exp(y hi + y lo) by the same 2^(k/128)-table method as sigp_exp
//         (sigp_exp_table.h), with the low part added to the reduced
//         argument.
// The fused multiply-adds sit where glibc's FMA build (-mfma, with the
// compiler's default contraction) has them; which of the placements that
// the contraction rules leave open matter was settled by sweeping against
// the host's libm: the final lo sum is fused (unfused differs once in 4e7
// calls near x = 1), the subnormal-range rescaling of exp is not (fused
// differs in 7e-5 of calls with subnormal results).
//
// Special cases follow pow.c: y = 0, +-inf, NaN; x = 0, +-inf, NaN, negative
// (integer y keeps the sign), subnormal x (normalised), |y| outside
// [2^-65, 2^63).  Overflow gives +-inf, underflow +-0.
//
// Licence of the method and coefficients: optimized-routines, Copyright (c)
// 2018 Arm Limited, SPDX-License-Identifier: MIT OR Apache-2.0 WITH
// LLVM-exception.  The code below is a restatement for this library.
//
// Compile with -ffp-contract=off: every fused multiply-add is explicit.
#pragma once
#include <math.h>
#include <stdint.h>

#ifdef __HIPCC__
#define SIGP_POW_FN __host__ __device__ __forceinline__
#else
#define SIGP_POW_FN static inline
#endif

namespace sigp_pow {

SIGP_POW_FN double asd(uint64_t u) { return __builtin_bit_cast(double, u); }
SIGP_POW_FN uint64_t asu(double x) { return __builtin_bit_cast(uint64_t, x); }

constexpr uint64_t kOff = 0x3fe6955500000000ull;
constexpr uint64_t kSignBias = 0x800ull << 7;       // sign bit, pre-shifted for the exp table index

// log(x) for x = asd(ix) normal positive (or subnormal pre-scaled by 2^52
// with its exponent field lowered by 52): returns hi, *tail = lo.
SIGP_POW_FN double log_inline(uint64_t ix, double *tail, const double *tab) {
    const double ln2hi = 0x1.62e42fefa3800p-1, ln2lo = 0x1.ef35793c76730p-45;
    const double a0 = -0x1p-1;
    const double a1 = -0x1.555555555556p-1;        // 0x1.555555555556p-2 * -2
    const double a2 = 0x1.0000000000006p-1;        // -0x1.0000000000006p-2 * -2
    const double a3 = 0x1.999999959554ep-1;        // 0x1.999999959554ep-3 * 4
    const double a4 = -0x1.555555529a47ap-1;       // -0x1.555555529a47ap-3 * 4
    const double a5 = -0x1.2495b9b4845e9p0;        // 0x1.2495b9b4845e9p-3 * -8
    const double a6 = 0x1.0002b8b263fc3p0;         // -0x1.0002b8b263fc3p-3 * -8
    const uint64_t tmp = ix - kOff;
    const int i = (int)((tmp >> 45) & 127u);
    const int k = (int)((int64_t)tmp >> 52);
    const uint64_t iz = ix - (tmp & (0xfffull << 52));
    const double z = asd(iz), kd = (double)k;
    const double invc = tab[3 * i], logc = tab[3 * i + 1], logctail = tab[3 * i + 2];
    const double r = fma(z, invc, -1.0);
    const double t1 = fma(kd, ln2hi, logc);        // exact: kd * ln2hi has spare bits
    const double t2 = t1 + r;
    const double lo1 = fma(kd, ln2lo, logctail);
    const double lo2 = t1 - t2 + r;
    const double ar = a0 * r;
    const double ar2 = r * ar;
    const double ar3 = r * ar2;
    const double hi = t2 + ar2;
    const double lo3 = fma(ar, r, -ar2);
    const double lo4 = t2 - hi + ar2;
    const double q = fma(ar2, fma(ar2, fma(r, a6, a5), fma(r, a4, a3)), fma(r, a2, a1));
    const double lo = fma(ar3, q, lo1 + lo2 + lo3 + lo4);
    const double y = hi + lo;
    *tail = hi - y + lo;
    return y;
}

// exp(x + xtail), |xtail| tiny against x; sign_bias = kSignBias negates.
SIGP_POW_FN double exp_inline(double x, double xtail, uint64_t sign_bias, const uint64_t *etab) {
    const double inv_ln2_n = 0x1.71547652b82fep0 * 128.0;
    const double shift = 0x1.8p52;
    const double neg_ln2_hi_n = -0x1.62e42fefa0000p-8;
    const double neg_ln2_lo_n = -0x1.cf79abc9e3b3ap-47;
    const double c2 = 0x1.ffffffffffdbdp-2, c3 = 0x1.555555555543cp-3;
    const double c4 = 0x1.55555cf172b91p-5, c5 = 0x1.1111167a4d017p-7;
    const uint32_t abstop = (uint32_t)(asu(x) >> 52) & 0x7ffu;
    bool large = false;
    if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {
        if (abstop - 0x3c9u >= 0x80000000u) {               // |x| < 2^-54
            const double one = 1.0 + x;
            return sign_bias ? -one : one;
        }
        if (abstop >= 0x409u) {                             // |x| >= 1024
            const double z = (asu(x) >> 63) ? 0.0 : HUGE_VAL;
            return sign_bias ? -z : z;
        }
        large = true;                                       // 512 <= |x| < 1024
    }
    double kd = fma(inv_ln2_n, x, shift);
    const uint64_t ki = asu(kd);
    kd -= shift;
    double r = fma(kd, neg_ln2_lo_n, fma(kd, neg_ln2_hi_n, x));
    r += xtail;
    const uint32_t idx = 2u * (uint32_t)(ki & 127u);
    const uint64_t top = (ki + sign_bias) << 45;
    const double tail = asd(etab[idx]);
    uint64_t sbits = etab[idx + 1] + top;
    const double r2 = r * r;
    const double tmp = fma(r2 * r2, fma(r, c5, c4), fma(r2, fma(r, c3, c2), tail + r));
    if (large) {
        if ((ki & 0x80000000ull) == 0) {                    // result may overflow
            sbits -= 1009ull << 52;
            const double scale = asd(sbits);
            return 0x1p1009 * fma(scale, tmp, scale);
        }
        sbits += 1022ull << 52;                             // result may be subnormal
        const double scale = asd(sbits);
        const double st = scale * tmp;
        double y = scale + st;
        if (fabs(y) < 1.0) {
            const double one = y < 0.0 ? -1.0 : 1.0;
            double lo = scale - y + st;
            const double hi = one + y;
            lo = one - hi + y + lo;
            y = (hi + lo) - one;
            if (y == 0.0) y = asd(sbits & 0x8000000000000000ull);
        }
        return 0x1p-1022 * y;
    }
    const double scale = asd(sbits);
    return fma(scale, tmp, scale);
}

// 0 if y is not an integer, 1 if odd, 2 if even.
SIGP_POW_FN int check_int(uint64_t iy) {
    const int e = (int)(iy >> 52 & 0x7ff);
    if (e < 0x3ff) return 0;
    if (e > 0x3ff + 52) return 2;
    if (iy & ((1ull << (0x3ff + 52 - e)) - 1)) return 0;
    if (iy & (1ull << (0x3ff + 52 - e))) return 1;
    return 2;
}

SIGP_POW_FN bool zero_inf_nan(uint64_t i) { return 2 * i - 1 >= 2 * asu(HUGE_VAL) - 1; }

SIGP_POW_FN double pow(double x, double y, const double *log_tab, const uint64_t *exp_tab) {
    uint64_t sign_bias = 0;
    uint64_t ix = asu(x);
    const uint64_t iy = asu(y);
    uint32_t topx = (uint32_t)(ix >> 52);
    const uint32_t topy = (uint32_t)(iy >> 52);
    if (topx - 0x001u >= 0x7ffu - 0x001u || (topy & 0x7ffu) - 0x3beu >= 0x43eu - 0x3beu) {
        if (zero_inf_nan(iy)) {
            if (2 * iy == 0) return 1.0;
            if (ix == asu(1.0)) return 1.0;
            if (2 * ix > 2 * asu(HUGE_VAL) || 2 * iy > 2 * asu(HUGE_VAL)) return x + y;
            if (2 * ix == 2 * asu(1.0)) return 1.0;
            if ((2 * ix < 2 * asu(1.0)) == !(iy >> 63)) return 0.0;
            return y * y;
        }
        if (zero_inf_nan(ix)) {
            double x2 = x * x;
            if ((ix >> 63) && check_int(iy) == 1) x2 = -x2;
            return (iy >> 63) ? 1.0 / x2 : x2;
        }
        if (ix >> 63) {                                     // negative finite x
            const int yint = check_int(iy);
            if (yint == 0) return (x - x) / (x - x);
            if (yint == 1) sign_bias = kSignBias;
            ix &= 0x7fffffffffffffffull;
            topx &= 0x7ffu;
        }
        if ((topy & 0x7ffu) - 0x3beu >= 0x43eu - 0x3beu) {   // |y| tiny or huge
            if (ix == asu(1.0)) return 1.0;
            if ((topy & 0x7ffu) < 0x3beu) return ix > asu(1.0) ? 1.0 + y : 1.0 - y;
            return ((ix > asu(1.0)) == (topy < 0x800u)) ? HUGE_VAL : 0.0;
        }
        if (topx == 0) {                                    // subnormal x
            ix = asu(x * 0x1p52);
            ix &= 0x7fffffffffffffffull;
            ix -= 52ull << 52;
        }
    }
    double lo;
    const double hi = log_inline(ix, &lo, log_tab);
    const double ehi = y * hi;
    const double elo = fma(y, lo, fma(y, hi, -ehi));
    return exp_inline(ehi, elo, sign_bias, exp_tab);
}

}  // namespace sigp_pow
