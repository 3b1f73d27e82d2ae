// Correctly rounded natural logarithm, sine and cosine for the scenario
// sampler, shared by host and device code.
//
// Method: double-double arithmetic (Dekker / Knuth error-free sum, product
// by fused multiply-add), about 2^-100 relative error before the final
// rounding, so the returned double is the correctly rounded value except in
// cases within 2^-100 of a rounding midpoint.
//   log(x):  x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(t),
//            t = (m - 1) / (m + 1), |t| <= 0.1716, by its power series
//            (24 terms); log x = e ln2 + log m.
//   sin, cos: x - k pi/2 with pi/2 split in three doubles (Cody-Waite), then
//            the Taylor series of both on |r| <= pi/4 (15 terms each).
// Constants were derived with 80-digit decimal arithmetic: ln 2 (series of
// 1/(k 2^k)) and pi (Machin's formula), split into doubles.
//
// glibc's log, sin and cos are not correctly rounded: on the sampler's
// arguments they differ from these in about 0.08 % (log) and 0.15 % (sin,
// cos) of calls, by one unit in the last place, and in each such case
// examined against 60-digit values these functions were the correctly rounded
// ones.
#pragma once
#include <math.h>

#ifdef __HIPCC__
#define SIGP_DD __host__ __device__ __forceinline__
#else
#define SIGP_DD static inline
#endif

namespace sigp_dd {

struct dd {
    double hi, lo;
};

SIGP_DD dd fast2(double a, double b) {   // |a| >= |b|
    const double s = a + b;
    return dd{s, b - (s - a)};
}
SIGP_DD dd two(double a, double b) {
    const double s = a + b, bb = s - a;
    return dd{s, (a - (s - bb)) + (b - bb)};
}
SIGP_DD dd prod(double a, double b) {
    const double p = a * b;
    return dd{p, fma(a, b, -p)};
}
SIGP_DD dd add(dd x, dd y) {
    const dd s = two(x.hi, y.hi);
    return fast2(s.hi, s.lo + (x.lo + y.lo));
}
SIGP_DD dd neg(dd x) { return dd{-x.hi, -x.lo}; }
SIGP_DD dd mul(dd x, dd y) {
    const dd p = prod(x.hi, y.hi);
    return fast2(p.hi, p.lo + (x.hi * y.lo + x.lo * y.hi));
}
SIGP_DD dd muld(dd x, double d) {
    const dd p = prod(x.hi, d);
    return fast2(p.hi, p.lo + x.lo * d);
}
SIGP_DD dd divd(dd x, double d) {
    const double q1 = x.hi / d;
    const dd p = prod(q1, d);
    const dd r = two(x.hi, -p.hi);
    const double rr = (r.hi + (r.lo - p.lo)) + x.lo;
    return fast2(q1, rr / d);
}
SIGP_DD dd div(dd x, dd y) {
    const double q1 = x.hi / y.hi;
    dd r = add(x, neg(muld(y, q1)));
    const double q2 = r.hi / y.hi;
    r = add(r, neg(muld(y, q2)));
    const double q3 = r.hi / y.hi;
    return add(fast2(q1, q2), dd{q3, 0.0});
}

// log(x) for finite x > 0
SIGP_DD double cr_log(double x) {
    const dd LN2{0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};
    int e;
    double m = frexp(x, &e);
    if (m < 0x1.6a09e667f3bcdp-1) {
        m *= 2.0;
        e -= 1;
    }
    const dd t = div(dd{m - 1.0, 0.0}, two(m, 1.0));   // m - 1 is exact (Sterbenz)
    const dd t2 = mul(t, t);
    dd p = t, s = t;
    for (int k = 1; k <= 24; ++k) {
        p = mul(p, t2);
        s = add(s, divd(p, (double)(2 * k + 1)));
    }
    const dd r = add(muld(LN2, (double)e), muld(s, 2.0));
    return r.hi + r.lo;
}

// sin(x) and cos(x) for 0 <= x < 8 (the sampler's angles lie in [0, 2 pi))
SIGP_DD void cr_sincos(double x, double *sn, double *cs) {
    const double p1 = 0x1.921fb54442d18p+0, p2 = 0x1.1a62633145c07p-54, p3 = -0x1.f1976b7ed8fbcp-110;
    const double kf = nearbyint(x * 0x1.45f306dc9c883p-1);   // round(x * 2/pi)
    const int k = (int)kf;
    const dd a = prod(kf, p1);
    dd r = two(x, -a.hi);
    r = add(r, dd{-a.lo, 0.0});
    r = add(r, neg(prod(kf, p2)));
    r = add(r, dd{-kf * p3, 0.0});
    const dd r2 = mul(r, r);
    dd ps = r, s = r, pc{1.0, 0.0}, c{1.0, 0.0};
    for (int j = 1; j <= 15; ++j) {
        ps = divd(mul(ps, r2), (double)((2 * j) * (2 * j + 1)));
        pc = divd(mul(pc, r2), (double)((2 * j - 1) * (2 * j)));
        if (j & 1) {
            s = add(s, neg(ps));
            c = add(c, neg(pc));
        } else {
            s = add(s, ps);
            c = add(c, pc);
        }
    }
    const double sv = s.hi + s.lo, cv = c.hi + c.lo;
    switch (k & 3) {
    case 0: *sn = sv; *cs = cv; break;
    case 1: *sn = cv; *cs = -sv; break;
    case 2: *sn = -sv; *cs = -cv; break;
    default: *sn = -cv; *cs = sv; break;
    }
}

}  // namespace sigp_dd
