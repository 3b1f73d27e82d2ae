// Terminal features of a dispatching decision, evaluated for one candidate
// customer `i` by every lane of a wavefront at once.  A wavefront simulates
// one rule over 64 scenarios: the candidate index i and the neighbour index j
// of every scan are wave-uniform, so instance data indexed by (i, j) is read
// with scalar loads, while the lane-private state (current POI, clock,
// visited set) lives in VGPRs.  The visited set and the per-decision
// feasibility sets are 32-bit words, NW words per lane.
//
// Each function restates one reference terminal in its operation order
// (engine/sim_terms.hpp, file:line given per function).  Two exact
// substitutions are used throughout:
//   * score[j] / dist(a, j) is read from the instance table rk (the same
//     IEEE division done once per instance, engine/sim_terms.hpp:31 etc.);
//   * "feasible from the current POI now" is read from the per-decision
//     bitmask fnow, which holds the identical test of pick_best
//     (engine/sim_core.hpp:78-82), and the close-time-only test from cnow.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "sigp_libm.h"

namespace sigp {

// Read-only instance memory in the constant address space: uniform loads
// become scalar (SMEM) loads.
typedef const __attribute__((address_space(4))) int32_t *cip;
typedef const __attribute__((address_space(4))) char *cbase;

// An instance table: a 32-bit byte offset from the base address of the
// allocation that holds it.  Kernels give every table of one allocation the
// same base value (dev_of, tabs_of), so the tables cost one scalar register
// each plus one pair per allocation, instead of a pair each, and scalar
// loads take the SGPR-offset form.  An instance's two allocations stay far
// below 4 GiB (n <= 512: about 8.5 MB and 25 MB).
struct cdp {
    cbase base;
    uint32_t off;
    __host__ __device__ __forceinline__ cdp operator+(uint32_t k) const { return cdp{base, off + 8u * k}; }
    __device__ __forceinline__ double operator[](uint32_t i) const {
        return *(const __attribute__((address_space(4))) double *)(base + (off + 8u * i));
    }
};
struct cid {
    cbase base;
    uint32_t off;
    __device__ __forceinline__ int32_t operator[](uint32_t i) const {
        return *(const __attribute__((address_space(4))) int32_t *)(base + (off + 4u * i));
    }
};
typedef const __attribute__((address_space(4))) uint64_t *cup;

constexpr double kSigma = 0.2;                  // engine/sim_terms.hpp:15
constexpr double kSqrt2 = 1.4142135623730951;   // engine/sim_terms.hpp:16

struct DevInst {
    int n, m, ld, K;
    double tmax;
    cdp dist;    // (n+1)^2, dist[a*ld + b]
    cdp distT;   // its transpose (per-lane row gathers become column reads)
    cdp rk;      // (n+1) x n: score[q] / max(dist[a][q+1], 1e-10)
    cdp rkT;     // n x (n+1) transpose of rk
    cdp open, close, dur, d2d, score;
    cdp durrob;  // dur * (1 + sigma)
    cdp sdv;     // 0.2 * dur
    cid dcls;    // duration class of each customer
    cdp pw2;     // K x K: pow(sd_c*sd_c + sd_j*sd_j, 0.5)
    cdp pw1;     // K: pow(0.0 + sd_j*sd_j, 0.5)
    // ATRISK saturation thresholds (engine/sim_terms.hpp:221-252): for a
    // divisor s, norm_cdf(d / s) is exactly 1 for d >= hi(s) and exactly 0
    // for d <= lo(s) -- erf returns +-1 for arguments of magnitude >= 6
    // and (d / s) * 0.7071067811865476 is monotone in d.  Per duration class
    // for s = 0.2 * dur (atp1 / atp0), per pw2 entry (atq2: K x K pairs hi,
    // lo) and per pw1 entry (atq1: K pairs hi, lo).
    cdp atp1, atp0;
    cdp atq2, atq1;
    // ATRISK's p_bud of customer j for a mean arrival before j opens
    // (engine/sim_terms.hpp:243-252 with mean_start = open_j, var_start = 0):
    // it depends on j alone, so the host computes it once per instance with
    // the C library's erf -- the reference's own value.
    cdp atpb0;
    cdp zp;      // 10 z-statistics
};

// A kernel's view of an instance record: every table on the base of the
// first (they share one allocation).
__device__ __forceinline__ DevInst dev_of(const DevInst &src) {
    DevInst I = src;
    const cbase b = src.dist.base;
    I.dist.base = b; I.distT.base = b; I.rk.base = b; I.rkT.base = b;
    I.open.base = b; I.close.base = b; I.dur.base = b; I.d2d.base = b; I.score.base = b;
    I.durrob.base = b; I.sdv.base = b; I.dcls.base = b; I.pw2.base = b; I.pw1.base = b;
    I.atp1.base = b; I.atp0.base = b; I.atq2.base = b; I.atq1.base = b; I.atpb0.base = b; I.zp.base = b;
    return I;
}

template <int NW>
__device__ __forceinline__ uint32_t wsel(const uint32_t (&a)[NW], int w) {
    // An and/or chain rather than a select chain: selects between loads of
    // a[k] fold into one load from a computed address, which keeps the
    // array in scratch memory instead of registers.
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) r |= a[k] & (0u - (uint32_t)(w == k));
    return r;
}

template <int NW>
__device__ __forceinline__ void wset(uint32_t (&a)[NW], int idx) {
    const int w = idx >> 5;
    const uint32_t b = 1u << (idx & 31);
#pragma unroll
    for (int k = 0; k < NW; ++k) a[k] |= (w == k) ? b : 0u;
}

template <int NW>
__device__ __forceinline__ bool wtest(const uint32_t (&a)[NW], int idx) {
    return (wsel<NW>(a, idx >> 5) >> (idx & 31)) & 1u;
}

// Common neighbour test of the NS family from POI i+1 at time f:
// t >= 1e-10, arrival <= close, start + dur + d2d <= tmax.
// `k1` scales the neighbour's duration (1 for plain, 1 + m*sigma for robust).

// engine/sim_terms.hpp:18-34 (t_ns) with rk for score/t.
template <int NW>
__device__ __forceinline__ double t_ns(const DevInst &I, int i, double fin, const uint32_t (&vis)[NW]) {
    const cdp drow = I.dist + (size_t)(i + 1) * I.ld + 1;
    const cdp rrow = I.rk + (size_t)(i + 1) * I.n;
    double total = 0.0;
    uint32_t vw = 0;
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) vw = wsel<NW>(vis, j >> 5);
        const double t = drow[j];
        if (j == i || t < 1e-10) continue;
        if ((vw >> (j & 31)) & 1u) continue;
        const double arr = fin + t;
        if (arr > I.close[j]) continue;
        const double st = arr > I.open[j] ? arr : I.open[j];
        if (st + I.dur[j] + I.d2d[j] > I.tmax) continue;
        total += rrow[j];
    }
    return total;
}

// engine/sim_terms.hpp:56-74 (t_ns_rob, general = false) and :295-314
// (t_ns_rob_m, general = true, ms = m * sigma).
template <int NW>
__device__ __forceinline__ double t_ns_rob_m(const DevInst &I, int i, double fin, double dur_i,
                             const uint32_t (&vis)[NW], double msig, bool general) {
    const double ms = general ? msig * kSigma : kSigma;
    const double frob = fin + dur_i * ms;
    const double k1 = 1.0 + ms;
    const cdp drow = I.dist + (size_t)(i + 1) * I.ld + 1;
    const cdp rrow = I.rk + (size_t)(i + 1) * I.n;
    double total = 0.0;
    uint32_t vw = 0;
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) vw = wsel<NW>(vis, j >> 5);
        const double t = drow[j];
        if (j == i || t < 1e-10) continue;
        if ((vw >> (j & 31)) & 1u) continue;
        const double arr = frob + t;
        if (arr > I.close[j]) continue;
        const double st = arr > I.open[j] ? arr : I.open[j];
        if (st + I.dur[j] * k1 + I.d2d[j] > I.tmax) continue;
        total += rrow[j];
    }
    return total;
}

// engine/sim_terms.hpp:76-93
template <int NW>
__device__ __forceinline__ double t_maxn(const DevInst &I, int i, double fin, const uint32_t (&vis)[NW]) {
    const cdp drow = I.dist + (size_t)(i + 1) * I.ld + 1;
    const cdp rrow = I.rk + (size_t)(i + 1) * I.n;
    double best = 0.0;
    uint32_t vw = 0;
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) vw = wsel<NW>(vis, j >> 5);
        const double t = drow[j];
        if (j == i || t < 1e-10) continue;
        if ((vw >> (j & 31)) & 1u) continue;
        const double arr = fin + t;
        if (arr > I.close[j]) continue;
        const double st = arr > I.open[j] ? arr : I.open[j];
        if (st + I.dur[j] + I.d2d[j] > I.tmax) continue;
        const double v = rrow[j];
        if (v > best) best = v;
    }
    return best;
}

// engine/sim_terms.hpp:95-127 (rob = false) and :156-190 (rob = true).  The
// second leg starts from the lane's best first neighbour: its row is a
// per-lane gather through the transposed tables.
template <int NW>
__device__ __forceinline__ double t_reach2x(const DevInst &I, int i, double fin, double dur_i,
                            const uint32_t (&vis)[NW], bool rob) {
    const double k1 = 1.0 + kSigma;
    const double f0 = rob ? fin + dur_i * kSigma : fin;
    const cdp drow = I.dist + (size_t)(i + 1) * I.ld + 1;
    const cdp rrow = I.rk + (size_t)(i + 1) * I.n;
    double bv = 0.0, fj = 0.0;
    int bj = -1;
    uint32_t vw = 0;
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) vw = wsel<NW>(vis, j >> 5);
        const double t = drow[j];
        if (j == i || t < 1e-10) continue;
        if ((vw >> (j & 31)) & 1u) continue;
        const double arr = f0 + t;
        if (arr > I.close[j]) continue;
        const double st = arr > I.open[j] ? arr : I.open[j];
        const double lim = rob ? st + I.dur[j] * k1 + I.d2d[j] : st + I.dur[j] + I.d2d[j];
        if (lim > I.tmax) continue;
        const double v = rrow[j];
        if (v > bv) {
            bv = v;
            bj = j;
            fj = rob ? st + I.dur[j] * k1 : st + I.dur[j];
        }
    }
    if (bj < 0) return 0.0;
    // second leg: dist[bj+1][k+1] = distT[(k+1)*ld + bj+1], rk[bj+1][k] = rkT[k*ld + bj+1]
    double bv2 = 0.0, s2 = 0.0;
    const int pj = bj + 1;
    uint32_t vw2 = 0;
    for (int k = 0; k < I.n; ++k) {
        if ((k & 31) == 0) vw2 = wsel<NW>(vis, k >> 5);
        if (k == i || k == bj || ((vw2 >> (k & 31)) & 1u)) continue;
        const double t = I.distT[(size_t)(k + 1) * I.ld + pj];
        if (t < 1e-10) continue;
        const double arr = fj + t;
        if (arr > I.close[k]) continue;
        const double st = arr > I.open[k] ? arr : I.open[k];
        const double lim = rob ? st + I.dur[k] * k1 + I.d2d[k] : st + I.dur[k] + I.d2d[k];
        if (lim > I.tmax) continue;
        const double v = I.rkT[(size_t)k * I.ld + pj];
        if (v > bv2) { bv2 = v; s2 = I.score[k]; }
    }
    return I.score[bj] + s2;
}

// engine/sim_terms.hpp:36-54: the "now" half of each test is fnow[j].
template <int NW>
__device__ __forceinline__ double t_regret(const DevInst &I, int i, double fin_i, const uint32_t (&vis)[NW],
                           const uint32_t (&fnow)[NW]) {
    const cdp drow = I.dist + (size_t)(i + 1) * I.ld + 1;
    double total = 0.0;
    uint32_t vw = 0, fw = 0;
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) { vw = wsel<NW>(vis, j >> 5); fw = wsel<NW>(fnow, j >> 5); }
        if (j == i) continue;
        const uint32_t b = 1u << (j & 31);
        if ((vw & b) || !(fw & b)) continue;
        const double aa = fin_i + drow[j];
        if (aa > I.close[j]) { total += I.score[j]; continue; }
        const double sa = aa > I.open[j] ? aa : I.open[j];
        if (sa + I.dur[j] + I.d2d[j] > I.tmax) total += I.score[j];
    }
    return total;
}

// engine/sim_terms.hpp:129-154: the first scan runs from the lane's current
// POI (column gathers through distT / rkT).
template <int NW>
__device__ __forceinline__ double t_skiploss(const DevInst &I, int i, int cur, double ct,
                             const uint32_t (&vis)[NW]) {
    double bv = 0.0, fq = 0.0;
    int bq = -1;
    uint32_t vw = 0;
    for (int q = 0; q < I.n; ++q) {
        if ((q & 31) == 0) vw = wsel<NW>(vis, q >> 5);
        if (q == i || ((vw >> (q & 31)) & 1u)) continue;
        const double t = I.distT[(size_t)(q + 1) * I.ld + cur];
        if (t < 1e-10) continue;
        const double arr = ct + t;
        if (arr > I.close[q]) continue;
        const double st = arr > I.open[q] ? arr : I.open[q];
        if (st + I.dur[q] + I.d2d[q] > I.tmax) continue;
        const double v = I.rkT[(size_t)q * I.ld + cur];
        if (v > bv) { bv = v; bq = q; fq = st + I.dur[q]; }
    }
    if (bq < 0) return 0.0;
    const double tqi = I.dist[(size_t)(bq + 1) * I.ld + (i + 1)];
    const double ai = fq + tqi;
    if (ai > I.close[i]) return I.score[i];
    const double si = ai > I.open[i] ? ai : I.open[i];
    if (si + I.dur[i] + I.d2d[i] > I.tmax) return I.score[i];
    return 0.0;
}

// engine/sim_terms.hpp:192-213
template <int NW>
__device__ __forceinline__ double t_nstw(const DevInst &I, int i, double fin, const uint32_t (&vis)[NW]) {
    const cdp drow = I.dist + (size_t)(i + 1) * I.ld + 1;
    const cdp rrow = I.rk + (size_t)(i + 1) * I.n;
    double total = 0.0;
    uint32_t vw = 0;
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) vw = wsel<NW>(vis, j >> 5);
        const double t = drow[j];
        if (j == i || t < 1e-10) continue;
        if ((vw >> (j & 31)) & 1u) continue;
        const double arr = fin + t;
        if (arr > I.close[j]) continue;
        const double st = arr > I.open[j] ? arr : I.open[j];
        if (st + I.dur[j] + I.d2d[j] > I.tmax) continue;
        double wait = I.open[j] - arr;
        if (wait < 0.0) wait = 0.0;
        const double win = I.close[j] - I.open[j];
        double w = win > 1e-10 ? 1.0 - wait / win : 1.0;
        if (w < 0.0) w = 0.0;
        total += rrow[j] * w;
    }
    return total;
}

// engine/sim_terms.hpp:215-219
__device__ __forceinline__ double t_rand(int i) {
    long long x = (((long long)i + 1) * 1103515245LL + 12345LL) & 0x7FFFFFFFLL;
    x = (x * 1103515245LL + 12345LL) & 0x7FFFFFFFLL;
    return (double)(x % 100000LL) / 1000.0;
}

// engine/sim_terms.hpp:221-223
__device__ __forceinline__ double norm_cdf(double z) {
    return 0.5 * (1.0 + sigp_erf(z * 0.7071067811865476));
}

// engine/sim_terms.hpp:225-257.  "arrives before close from the current POI"
// is cnow[j]; pow(var, 0.5) comes from the per-instance duration-class table
// (the identical libm call, made once per instance).
template <int NW>
__device__ __forceinline__ double t_eloss(const DevInst &I, int c, double fin_c, const uint32_t (&vis)[NW],
                          const uint32_t (&cnow)[NW]) {
    const double sdc = I.sdv[c];
    const double sdc2 = sdc * sdc;
    const cdp drow = I.dist + (size_t)(c + 1) * I.ld + 1;
    const cdp pwrow = I.pw2 + (size_t)I.dcls[c] * I.K;
    double eloss = 0.0;
    uint32_t vw = 0, cw = 0;
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) { vw = wsel<NW>(vis, j >> 5); cw = wsel<NW>(cnow, j >> 5); }
        if (j == c) continue;
        const uint32_t b = 1u << (j & 31);
        if ((vw & b) || !(cw & b)) continue;
        const double ma = fin_c + drow[j];
        const double pw = sdc > 1e-9 ? norm_cdf((I.close[j] - ma) / sdc)
                                     : (ma <= I.close[j] ? 1.0 : 0.0);
        const bool late = ma > I.open[j];
        const double ms = late ? ma : I.open[j];
        const double vs = late ? sdc2 : 0.0;
        const double sdj = I.sdv[j];
        const double vb = vs + sdj * sdj;
        const double mb = ms + I.dur[j] + I.d2d[j];
        double pb;
        if (vb > 1e-18) {
            const int kj = I.dcls[j];
            const double root = late ? pwrow[kj] : I.pw1[kj];
            pb = norm_cdf((I.tmax - mb) / root);
        } else {
            pb = mb <= I.tmax ? 1.0 : 0.0;
        }
        eloss += I.score[j] * (1.0 - pw * pb);
    }
    return eloss;
}

// engine/sim_terms.hpp:259-275 (half = false: score/t^2) and :277-293
// (half = true: score/sqrt(t)).
template <int NW>
__device__ __forceinline__ double t_ns_alpha(const DevInst &I, int i, double fin, const uint32_t (&vis)[NW],
                             bool half) {
    const cdp drow = I.dist + (size_t)(i + 1) * I.ld + 1;
    double total = 0.0;
    uint32_t vw = 0;
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) vw = wsel<NW>(vis, j >> 5);
        const double t = drow[j];
        if (j == i || t < 1e-10) continue;
        if ((vw >> (j & 31)) & 1u) continue;
        const double arr = fin + t;
        if (arr > I.close[j]) continue;
        const double st = arr > I.open[j] ? arr : I.open[j];
        if (st + I.dur[j] + I.d2d[j] > I.tmax) continue;
        total += half ? I.score[j] / sqrt(t) : I.score[j] / (t * t);
    }
    return total;
}

// engine/sim_terms.hpp:316-334
template <int NW>
__device__ __forceinline__ double t_nnfwd(const DevInst &I, int i, double fin, const uint32_t (&vis)[NW]) {
    const cdp drow = I.dist + (size_t)(i + 1) * I.ld + 1;
    double dmin = -1.0;
    uint32_t vw = 0;
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) vw = wsel<NW>(vis, j >> 5);
        const double t = drow[j];
        if (j == i || t < 1e-10) continue;
        if ((vw >> (j & 31)) & 1u) continue;
        const double arr = fin + t;
        if (arr > I.close[j]) continue;
        const double st = arr > I.open[j] ? arr : I.open[j];
        if (st + I.dur[j] + I.d2d[j] > I.tmax) continue;
        if (dmin < 0.0 || t < dmin) dmin = t;
    }
    if (dmin < 0.0) return 0.0;
    return 1.0 / (1.0 + dmin);
}

// engine/sim_terms.hpp:336-359
template <int NW>
__device__ __forceinline__ double t_nnk3(const DevInst &I, int i, double fin, const uint32_t (&vis)[NW]) {
    const cdp drow = I.dist + (size_t)(i + 1) * I.ld + 1;
    double d1 = -1.0, d2 = -1.0, d3 = -1.0;
    uint32_t vw = 0;
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) vw = wsel<NW>(vis, j >> 5);
        const double t = drow[j];
        if (j == i || t < 1e-10) continue;
        if ((vw >> (j & 31)) & 1u) continue;
        const double arr = fin + t;
        if (arr > I.close[j]) continue;
        const double st = arr > I.open[j] ? arr : I.open[j];
        if (st + I.dur[j] + I.d2d[j] > I.tmax) continue;
        if (d1 < 0.0 || t < d1) { d3 = d2; d2 = d1; d1 = t; }
        else if (d2 < 0.0 || t < d2) { d3 = d2; d2 = t; }
        else if (d3 < 0.0 || t < d3) { d3 = t; }
    }
    if (d1 < 0.0) return 0.0;
    double tot = d1, k = 1.0;
    if (d2 >= 0.0) { tot += d2; k += 1.0; }
    if (d3 >= 0.0) { tot += d3; k += 1.0; }
    return 1.0 / (1.0 + tot / k);
}

// engine/sim_terms.hpp:361-381 (dist(current, j) is a column gather).
template <int NW>
__device__ __forceinline__ double t_detour(const DevInst &I, int i, int cur, double fin, double ta_i,
                           const uint32_t (&vis)[NW]) {
    const cdp drow = I.dist + (size_t)(i + 1) * I.ld + 1;
    double total = 0.0, cnt = 0.0;
    uint32_t vw = 0;
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) vw = wsel<NW>(vis, j >> 5);
        const double t = drow[j];
        if (j == i || t < 1e-10) continue;
        if ((vw >> (j & 31)) & 1u) continue;
        const double arr = fin + t;
        if (arr > I.close[j]) continue;
        const double st = arr > I.open[j] ? arr : I.open[j];
        if (st + I.dur[j] + I.d2d[j] > I.tmax) continue;
        total += ta_i + t - I.distT[(size_t)(j + 1) * I.ld + cur];
        cnt += 1.0;
    }
    if (cnt == 0.0) return 0.0;
    return total / cnt;
}

// engine/sim_terms.hpp:383-401
template <int NW>
__device__ __forceinline__ double t_scorerank(const DevInst &I, int i, double fin, const uint32_t (&vis)[NW]) {
    const cdp drow = I.dist + (size_t)(i + 1) * I.ld + 1;
    const double si = I.score[i];
    double below = 0.0, tot = 0.0;
    uint32_t vw = 0;
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) vw = wsel<NW>(vis, j >> 5);
        const double t = drow[j];
        if (j == i || t < 1e-10) continue;
        if ((vw >> (j & 31)) & 1u) continue;
        const double arr = fin + t;
        if (arr > I.close[j]) continue;
        const double st = arr > I.open[j] ? arr : I.open[j];
        if (st + I.dur[j] + I.d2d[j] > I.tmax) continue;
        tot += 1.0;
        if (I.score[j] < si) below += 1.0;
    }
    if (tot == 0.0) return 1.0;
    return below / tot;
}

// engine/sim_terms.hpp:405-441 (rob = false) and :443-479 (rob = true):
// K greedy waypoints from candidate c (t0 = start_c + dur_c [* (1+sigma)]).
// The first step scans row c+1 (uniform); later steps scan the lane's own
// waypoint row through the transposed tables.
template <int NW>
__device__ __forceinline__ double t_reach_k(const DevInst &I, int c, double st_c, const uint32_t (&vis)[NW],
                            int K, bool rob) {
    const double k1 = 1.0 + kSigma;
    uint32_t scr[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) scr[w] = vis[w];
    wset<NW>(scr, c);
    int cur = c + 1;
    double t = rob ? st_c + I.dur[c] * k1 : st_c + I.dur[c];
    double total = 0.0;
    for (int step = 0; step < K; ++step) {
        int best = -1;
        double bv = -1.0;
        uint32_t sw = 0;
        if (step == 0) {
            const cdp drow = I.dist + (size_t)cur * I.ld + 1;
            const cdp rrow = I.rk + (size_t)cur * I.n;
            for (int q = 0; q < I.n; ++q) {
                if ((q & 31) == 0) sw = wsel<NW>(scr, q >> 5);
                if ((sw >> (q & 31)) & 1u) continue;
                const double ta = drow[q];
                const double arr = t + ta;
                if (arr > I.close[q]) continue;
                const double st = arr > I.open[q] ? arr : I.open[q];
                const double lim = rob ? st + I.dur[q] * k1 + I.d2d[q] : st + I.dur[q] + I.d2d[q];
                if (lim > I.tmax) continue;
                const double v = rrow[q];
                if (v > bv) { bv = v; best = q; }
            }
        } else {
            for (int q = 0; q < I.n; ++q) {
                if ((q & 31) == 0) sw = wsel<NW>(scr, q >> 5);
                if ((sw >> (q & 31)) & 1u) continue;
                const double ta = I.distT[(size_t)(q + 1) * I.ld + cur];
                const double arr = t + ta;
                if (arr > I.close[q]) continue;
                const double st = arr > I.open[q] ? arr : I.open[q];
                const double lim = rob ? st + I.dur[q] * k1 + I.d2d[q] : st + I.dur[q] + I.d2d[q];
                if (lim > I.tmax) continue;
                const double v = I.rkT[(size_t)q * I.ld + cur];
                if (v > bv) { bv = v; best = q; }
            }
        }
        if (best < 0) break;
        const double ta = I.dist[(size_t)cur * I.ld + best + 1];
        const double arr = t + ta;
        const double st = arr > I.open[best] ? arr : I.open[best];
        t = rob ? st + I.dur[best] * k1 : st + I.dur[best];
        cur = best + 1;
        total += I.score[best];
        wset<NW>(scr, best);
    }
    return total;
}

// engine/sim_terms.hpp:481-519
template <int NW>
__device__ __forceinline__ double t_nnchain(const DevInst &I, int c, double st_c, const uint32_t (&vis)[NW], int K) {
    uint32_t scr[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) scr[w] = vis[w];
    wset<NW>(scr, c);
    int cur = c + 1;
    double t = st_c + I.dur[c];
    double total_d = 0.0;
    int done = 0;
    for (int step = 0; step < K; ++step) {
        int best = -1;
        double bd = -1.0;
        uint32_t sw = 0;
        for (int q = 0; q < I.n; ++q) {
            if ((q & 31) == 0) sw = wsel<NW>(scr, q >> 5);
            if ((sw >> (q & 31)) & 1u) continue;
            const double ta = I.distT[(size_t)(q + 1) * I.ld + cur];
            if (ta < 1e-10) continue;
            const double arr = t + ta;
            if (arr > I.close[q]) continue;
            const double st = arr > I.open[q] ? arr : I.open[q];
            if (st + I.dur[q] + I.d2d[q] > I.tmax) continue;
            if (bd < 0.0 || ta < bd) { bd = ta; best = q; }
        }
        if (best < 0) break;
        const double arr = t + bd;
        const double st = arr > I.open[best] ? arr : I.open[best];
        t = st + I.dur[best];
        cur = best + 1;
        total_d += bd;
        done++;
        wset<NW>(scr, best);
    }
    if (done == 0) return 0.0;
    return 1.0 / (1.0 + total_d / done);
}

// engine/sim_terms.hpp:533-580: FRAGCNT, FRAGSCORE, DOWNSLACK, PBUST.
// feas_now(q) is fnow[q]; feas_after(q) is tested from POI c+1 at fin_c.
template <int NW>
__device__ __forceinline__ void struct_feats(const DevInst &I, int c, double fin_c, const uint32_t (&vis)[NW],
                             const uint32_t (&fnow)[NW], double out[4]) {
    const cdp drow = I.dist + (size_t)(c + 1) * I.ld + 1;
    const cdp rrow = I.rk + (size_t)(c + 1) * I.n;
    double fcnt = 0.0, fscore = 0.0, dslack = 1e18, bv = -1.0;
    int bq = -1;
    uint32_t vw = 0, fw = 0;
    for (int q = 0; q < I.n; ++q) {
        if ((q & 31) == 0) { vw = wsel<NW>(vis, q >> 5); fw = wsel<NW>(fnow, q >> 5); }
        const uint32_t b = 1u << (q & 31);
        if (q == c || (vw & b)) continue;
        const bool now = (fw & b) != 0;
        const double ta = drow[q];
        const double arr = fin_c + ta;
        bool after = !(arr > I.close[q]);
        const double st = arr > I.open[q] ? arr : I.open[q];
        after = after && !(st + I.dur[q] + I.d2d[q] > I.tmax);
        if (now && !after) { fcnt += 1.0; fscore += I.score[q]; }
        if (after) {
            const double slack = I.close[q] - arr;
            if (slack < dslack) dslack = slack;
            if (ta > 1e-10) {
                const double v = rrow[q];
                if (v > bv) { bv = v; bq = q; }
            }
        }
    }
    if (dslack > 1e17) dslack = 0.0;
    double pbust = 0.0;
    if (bq >= 0) {
        const double ab = fin_c + I.distT[(size_t)(bq + 1) * I.ld + (c + 1)];
        const double sd = kSigma * I.dur[c];
        if (sd > 1e-12) {
            const double z = (I.close[bq] - ab) / (sd * kSqrt2);
            pbust = 0.5 * sigp_erfc(z);
        } else {
            pbust = ab > I.close[bq] ? 1.0 : 0.0;
        }
    }
    out[0] = fcnt;
    out[1] = fscore;
    out[2] = dslack;
    out[3] = pbust;
}

}  // namespace sigp
