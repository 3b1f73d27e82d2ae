// SI-GP rule x scenario evaluator for MI355X (gfx950): kernels and the C ABI
// declared in include/sigp.h.
//
// Work layout.  One wavefront simulates one rule on 64 scenarios (lane =
// scenario, common random numbers across rules).  The rule's postfix code,
// the candidate index of every decision and the neighbour index of every
// terminal scan are wave-uniform, so instance data is read through the scalar
// cache and the interpreter stack is indexed with a uniform stack pointer
// (s_set_gpr_idx).  Lanes hold their own route state: current POI, clock,
// reward and the visited set as NW 32-bit words.  A decision first builds,
// per lane, the bitmask of customers feasible from the current POI (fnow) and
// of those reachable before their close time (cnow), then walks the
// candidates in index order; a candidate no lane can take is skipped by the
// whole wave.
//
// Exactness.  Every floating-point expression follows the reference's
// operand order (engine/sim_core.hpp, engine/sim_terms.hpp); the file is
// compiled with -ffp-contract=off.  erf/erfc are glibc-compatible
// (sigp_libm.h); pow(var, 0.5) of ATRISK is tabulated per instance with the
// host libm call the reference makes; rewards are sums of scores, reduced in
// scenario order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <unordered_map>
#include <cstdlib>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sigp.h"
#include "sigp_libm.h"
#include "sigp_terms.h"
#include "sigp_fused.h"
#include "sigp_internal.h"

using namespace sigp;

// ---------------------------------------------------------------------------
// error state
static thread_local std::string g_err;
static sigp_status fail(sigp_status st, const std::string &msg) {
    g_err = msg;
    return st;
}
#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? SIGP_OOM : SIGP_DEVICE_ERROR,       \
                        std::string(#expr) + ": " + hipGetErrorString(e_));             \
    } while (0)

extern "C" const char *sigp_last_error(void) { return g_err.c_str(); }

// Error reporting for the library's other translation units.
sigp_status sigp_internal_fail(sigp_status st, const std::string &msg) { return fail(st, msg); }

constexpr uint64_t B(int s) { return 1ull << s; }

// ---------------------------------------------------------------------------
// instance setup kernels

// distT = dist^T, rk[a][q] = score[q] / max(dist[a][q+1], 1e-10) and its
// transpose (engine/sim_terms.hpp:31,89,109,427-428,558: the same divisions).
__global__ void prep_tables_kernel(int n, const double *__restrict__ dist,
                                   const double *__restrict__ score, double *__restrict__ distT,
                                   double *__restrict__ rk, double *__restrict__ rkT) {
    const int ld = n + 1;
    const long long total = (long long)ld * ld;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int a = (int)(idx / ld), b = (int)(idx % ld);
        distT[(size_t)b * ld + a] = dist[idx];
        if (b >= 1) {
            const int q = b - 1;
            const double d = dist[idx];
            const double v = score[q] / (d > 1e-10 ? d : 1e-10);
            rk[(size_t)a * n + q] = v;
            rkT[(size_t)q * ld + a] = v;
        }
    }
}

// Order-preserving map from doubles to unsigned keys and back.
__device__ __forceinline__ uint64_t okey(double x) {
    const uint64_t u = sigp_bits(x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double unkey(uint64_t k) {
    return sigp_dbl((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k);
}

// Largest finite double f with pred(f) true for a predicate that is true
// below some point and false above it; -inf when it is never true.
template <class Pred>
__device__ double last_true(Pred pred) {
    const double lo_v = -1.7976931348623157e308, hi_v = 1.7976931348623157e308;
    if (!pred(lo_v)) return -__builtin_inf();
    if (pred(hi_v)) return hi_v;
    uint64_t lo = okey(lo_v), hi = okey(hi_v);
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (pred(unkey(mid))) lo = mid;
        else hi = mid;
    }
    return unkey(lo);
}

// Feasibility thresholds and ratios per (POI a, customer j): rows [a][j]
// (row length ldp) and columns [j][a] (ldp rows of length ldq; rows j >= n
// are padding).  The predicates are the reference's neighbour tests verbatim
// (engine/sim_terms.hpp:27-30, 67-70, 522-531): f + d > close fails;
// max(f + d, open) + dur + d2d > tmax fails (dur * (1 + sigma) for the
// robust variant); the NS-family variants also fail when d < 1e-10.
__global__ void prep_thresholds_kernel(int n, int ldp, int ldq, const double *__restrict__ dist,
                                       const double *__restrict__ open, const double *__restrict__ close,
                                       const double *__restrict__ dur, const double *__restrict__ durrob,
                                       const double *__restrict__ d2d, const double *__restrict__ score,
                                       double tmax, double *t0, double *t0n, double *t2, double *t2n,
                                       double *r, double *d, double *t0c, double *t0nc, double *t1c,
                                       double *t2c, double *t2nc, double *rc) {
    const int ld = n + 1;
    const long long total = (long long)ld * ldp;
    const double ninf = -__builtin_inf();
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int a = (int)(idx / ldp), j = (int)(idx % ldp);
        double v0 = ninf, v1 = ninf, v2 = ninf, vr = 0.0, vd = __builtin_inf();
        bool near = true;
        if (j < n) {
            const double dd = dist[(size_t)a * ld + j + 1];
            const double cl = close[j], op = open[j], du = dur[j], dr = durrob[j], dd2 = d2d[j];
            v1 = last_true([&](double f) { return !(f + dd > cl); });
            v0 = last_true([&](double f) {
                const double arr = f + dd;
                if (arr > cl) return false;
                const double st = arr > op ? arr : op;
                return !(st + du + dd2 > tmax);
            });
            v2 = last_true([&](double f) {
                const double arr = f + dd;
                if (arr > cl) return false;
                const double st = arr > op ? arr : op;
                return !(st + dr + dd2 > tmax);
            });
            vr = score[j] / (dd > 1e-10 ? dd : 1e-10);
            vd = dd;
            near = dd < 1e-10;
        }
        const double v0n = near ? ninf : v0, v2n = near ? ninf : v2;
        t0[idx] = v0; t0n[idx] = v0n; t2[idx] = v2; t2n[idx] = v2n; r[idx] = vr; d[idx] = vd;
        const size_t c = (size_t)j * ldq + a;
        t0c[c] = v0; t0nc[c] = v0n; t1c[c] = v1; t2c[c] = v2; t2nc[c] = v2n; rc[c] = vr;
    }
}

// z-statistic features of every customer at the initial state
// (engine/sim_core.hpp:277-304): one lane per customer.
template <int NW>
__global__ void zfeat_kernel(DevInst I, double *__restrict__ feat, int *__restrict__ ok) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t vis[NW], fnow[NW], cnow[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) { vis[w] = 0; fnow[w] = 0; cnow[w] = 0; }
    for (int j = 0; j < I.n; ++j) {   // feasibility from the depot at time 0
        const double a = 0.0 + I.dist[j + 1];
        if (a > I.close[j]) continue;
        wset<NW>(cnow, j);
        const double st = a > I.open[j] ? a : I.open[j];
        if (st + I.dur[j] + I.d2d[j] > I.tmax) continue;
        wset<NW>(fnow, j);
    }
    if (i >= I.n) return;
    const double ta = I.dist[i + 1];
    const double arr = ta;
    int good = 1;
    if (arr > I.close[i]) good = 0;
    const double st = arr > I.open[i] ? arr : I.open[i];
    if (good && st + I.dur[i] + I.d2d[i] > I.tmax) good = 0;
    ok[i] = good;
    if (!good) return;
    const double to = I.open[i];
    const double tsv = to > ta ? to : ta;
    const double fin = tsv + I.dur[i];
    const double ac = 0.0 + ta;
    const double sc = ac > I.open[i] ? ac : I.open[i];
    feat[(size_t)i * 5 + 0] = I.close[i];
    feat[(size_t)i * 5 + 1] = t_reach2x<NW>(I, i, fin, 0.0, vis, false);
    feat[(size_t)i * 5 + 2] = t_regret<NW>(I, i, fin, vis, fnow);
    feat[(size_t)i * 5 + 3] = t_eloss<NW>(I, i, sc + I.dur[i], vis, cnow);
    feat[(size_t)i * 5 + 4] = I.d2d[i];
}

// Sequential accumulation in customer order (engine/sim_core.hpp:298-311).
__global__ void zreduce_kernel(int n, const double *__restrict__ feat, const int *__restrict__ ok,
                               double *__restrict__ zp) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double zs[5] = {0, 0, 0, 0, 0}, zq[5] = {0, 0, 0, 0, 0};
    double cnt = 0.0;
    for (int i = 0; i < n; ++i) {
        if (!ok[i]) continue;
        for (int k = 0; k < 5; ++k) {
            const double f = feat[(size_t)i * 5 + k];
            zs[k] += f;
            zq[k] += f * f;
        }
        cnt += 1.0;
    }
    if (cnt < 1.0) cnt = 1.0;
    for (int k = 0; k < 5; ++k) {
        const double mean = zs[k] / cnt;
        double var = zq[k] / cnt - mean * mean;
        if (var < 0.0) var = 0.0;
        zp[k] = mean;
        zp[5 + k] = sqrt(var) + 1e-9;
    }
}

// ---------------------------------------------------------------------------
// simulation kernel

struct RuleSet {
    cip ops, args, off, len;
    cup rd;      // per rule: slots the rule reads (its leaves index them compactly)
    cup flags;   // per rule: fused computations (F_*) and REACH path lengths
    cip order;   // launch order of (instance, rule) pairs k * P + p, longest first; null = k = 0, p = index
    int P;
    int KL;      // value-store rows per wavefront: max over rules of popcount(rd), >= 1
    int SL;      // stack rows per wavefront: max postfix depth - 1, >= 1
};

struct ScenSpec {
    const double *sampled;   // rows of n durations (instance k's rows start at its samp_off)
    int T;                   // scenarios per rule
    int group_size;          // rule p reads rows (p / group_size) * T + s of its instance
    int nblk;                // ceil(T / 64)
    // Scenario order of scenario lanes (null: index order): lane j of pair
    // (k, p)'s scenario blocks simulates scenario perm[perm_off + (g * n +
    // c) * T + j], g = p / group_size, c = first[k * P + p] (the rule's first
    // customer; -1: index order).  The result still goes to slot (p, s).
    const int32_t *perm;
    int32_t *first;
    int first_only;          // candidate lanes: only record each pair's first decision in `first`
    long long n_ranks;       // scenario lanes: pairs of the launch (the grid is padded, see sim_kernel)
    const int32_t *rep;      // null, or rep[p] != p: rule p repeats rule rep[p] and is skipped
    // Protocol: 0 = vehicle by vehicle (engine/sim_core.hpp:207-238); > 0 =
    // round robin (engine/sim_core.hpp:241-274) with route states of up to
    // `rr` vehicles, 2 * rr LDS rows per wavefront after the W wavefronts'
    // value stores and stacks (see rr_rotate).
    int rr;
};

// Workgroups are dispatched to the 8 XCDs round-robin by index.  Scenario
// lanes give XCD x the pairs x, x + 8, ... of the launch order, each with all
// its scenario blocks in a row, so a rule's blocks share one L2 (its rewards'
// cache lines, written in scenario order or permuted, merge there) and the
// longest-first order holds on every XCD.
constexpr int N_XCD = 8;

// One instance of a launch: its tables and where its scenario rows and
// results start.  A single-instance batch has one slot; a multi-instance
// batch one per instance, all evaluating the same rules.
struct InstSlot {
    DevInst dev;
    FusedTabs tabs;
    long long samp_off;   // first double of the instance's scenario rows
    long long out_off;    // first result: instance k's pairs are out[out_off + p * T + s]
    long long perm_off;   // first entry of the instance's scenario orders
};
// Kernels take the slots as a read-only, non-aliased argument, so that their
// wave-uniform reads of it are scalar loads.
typedef const InstSlot *__restrict__ cslot;

// Postfix interpreter (engine/sim_core.hpp:11-41).  Leaves read the
// candidate's terminal values from the wave's LDS value store (`tv`, one
// 64-lane row per compact slot index); the stack lives in LDS below them
// (`st`, one row per level) with its top element in a register.  The stack
// depth is wave-uniform, so every LDS access is one conflict-free row.
// RMOD = false compiles out the rmod operator (opcode 6) for batches with no
// such node, so that the other batches carry no pow code or registers.
template <bool RMOD>
__device__ __forceinline__ double eval_rule(const cip ops, const cip args, int L,
                                            const double *tv, double *st) {
    double top = 0.0;
    int d = 0;   // values on the stack; the topmost is `top`
    for (int k = 0; k < L; ++k) {
        const int o = ops[k];
        if (o < 0) {
            double v;
            if (o == -1) v = tv[args[2 * k] * 64];
            else v = args[2 * k] / 1000.0;
            if (d > 0) st[(d - 1) * 64] = top;
            top = v;
            ++d;
        } else {
            const double b = top;
            const double a = st[(d - 2) * 64];
            --d;
            double r;
            if (o == 0) r = a + b;
            else if (o == 1) r = a - b;
            else if (o == 2) r = a * b;
            else if (o == 3) r = fabs(b) < 1e-10 ? 1.0 : a / b;
            else if (o == 4) r = a < b ? a : b;
            else if (o == 5 || !RMOD) r = a > b ? a : b;
            else {
                const double base = a > 0.0 ? a : 0.0;
                double e = b;
                if (e < 0.0) e = 0.0;
                else if (e > 1.0) e = 1.0;
                r = sigp_pow_glibc(base, 1.0 + e);
            }
            top = r;
        }
    }
    return top;
}

// One decision of pick_best (engine/sim_core.hpp:43-204) for every running lane.
// Candidates i0, i0 + istep, ... of one decision; returns this subset's
// first maximum (index, priority) -- the reference's strict ">" scan in index
// order -- so that subsets evaluated by cooperating wavefronts merge exactly.
// `rd` holds the slots the rule reads; slot s is stored at compact index
// popcount(rd & (bit(s) - 1)) of the value store, as its leaves were
// renumbered when the batch was built.
// RARE = false compiles out the slots no scan family produces (44, 49, 50,
// 53-57, 59, 60) for batches whose rules read none of them.
// Index of the r-th (0-based) set bit of a, r < popcount(a).
__device__ __forceinline__ int select_bit(uint32_t a, int r) {
    int pos = 0;
    int c = __builtin_popcount(a & 0xffffu);
    if (r >= c) { r -= c; a >>= 16; pos += 16; }
    c = __builtin_popcount(a & 0xffu);
    if (r >= c) { r -= c; a >>= 8; pos += 8; }
    c = __builtin_popcount(a & 0xfu);
    if (r >= c) { r -= c; a >>= 4; pos += 4; }
    c = __builtin_popcount(a & 0x3u);
    if (r >= c) { r -= c; a >>= 2; pos += 2; }
    if (r >= (int)(a & 1u)) pos += 1;
    return pos;
}

// COMPACT: the route state (cur, ct, vis) is the same in every lane of the
// workgroup (candidate lanes), so the feasible candidates are too, and the
// k-th of them (in index order) goes to the lanes' k-th slot: k = i0,
// i0 + istep, ... -- no lane is handed an infeasible candidate.  SEGW < 64:
// the state is the same within each aligned segment of SEGW lanes (one
// simulation per segment), and the masks are built per segment.
template <int NW, bool RMOD, bool RARE, bool COMPACT = false, int SEGW = 64>
__device__ __forceinline__ int decide(const DevInst &I, const FusedTabs &F, const cip ops, const cip args, int L,
                      uint64_t rd, uint64_t fl, int cur, double ct, int rem,
                      const uint32_t (&vis)[NW], int i0, int istep, double *tv, double *st,
                      double &best_pri_out) {
    // Per-decision feasibility from the lane's current POI: fnow (the
    // candidate test, :78-82), cnow (close time only, ATRISK :239-240) and
    // the robust bits of NFEASROB (:68); thresholds read down column `cur`.
    uint32_t fnow[NW], cnow[NW], frob[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) { fnow[w] = 0; cnow[w] = 0; frob[w] = 0; }
    const bool want_c = fl & F_CNOW, want_r = fl & F_FNOWROB;
    if (COMPACT && SEGW < 64) {
        // segments of SEGW lanes, each with its own route state: one customer
        // per lane of the segment, the segment's SEGW bits of each ballot
        const int lane = (int)(threadIdx.x & 63);
        const int sl = lane % SEGW, base = lane - sl;
        const uint64_t keep = (1ull << SEGW) - 1;
        for (int b = 0; b < I.n; b += SEGW) {
            const int j = b + sl;
            bool f = false, c = false, r = false;
            if (j < I.n) {
                const size_t at = (size_t)j * F.ldq + cur;
                f = ct <= F.t0c[at];
                if (want_c) c = ct <= F.t1c[at];
                if (want_r) r = ct <= F.t2c[at];
            }
            const uint32_t sf = (uint32_t)((__builtin_amdgcn_ballot_w64(f) >> base) & keep),
                           sc = (uint32_t)((__builtin_amdgcn_ballot_w64(c) >> base) & keep),
                           sr = (uint32_t)((__builtin_amdgcn_ballot_w64(r) >> base) & keep);
            const int w0 = b >> 5, sh = b & 31;
#pragma unroll
            for (int k = 0; k < NW; ++k)
                if (k == w0) {
                    fnow[k] |= sf << sh;
                    cnow[k] |= sc << sh;
                    frob[k] |= sr << sh;
                }
        }
    } else if (COMPACT) {
        // one customer per lane, 64 bits per ballot (the state is the same in every lane)
        const int lane = (int)(threadIdx.x & 63);
        for (int b = 0; b < I.n; b += 64) {
            const int j = b + lane;
            bool f = false, c = false, r = false;
            if (j < I.n) {
                const size_t at = (size_t)j * F.ldq + cur;
                f = ct <= F.t0c[at];
                if (want_c) c = ct <= F.t1c[at];
                if (want_r) r = ct <= F.t2c[at];
            }
            const uint64_t mf = __builtin_amdgcn_ballot_w64(f), mc = __builtin_amdgcn_ballot_w64(c),
                           mr = __builtin_amdgcn_ballot_w64(r);
            const int w0 = b >> 5;
#pragma unroll
            for (int k = 0; k < NW; ++k) {
                if (k == w0) {
                    fnow[k] = (uint32_t)mf;
                    cnow[k] = (uint32_t)mc;
                    frob[k] = (uint32_t)mr;
                } else if (k == w0 + 1) {
                    fnow[k] = (uint32_t)(mf >> 32);
                    cnow[k] = (uint32_t)(mc >> 32);
                    frob[k] = (uint32_t)(mr >> 32);
                }
            }
        }
    } else {
        uint32_t fw = 0, cw = 0, rw = 0;
        // four customers' loads in flight at n <= 128 (measured on MI355X:
        // plain workload -0.8 %; the n <= 288 kernel no faster)
        constexpr int UNROLL = NW == 4 ? 4 : 1;
#pragma unroll UNROLL
        for (int j = 0; j < I.n; ++j) {
            const int bit = j & 31;
            const size_t at = (size_t)j * F.ldq + cur;
            fw |= (ct <= F.t0c[at] ? 1u : 0u) << bit;
            if (want_c) cw |= (ct <= F.t1c[at] ? 1u : 0u) << bit;
            if (want_r) rw |= (ct <= F.t2c[at] ? 1u : 0u) << bit;
            if (bit == 31 || j == I.n - 1) {
                const int w = j >> 5;
#pragma unroll
                for (int k = 0; k < NW; ++k) {
                    fnow[k] = (w == k) ? fw : fnow[k];
                    cnow[k] = (w == k) ? cw : cnow[k];
                    frob[k] = (w == k) ? rw : frob[k];
                }
                fw = cw = rw = 0;
            }
        }
    }
    // VFRAC / NFEAS / FEASCORE / NFEASROB pre-pass (:56-71): counts are exact
    // integers; FEASCORE sums scores in customer order.
    double nvis = 0.0, nfeas = 0.0, fsc = 0.0, nfr = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        nvis += (double)__builtin_popcount(vis[w]);
        nfeas += (double)__builtin_popcount(fnow[w] & ~vis[w]);
        nfr += (double)__builtin_popcount(fnow[w] & frob[w] & ~vis[w]);
    }
    if (fl & F_FSCORE) {
        uint32_t m = 0;
        for (int j = 0; j < I.n; ++j) {
            if ((j & 31) == 0) m = wsel<NW>(fnow, j >> 5) & ~wsel<NW>(vis, j >> 5);
            if ((m >> (j & 31)) & 1u) fsc += I.score[j];
        }
    }
    const int kpath = (int)((fl >> 40) & 7), kpathr = (int)((fl >> 43) & 7);
    // Slots that do not depend on the candidate: stored once per decision.
    put_slot(tv, rd, 9, rem * I.tmax + (I.tmax - ct));
    put_slot(tv, rd, 24, I.tmax > 0.0 ? ct / I.tmax : 0.0);
    put_slot(tv, rd, 25, nvis / I.n);
    put_slot(tv, rd, 26, nfeas);
    put_slot(tv, rd, 27, (double)rem);
    put_slot(tv, rd, 36, 0.1);
    put_slot(tv, rd, 37, 0.2);
    put_slot(tv, rd, 38, 0.3);
    put_slot(tv, rd, 39, 0.4);
    put_slot(tv, rd, 40, 0.5);
    put_slot(tv, rd, 51, fsc);
    put_slot(tv, rd, 52, nfr);

    int best = -1;
    double best_pri = -1e300;
    const int n_feas = (int)nfeas;
    int wc = -1;
    uint32_t vwc = 0, fwc = 0;
    for (int k = i0; k < (COMPACT ? n_feas : I.n); k += istep) {
        int i = k;
        if (COMPACT) {
            int r = k;
            i = -1;
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                const uint32_t a = fnow[q] & ~vis[q];
                const int c = __builtin_popcount(a);
                if (i < 0) {
                    if (r < c) i = q * 32 + select_bit(a, r);
                    else r -= c;
                }
            }
        } else {
            const uint32_t b = 1u << (i & 31);
            const int w = i >> 5;
            bool feas;
            if constexpr (NW == 4) {
                // the words of candidate i's 32, selected once per 32 (measured
                // on MI355X: plain workload -0.8 %, the n <= 288 kernel 7 %
                // slower, so only at n <= 128)
                if (w != wc) {
                    wc = w;
                    vwc = wsel<NW>(vis, w);
                    fwc = wsel<NW>(fnow, w);
                }
                feas = !(vwc & b) && (fwc & b);
            } else {
                feas = !(wsel<NW>(vis, w) & b) && (wsel<NW>(fnow, w) & b);
            }
            if (__builtin_amdgcn_ballot_w64(feas) == 0) continue;
            if (!feas) continue;
        }
        const double ta = I.distT[(size_t)(i + 1) * I.ld + cur];
        const double oi = I.open[i], di = I.dur[i];
        const double to = oi - ct;
        const double tc = I.close[i] - ct;
        const double tsv = to > ta ? to : ta;
        const double tfv = tsv + di;
#define PUT(sl, val) put_slot(tv, rd, sl, val)
        PUT(0, I.score[i]);
        PUT(1, di);
        PUT(2, to);
        PUT(3, tc);
        PUT(4, ta);
        PUT(5, I.d2d[i]);
        PUT(6, tsv);
        PUT(7, tfv);
        PUT(8, tc - ta);
        PUT(16, t_rand(i));
        PUT(31, (tc - I.zp[0]) / I.zp[5]);
        PUT(35, (I.d2d[i] - I.zp[4]) / I.zp[9]);
        if (rd & B(42)) {
            double den = I.tmax - ct;
            if (den < 1.0) den = 1.0;
            PUT(42, (ta + di) / den);
        }
        PUT(58, I.distT[cur] + I.d2d[i] - ta);   // dist[cur][0]
        const double fin = ct + tfv;
        const double arr = ct + ta;
        const double st_c = arr > oi ? arr : oi;
        const double fin_c = st_c + di;

        if (fl & 0x7FFull) cand_terms<NW>(I, F, fl, rd, kpath, kpathr, i, ta, fin, st_c, fin_c, vis, fnow, cnow, tv);
        if (RARE && (rd & B(44))) PUT(44, t_skiploss<NW>(I, i, cur, ct, vis));
        if (RARE && (rd & B(49))) PUT(49, t_ns_alpha<NW>(I, i, fin, vis, false));
        if (RARE && (rd & B(50))) PUT(50, t_ns_alpha<NW>(I, i, fin, vis, true));
        if (RARE && (rd & B(53))) PUT(53, t_ns_rob_m<NW>(I, i, fin, di, vis, 0.5, true));
        if (RARE && (rd & B(54))) PUT(54, t_ns_rob_m<NW>(I, i, fin, di, vis, 2.0, true));
        if (RARE && (rd & B(55))) PUT(55, t_nnfwd<NW>(I, i, fin, vis));
        if (RARE && (rd & B(56))) PUT(56, t_nnk3<NW>(I, i, fin, vis));
        if (RARE && (rd & B(57))) PUT(57, t_nnchain<NW>(I, i, st_c, vis, 3));
        if (RARE && (rd & B(59))) PUT(59, t_detour<NW>(I, i, cur, fin, ta, vis));
        if (RARE && (rd & B(60))) PUT(60, t_scorerank<NW>(I, i, fin, vis));

#undef PUT
        double pri = eval_rule<RMOD>(ops, args, L, tv, st);
        if (pri != pri) pri = -1e300;
        if (pri > best_pri) { best_pri = pri; best = i; }
    }
    best_pri_out = best_pri;
    return best;
}

// Round-robin protocol (engine/sim_core.hpp:241-274): the vehicles advance
// one POI per round.  Each lane keeps the route state of every vehicle in its
// own column of 2 * m LDS rows (clock, then current POI), all zero at the
// start.  After vehicle `veh` has decided -- a move, or the vehicle closing
// (its bit cleared from `open`) -- its state is saved and the next vehicle
// still open after it, cyclically, resumes; that is the order of the
// reference's `for v` inside `while (n_active > 0)`.  Returns false once
// every vehicle has closed.
__device__ __forceinline__ double *rr_rows(double *lds_rows, const RuleSet &R, int W, int wv, int rr, int lane) {
    return lds_rows + (size_t)W * (R.KL + R.SL) * 64 + (size_t)wv * 2 * rr * 64 + lane;
}
__device__ __forceinline__ void rr_clear(double *rr, int m) {
    for (int v = 0; v < 2 * m; ++v) rr[(size_t)v * 64] = 0.0;
}
__device__ __forceinline__ bool rr_rotate(double *rr, uint32_t open, int &veh, int &cur, double &ct) {
    rr[(size_t)(2 * veh) * 64] = ct;
    rr[(size_t)(2 * veh + 1) * 64] = (double)cur;
    if (open == 0) return false;
    const uint32_t later = open & (~0u << (veh + 1));   // veh < SIGP_RR_MAX_M
    veh = later ? __builtin_ctz(later) : __builtin_ctz(open);
    ct = rr[(size_t)(2 * veh) * 64];
    cur = (int)rr[(size_t)(2 * veh + 1) * 64];
    return true;
}

// One workgroup = one rule x a block of 64 scenarios (engine/sim_core.hpp:
// 207-238 per lane), simulated by W = blockDim.x / 64 cooperating
// wavefronts: wavefront w evaluates candidates w, w + W, ... of every
// decision, and the W partial first-maxima are merged through LDS (largest
// priority, then lowest index), which is exactly the reference's argmax.  All
// W wavefronts hold identical route state, so they leave the loop together.
// Dynamic LDS: per wavefront KL + SL rows of 64 doubles (value store, then
// stack), so a launch takes W * (KL + SL) * 512 bytes.
template <int NW, bool RMOD, bool RARE>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 256),
                          amdgpu_waves_per_eu(1)))
void sim_kernel(cslot slots, RuleSet R, ScenSpec S, double *__restrict__ out) {
    __shared__ double lpri[4][64];
    __shared__ int lidx[4][64];
    extern __shared__ double lds_rows[];
    const int xcd = (int)(blockIdx.x % N_XCD);
    const int local = (int)(blockIdx.x / N_XCD);
    const int rank = xcd + N_XCD * (local / S.nblk);
    const int blk = local % S.nblk;
    if (rank >= S.n_ranks) return;   // padding of the last round (whole workgroup)
    const int pair = R.order ? R.order[rank] : rank;
    const int k = pair / R.P;
    const int p = pair - k * R.P;
    if (S.rep && S.rep[p] != p) return;   // a repeated rule (whole workgroup)
    const DevInst I = dev_of(slots[k].dev);
    const FusedTabs F = tabs_of<NW == 4>(slots[k].tabs, I);
    const int W = (int)(blockDim.x >> 6);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    int s = blk * 64 + lane;
    const bool active = s < S.T;
    const int g = p / S.group_size;
    if (S.perm && active) {
        // lanes of a block take scenarios adjacent in the rule's order, so
        // their routes part later (the first decision is the same in every
        // scenario; the first service time sets the clock after it)
        const int c1 = S.first[pair];
        if (c1 >= 0) s = S.perm[slots[k].perm_off + ((size_t)g * I.n + c1) * S.T + s];
    }
    const int off = R.off[p], L = R.len[p];
    const cip ops = R.ops + off;
    const cip args = R.args + 2 * (size_t)off;
    const uint64_t rd = R.rd[p];
    const uint64_t fl = R.flags[p];
    const size_t row = (size_t)g * S.T + (active ? s : 0);
    const double *__restrict__ samp = S.sampled + slots[k].samp_off + row * (size_t)I.n;
    double *tv = lds_rows + (size_t)wv * (R.KL + R.SL) * 64 + lane;
    double *st = tv + (size_t)R.KL * 64;

    uint32_t vis[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) vis[w] = 0;
    int cur = 0, veh = 0;
    double ct = 0.0, total = 0.0;
    bool running = active && I.m > 0;
    uint32_t open = 0;
    if (S.rr) {
        rr_clear(rr_rows(lds_rows, R, W, wv, S.rr, lane), I.m);
        open = (1u << I.m) - 1u;
    }
    // A lane that has closed vehicle `veh` holds until every running lane of
    // the block has closed it; then all start the next vehicle together.
    // Lanes at the same stage of the same vehicle share most of their
    // feasible candidates, so the block walks fewer candidate bodies than
    // when each lane moves on at once (modelled on the oracle over 16 rules
    // of the n = 288 workload: 36 % fewer bodies for 18 % more decisions).
    // A lane's own decisions are unchanged, so are its results.  (Round
    // robin never holds: its lanes rotate through the vehicles on their own.)
    bool held = false;
    while (__builtin_amdgcn_ballot_w64(running) != 0) {
        int best = -1;
        double bp = -1e300;
        const bool deciding = running && !held;
        if (deciding) {
            const int rem = I.m - veh - 1;
            best = decide<NW, RMOD, RARE>(I, F, ops, args, L, rd, fl, cur, ct, rem, vis, wv, W, tv, st, bp);
        }
        if (W > 1) {
            lpri[wv][lane] = bp;
            lidx[wv][lane] = best;
            __syncthreads();
            double m = -1e300;
            int bi = -1;
            for (int w = 0; w < W; ++w) {
                const int ix = lidx[w][lane];
                const double pr = lpri[w][lane];
                if (ix >= 0 && (pr > m || (pr == m && ix < bi))) { m = pr; bi = ix; }
            }
            best = bi;
            __syncthreads();
        }
        if (deciding) {
            if (best < 0) {
                if (S.rr) open &= ~(1u << veh);
                else held = true;
            } else {
                const int pb = best + 1;
                const double ta = I.dist[(size_t)cur * I.ld + pb];
                const double a = ct + ta;
                const double st = a > I.open[best] ? a : I.open[best];
                ct = st + samp[best];
                cur = pb;
                total += I.score[best];
                wset<NW>(vis, best);
            }
            if (S.rr) running = rr_rotate(rr_rows(lds_rows, R, W, wv, S.rr, lane), open, veh, cur, ct);
        }
        if (__builtin_amdgcn_ballot_w64(running && !held) == 0 && running) {
            ++veh;
            cur = 0;
            ct = 0.0;
            held = false;
            if (veh >= I.m) running = false;
        }
    }
    if (active && wv == 0) out[slots[k].out_off + (size_t)p * S.T + s] = total;
}

// Reduction step of the argmax below: does (pb, ib) beat (pa, ia)?  Index -1
// marks "no candidate"; a candidate's priority is always above -1e300.
__device__ __forceinline__ bool beats(double pb, int ib, double pa, int ia) {
    return ib >= 0 && (ia < 0 || pb > pa || (pb == pa && ib < ia));
}

// Workgroup-wide first maximum: the largest priority, the lowest candidate
// index among equals -- the reference's strict ">" scan in index order
// (engine/sim_core.hpp:199-203).  Every thread gets the winner (or -1).
__device__ __forceinline__ int block_argmax(double pri, int idx, int W, int wv, int lane,
                                            double *rpri, int *ridx) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double op = __shfl_xor(pri, off);
        const int oi = __shfl_xor(idx, off);
        if (beats(op, oi, pri, idx)) { pri = op; idx = oi; }
    }
    if (W > 1) {
        if (lane == 0) { rpri[wv] = pri; ridx[wv] = idx; }
        __syncthreads();
        pri = rpri[0];
        idx = ridx[0];
        for (int w = 1; w < W; ++w)
            if (beats(rpri[w], ridx[w], pri, idx)) { pri = rpri[w]; idx = ridx[w]; }
        __syncthreads();
    }
    return idx;
}

// Candidate-lane mapping: one workgroup = one (rule, scenario) simulation
// (engine/sim_core.hpp:207-238); thread t evaluates the feasible candidates
// t, t + 64W, ... (in index order) of every decision and the workgroup takes
// the first maximum.  Route
// state is identical in every thread.  Suited to few scenarios per rule (the
// reference's GP loop evaluates one per generation, engine/gp.hpp:32), where
// scenario lanes would sit idle.  Dynamic LDS as in sim_kernel.
template <int NW, bool RMOD, bool RARE>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 256)))
void sim_cand_kernel(cslot slots, RuleSet R, ScenSpec S, double *__restrict__ out) {
    __shared__ double rpri[4];
    __shared__ int ridx[4];
    extern __shared__ double lds_rows[];
    const int rank = (int)blockIdx.x / S.T;
    const int s = (int)blockIdx.x - rank * S.T;
    const int pair = R.order ? R.order[rank] : rank;
    const int k = pair / R.P;
    const int p = pair - k * R.P;
    if (S.rep && S.rep[p] != p) return;   // a repeated rule (whole workgroup)
    const DevInst I = dev_of(slots[k].dev);
    const FusedTabs F = tabs_of<false>(slots[k].tabs, I);
    const int W = (int)(blockDim.x >> 6);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int off = R.off[p], L = R.len[p];
    const cip ops = R.ops + off;
    const cip args = R.args + 2 * (size_t)off;
    const uint64_t rd = R.rd[p];
    const uint64_t fl = R.flags[p];
    const size_t row = (size_t)(p / S.group_size) * S.T + s;
    const double *__restrict__ samp = S.sampled + slots[k].samp_off + row * (size_t)I.n;
    double *tv = lds_rows + (size_t)wv * (R.KL + R.SL) * 64 + lane;
    double *st = tv + (size_t)R.KL * 64;

    uint32_t vis[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) vis[w] = 0;
    int cur = 0, veh = 0;
    double ct = 0.0, total = 0.0;
    bool running = I.m > 0;
    if (S.first_only && !running) {
        if (threadIdx.x == 0) S.first[pair] = -1;
        return;
    }
    uint32_t open = 0;
    if (S.rr) {   // route states of the round-robin protocol (same in every thread)
        rr_clear(rr_rows(lds_rows, R, W, wv, S.rr, lane), I.m);
        open = (1u << I.m) - 1u;
    }
    while (running) {
        double bp = -1e300;
        const int rem = I.m - veh - 1;
        int best = decide<NW, RMOD, RARE, true>(I, F, ops, args, L, rd, fl, cur, ct, rem, vis,
                                                (int)threadIdx.x, (int)blockDim.x, tv, st, bp);
        best = block_argmax(bp, best, W, wv, lane, rpri, ridx);
        if (S.first_only) {   // uniform: every thread holds the winner
            if (threadIdx.x == 0) S.first[pair] = best;
            return;
        }
        if (S.rr) {
            if (best < 0) {
                open &= ~(1u << veh);
            } else {
                const int pb = best + 1;
                const double ta = I.dist[(size_t)cur * I.ld + pb];
                const double a = ct + ta;
                const double stt = a > I.open[best] ? a : I.open[best];
                ct = stt + samp[best];
                cur = pb;
                total += I.score[best];
                wset<NW>(vis, best);
            }
            running = rr_rotate(rr_rows(lds_rows, R, W, wv, S.rr, lane), open, veh, cur, ct);
        } else if (best < 0) {
            ++veh;
            cur = 0;
            ct = 0.0;
            if (veh >= I.m) running = false;
        } else {
            const int pb = best + 1;
            const double ta = I.dist[(size_t)cur * I.ld + pb];
            const double a = ct + ta;
            const double stt = a > I.open[best] ? a : I.open[best];
            ct = stt + samp[best];
            cur = pb;
            total += I.score[best];
            wset<NW>(vis, best);
        }
    }
    if (threadIdx.x == 0) out[slots[k].out_off + (size_t)p * S.T + s] = total;
}

// First maximum within each segment of SEGW lanes (the segments are
// aligned, so the butterfly never crosses one): every lane gets its
// segment's winner (or -1).
template <int SEGW>
__device__ __forceinline__ int seg_argmax(double pri, int idx) {
#pragma unroll
    for (int off = SEGW / 2; off >= 1; off >>= 1) {
        const double op = __shfl_xor(pri, off);
        const int oi = __shfl_xor(idx, off);
        if (beats(op, oi, pri, idx)) { pri = op; idx = oi; }
    }
    return idx;
}

// Scenario order for scenario lanes: workgroup b = g * n + c ranks group g's
// T scenarios by their duration of customer c (a total order on the bit
// patterns, ties by index, so the result is always a permutation) and
// writes perm[b * T + rank] = s.  T <= ORDER_MAX_T.
constexpr int ORDER_MAX_T = 4096;
__global__ __launch_bounds__(256) void order_rows_kernel(const double *__restrict__ sampled, int n, int T,
                                                         int32_t *__restrict__ perm) {
    __shared__ unsigned long long key[ORDER_MAX_T];
    const int b = (int)blockIdx.x;
    const int g = b / n, c = b - g * n;
    for (int s = threadIdx.x; s < T; s += blockDim.x) {
        const unsigned long long u = (unsigned long long)__double_as_longlong(sampled[((size_t)g * T + s) * n + c]);
        key[s] = (u >> 63) ? ~u : (u | (1ull << 63));
    }
    __syncthreads();
    for (int s = threadIdx.x; s < T; s += blockDim.x) {
        const unsigned long long ks = key[s];
        int r = 0;
        for (int j = 0; j < T; ++j) {
            const unsigned long long kj = key[j];
            r += (kj < ks || (kj == ks && j < s)) ? 1 : 0;
        }
        perm[(size_t)b * T + r] = s;
    }
}

// Segmented candidate lanes: one wavefront simulates 64 / SEGW scenarios of
// one rule, segment g (lanes g*SEGW ..) the scenario blk * (64/SEGW) + g;
// within a segment, as in sim_cand_kernel, lane t evaluates the feasible
// candidates t, t + SEGW, ... and the segment takes the first maximum.
// Decisions with few feasible candidates -- most of them -- then fill the
// wavefront with several simulations instead of one.  Workgroup = one
// wavefront; dynamic LDS as in sim_kernel.
template <int NW, bool RMOD, bool RARE, int SEGW>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 64), amdgpu_waves_per_eu(2)))
void sim_seg_kernel(cslot slots, RuleSet R, ScenSpec S, double *__restrict__ out) {
    extern __shared__ double lds_rows[];
    constexpr int SEGS = 64 / SEGW;
    const int nsb = (S.T + SEGS - 1) / SEGS;
    const int rank = (int)blockIdx.x / nsb;
    const int blk = (int)blockIdx.x - rank * nsb;
    const int pair = R.order ? R.order[rank] : rank;
    const int k = pair / R.P;
    const int p = pair - k * R.P;
    if (S.rep && S.rep[p] != p) return;   // a repeated rule (whole workgroup)
    const DevInst I = dev_of(slots[k].dev);
    const FusedTabs F = tabs_of<false>(slots[k].tabs, I);
    const int lane = threadIdx.x & 63;
    const int sl = lane % SEGW;
    int s = blk * SEGS + lane / SEGW;
    const bool active = s < S.T;
    const int g = p / S.group_size;
    if (S.perm && active) {   // scenario order, as in sim_kernel
        const int c1 = S.first[pair];
        if (c1 >= 0) s = S.perm[slots[k].perm_off + ((size_t)g * I.n + c1) * S.T + s];
    }
    const int off = R.off[p], L = R.len[p];
    const cip ops = R.ops + off;
    const cip args = R.args + 2 * (size_t)off;
    const uint64_t rd = R.rd[p];
    const uint64_t fl = R.flags[p];
    const size_t row = (size_t)g * S.T + (active ? s : 0);
    const double *__restrict__ samp = S.sampled + slots[k].samp_off + row * (size_t)I.n;
    double *tv = lds_rows + lane;
    double *st = tv + (size_t)R.KL * 64;

    uint32_t vis[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) vis[w] = 0;
    int cur = 0, veh = 0;
    double ct = 0.0, total = 0.0;
    bool running = active && I.m > 0;
    uint32_t open = 0;
    if (S.rr) {   // round robin, as in sim_kernel
        rr_clear(rr_rows(lds_rows, R, 1, 0, S.rr, lane), I.m);
        open = (1u << I.m) - 1u;
    }
    // Vehicle boundaries are kept together as in sim_kernel: a simulation
    // that has closed vehicle `veh` holds until the other has closed it too
    // (modelled on 12 rules x 500 scenarios at n = 100: 10 % fewer candidate
    // passes for 2 % more decisions).  Round robin never holds.
    bool held = false;
    while (__builtin_amdgcn_ballot_w64(running) != 0) {
        double bp = -1e300;
        const int rem = I.m - veh - 1;
        const bool deciding = running && !held;
        // a finished or holding segment walks no candidates (its first slot
        // lies past any count)
        int best = decide<NW, RMOD, RARE, true, SEGW>(I, F, ops, args, L, rd, fl, cur, ct, rem, vis,
                                                      deciding ? sl : (1 << 20), SEGW, tv, st, bp);
        best = seg_argmax<SEGW>(bp, best);
        if (deciding) {
            if (best < 0) {
                if (S.rr) open &= ~(1u << veh);
                else held = true;
            } else {
                const int pb = best + 1;
                const double ta = I.dist[(size_t)cur * I.ld + pb];
                const double a = ct + ta;
                const double stt = a > I.open[best] ? a : I.open[best];
                ct = stt + samp[best];
                cur = pb;
                total += I.score[best];
                wset<NW>(vis, best);
            }
            if (S.rr) running = rr_rotate(rr_rows(lds_rows, R, 1, 0, S.rr, lane), open, veh, cur, ct);
        }
        if (__builtin_amdgcn_ballot_w64(running && !held) == 0 && running) {
            ++veh;
            cur = 0;
            ct = 0.0;
            held = false;
            if (veh >= I.m) running = false;
        }
    }
    if (active && sl == 0) out[slots[k].out_off + (size_t)p * S.T + s] = total;
}

// Decision situations of the phenotype: per situation the state of a route
// under construction and the candidates to rank.
struct PhenoSpec {
    int n_sit, n_cand;
    const int *cur, *rem, *cand;   // n_sit, n_sit, n_sit x n_cand
    const double *ct;              // n_sit
    const uint32_t *vis;           // n_sit x 16 words: visited customers
};

// A rule's phenotype: the rank of each candidate of each situation by the
// rule's priority, 0 = the candidate pick_best would choose.  Priorities
// come from the same per-candidate evaluation as the simulation
// (engine/sim_core.hpp:73-203, via decide on one candidate), clamped below
// at -1e300 (NaN counts as -1e300); equal priorities rank by candidate
// position, as the strict ">" scan would.  One workgroup per rule; thread t
// evaluates the (situation, candidate) pairs t, t + 64W, ...
template <int NW, bool RMOD, bool RARE>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 256)))
void pheno_kernel(cslot slots, RuleSet R, PhenoSpec Ph, int32_t *__restrict__ out) {
    __shared__ double spri[1024];
    extern __shared__ double lds_rows[];
    const int p = (int)blockIdx.x;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const DevInst I = dev_of(slots[0].dev);
    const FusedTabs F = tabs_of<false>(slots[0].tabs, I);
    const int off = R.off[p], L = R.len[p];
    const cip ops = R.ops + off;
    const cip args = R.args + 2 * (size_t)off;
    const uint64_t rd = R.rd[p];
    const uint64_t fl = R.flags[p];
    double *tv = lds_rows + (size_t)wv * (R.KL + R.SL) * 64 + lane;
    double *st = tv + (size_t)R.KL * 64;
    const int pairs = Ph.n_sit * Ph.n_cand;
    for (int q0 = 0; q0 < pairs; q0 += (int)blockDim.x) {
        const int q = q0 + (int)threadIdx.x;
        const bool live = q < pairs;
        const int si = live ? q / Ph.n_cand : 0;
        const int c = live ? Ph.cand[q] : 0;
        uint32_t vis[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) vis[w] = Ph.vis[(size_t)si * 16 + w];
        double bp = -1e300;
        if (live)
            (void)decide<NW, RMOD, RARE>(I, F, ops, args, L, rd, fl, Ph.cur[si], Ph.ct[si], Ph.rem[si],
                                         vis, c, I.n, tv, st, bp);
        if (live) spri[q] = bp;
    }
    __syncthreads();
    for (int q = (int)threadIdx.x; q < pairs; q += (int)blockDim.x) {
        const int base = (q / Ph.n_cand) * Ph.n_cand, j = q - base;
        const double mine = spri[q];
        int rank = 0;
        for (int k = 0; k < Ph.n_cand; ++k) {
            const double o = spri[base + k];
            rank += (o > mine || (o == mine && k < j)) ? 1 : 0;
        }
        out[(size_t)p * pairs + q] = rank;
    }
}

// Per-rule mean in scenario order, then / T (engine/gp_engine.cu:76-81).
__global__ void mean_kernel(const double *__restrict__ pairs, int P, int T, double *__restrict__ mean) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double acc = 0.0;
    const double *row = pairs + (size_t)p * T;
    for (int s = 0; s < T; ++s) acc += row[s];
    mean[p] = acc / T;
}

// Device math self-test (sigp_device_math).
__global__ void math_kernel(int which, const double *__restrict__ x, const double *__restrict__ y,
                            long long count, double *__restrict__ out) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const double v = x[k];
    double r;
    switch (which) {
    case 0: r = sigp_erf(v); break;
    case 1: r = sigp_erfc(v); break;
    case 2: r = sigp_exp(v); break;
    case 3: r = sqrt(v); break;
    default: r = sigp_pow_glibc(v, y[k]); break;
    }
    out[k] = r;
}

// ---------------------------------------------------------------------------
// host side

struct sigp_instance {
    int n = 0, m = 0, K = 0;
    double tmax = 0.0;
    void *dbuf = nullptr;           // one device allocation holding every table
    void *tbuf = nullptr;           // feasibility-threshold tables
    DevInst dev{};
    FusedTabs tabs{};
    std::vector<double> zp;
    double *scen_buf = nullptr;     // cached staging for sigp_evaluate
    size_t scen_cap = 0;
    double *out_buf = nullptr;
    size_t out_cap = 0;
    double *mean_buf = nullptr;
    size_t mean_cap = 0;
    int waves_per_item = 0;         // default launch mapping of new batches (0 = per batch)
    int mapping = SIGP_MAP_AUTO;
    bool skip_repeats = false;      // new batches simulate each distinct rule once
    int protocol = SIGP_PROTOCOL_VEHICLES;   // protocol of new batches
    bool one_launch = false;        // multi-instance batches: every size class in one launch
};

// Launch order and tuned mapping of the instances of one size class (one
// visited-set width NW of the kernels).
struct LaunchUnit {
    int nwb = 4;                    // NW of the class's kernels
    int max_n = 0;                  // largest instance of the class
    long long n_pairs = 0;          // (instance, rule) pairs of the class
    cip order = nullptr;            // their launch order, pair = k * P + p, longest first
    int mapping = SIGP_MAP_AUTO;    // tuned for this class; AUTO = follow the batch
    int waves = 0;                  // tuned for this class; 0 = follow the batch
    cip tune_order = nullptr;       // every s-th pair of `order`: the autotuner's sample
    long long tune_pairs = 0;
    cip check_order = nullptr;      // a 4x larger sample: the autotuner's check of its ranking
    long long check_pairs = 0;
};

struct sigp_batch {
    std::vector<sigp_instance *> insts;   // one, or the instances of a multi-instance batch
    bool multi = false;
    int P = 0, T = 0, group_size = 1;
    int waves_per_item = 0;         // wavefronts per workgroup; 0 = default of the mapping
    int mapping = SIGP_MAP_AUTO;    // SIGP_MAP_*; AUTO resolves from T at launch
    int64_t rows_needed = 0;        // scenario rows sigp_batch_run needs
    long long n_out = 0;            // results: instances x P x T
    bool has_rmod = false;          // some rule has an rmod node
    bool has_rare = false;          // some rule reads a slot outside the scan families
    void *dbuf = nullptr;           // rules, instance slots and launch orders
    RuleSet rs{};                   // rule arrays (order: set per launch unit)
    cslot slots = nullptr;
    int n_units = 0;
    LaunchUnit units[3];            // heaviest size class first
    double *pairs_tmp = nullptr;    // used when the caller wants means only
    hipStream_t side[2] = {nullptr, nullptr};   // size classes 2 and 3 run here, concurrently
    hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
    // scenario order of scenario lanes (ScenSpec::perm): each pair's first
    // customer, computed once, and per run the order of every group's rows
    // by each customer's duration; null when the batch does not qualify
    int32_t *first = nullptr;
    int32_t *perm = nullptr;
    bool order_on = false;
    const int32_t *rep = nullptr;   // repeated rules (sigp_device_rules::rep, or rep_owned)
    int32_t *rep_owned = nullptr;   // host-detected repeats (sigp_instance_skip_repeats); filled by the batch
    int G = 1;                      // scenario groups per instance
    int rr = 0;                     // ScenSpec::rr: 0 = vehicle by vehicle, else round robin over <= rr vehicles
};

static sigp_status check_device(int *dev_out) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(SIGP_NO_DEVICE, "no HIP device visible");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SIGP_NO_DEVICE, std::string("device is ") + prop.gcnArchName +
                                        ", this library is built for gfx950 only");
    if (dev_out) *dev_out = dev;
    return SIGP_OK;
}

extern "C" sigp_status sigp_device_info(char *name_buf, size_t buf_len, int32_t *n_cu,
                                        double *clock_ghz) {
    int dev = 0;
    sigp_status st = check_device(&dev);
    if (st != SIGP_OK) return st;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (name_buf && buf_len) std::snprintf(name_buf, buf_len, "%s (%s)", prop.name, prop.gcnArchName);
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (clock_ghz) *clock_ghz = prop.clockRate / 1e6;
    return SIGP_OK;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static sigp_status validate_desc(const sigp_instance_desc *d) {
    if (!d) return fail(SIGP_INVALID_ARG, "desc is NULL");
    if (d->n < 0 || d->n > SIGP_MAX_N)
        return fail(SIGP_INVALID_ARG, "n must be in [0, " + std::to_string(SIGP_MAX_N) + "]");
    if (d->m < 0) return fail(SIGP_INVALID_ARG, "m must be >= 0");
    if (d->n > 0 && (!d->scores || !d->durations || !d->open_t || !d->close_t ||
                     !d->dist_to_depot))
        return fail(SIGP_INVALID_ARG, "customer arrays must be non-NULL");
    if (!d->dist) return fail(SIGP_INVALID_ARG, "dist must be non-NULL");
    return SIGP_OK;
}

template <int NW>
static void launch_zstats(const DevInst &I, double *feat, int *ok, double *zp) {
    const int threads = 64 * ((I.n + 63) / 64 > 0 ? (I.n + 63) / 64 : 1);
    zfeat_kernel<NW><<<1, threads>>>(I, feat, ok);
    zreduce_kernel<<<1, 64>>>(I.n, feat, ok, zp);
}

static int nw_bucket(int n) { return n <= 128 ? 4 : (n <= 288 ? 9 : 16); }

// Builds the device tables of an instance into one allocation.
// Order-preserving 64-bit keys of finite doubles, for bisection over them.
static uint64_t dkey(double x) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
static double dkey_inv(uint64_t k) {
    const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double x;
    std::memcpy(&x, &u, 8);
    return x;
}
// The erf argument of norm_cdf(d / s) (engine/sim_terms.hpp:221-223), with
// its two roundings; non-decreasing in d for s > 0.
static double ncdf_arg(double d, double s) {
    const double z = d / s;
    return z * 0.7071067811865476;
}
// Smallest d with ncdf_arg(d, s) >= 6: from there erf returns 1 - 1e-300
// == 1 and norm_cdf exactly 1.
static double ncdf_one_from(double s) {
    uint64_t lo = dkey(-DBL_MAX), hi = dkey(DBL_MAX);
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (ncdf_arg(dkey_inv(mid), s) >= 6.0) hi = mid;
        else lo = mid;
    }
    return dkey_inv(hi);
}
// Largest d with ncdf_arg(d, s) <= -6: up to there erf returns 1e-300 - 1
// == -1 and norm_cdf exactly 0.
static double ncdf_zero_to(double s) {
    uint64_t lo = dkey(-DBL_MAX), hi = dkey(DBL_MAX);
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (ncdf_arg(dkey_inv(mid), s) <= -6.0) lo = mid;
        else hi = mid;
    }
    return dkey_inv(lo);
}

static sigp_status build_instance(const sigp_instance_desc *d, sigp_instance *I) {
    const int n = d->n, ld = n + 1;
    I->n = n;
    I->m = d->m;
    I->tmax = d->tmax;
    // duration classes for the ATRISK pow(var, 0.5) table (engine/sim_terms.hpp:229-252)
    std::vector<int32_t> cls(n);
    std::vector<double> cval;
    {
        std::map<uint64_t, int> seen;
        for (int i = 0; i < n; ++i) {
            uint64_t u;
            std::memcpy(&u, &d->durations[i], 8);
            auto it = seen.find(u);
            if (it == seen.end()) {
                const int k = (int)cval.size();
                seen.emplace(u, k);
                cval.push_back(d->durations[i]);
                cls[i] = k;
            } else {
                cls[i] = it->second;
            }
        }
    }
    const int K = std::max(1, (int)cval.size());
    I->K = K;
    std::vector<double> pw2((size_t)K * K, 0.0), pw1(K, 0.0);
    const double SIG = 0.2;
    for (int a = 0; a < (int)cval.size(); ++a) {
        const double sa = SIG * cval[a];
        for (int b = 0; b < (int)cval.size(); ++b) {
            const double sb = SIG * cval[b];
            pw2[(size_t)a * K + b] = std::pow(sa * sa + sb * sb, 0.5);
        }
        pw1[a] = std::pow(0.0 + sa * sa, 0.5);
    }
    std::vector<double> durrob(n), sdv(n);
    for (int i = 0; i < n; ++i) {
        durrob[i] = d->durations[i] * (1.0 + kSigma);
        sdv[i] = SIG * d->durations[i];
    }
    // ATRISK saturation thresholds of every divisor it uses
    std::vector<double> atp1(K, 0.0), atp0(K, 0.0), atq2((size_t)K * K * 2, 0.0), atq1((size_t)K * 2, 0.0);
    for (int a = 0; a < (int)cval.size(); ++a) {
        const double sa = SIG * cval[a];
        if (sa > 0.0) {
            atp1[a] = ncdf_one_from(sa);
            atp0[a] = ncdf_zero_to(sa);
        }
        for (int b = 0; b < (int)cval.size(); ++b) {
            const double r = pw2[(size_t)a * K + b];
            if (r > 0.0) {
                atq2[((size_t)a * K + b) * 2] = ncdf_one_from(r);
                atq2[((size_t)a * K + b) * 2 + 1] = ncdf_zero_to(r);
            }
        }
        if (pw1[a] > 0.0) {
            atq1[2 * a] = ncdf_one_from(pw1[a]);
            atq1[2 * a + 1] = ncdf_zero_to(pw1[a]);
        }
    }
    // ATRISK's p_bud when the mean arrival at j is not after open_j
    // (engine/sim_terms.hpp:243-252: mean_start = open_j, var_start = 0.0),
    // in the reference's expression with the host's erf
    std::vector<double> atpb0(n, 1.0);
    for (int j = 0; j < n; ++j) {
        const double sdj = SIG * d->durations[j];
        const double vb = 0.0 + sdj * sdj;
        const double mbd = d->open_t[j] + d->durations[j] + d->dist_to_depot[j];
        if (vb > 1e-18) atpb0[j] = 0.5 * (1.0 + std::erf(((d->tmax - mbd) / pw1[cls[j]]) * 0.7071067811865476));
        else atpb0[j] = mbd <= d->tmax ? 1.0 : 0.0;
    }

    const size_t nn = (size_t)ld * ld;
    const size_t np = (size_t)std::max(8, (n + 7) / 8 * 8);   // customer arrays padded for 4-wide reads
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return o; };
    const size_t o_dist = take(nn * 8), o_distT = take(nn * 8), o_rk = take((size_t)ld * n * 8 + 8),
                 o_rkT = take((size_t)n * ld * 8 + 8), o_open = take(np * 8),
                 o_close = take(np * 8), o_dur = take(np * 8), o_d2d = take(np * 8),
                 o_score = take(np * 8), o_durrob = take(np * 8), o_sdv = take(np * 8),
                 o_cls = take(n * 4 + 8), o_pw2 = take((size_t)K * K * 8), o_pw1 = take(K * 8),
                 o_atp1 = take(K * 8), o_atp0 = take(K * 8), o_atq2 = take((size_t)K * K * 16),
                 o_atq1 = take((size_t)K * 16), o_atpb0 = take(np * 8),
                 o_zp = take(10 * 8), o_feat = take((size_t)n * 5 * 8 + 8), o_ok = take(n * 4 + 8);
    char *base = nullptr;
    HIP_TRY(hipMalloc((void **)&base, off));
    I->dbuf = base;
    HIP_TRY(hipMemset(base, 0, off));
    auto up = [&](size_t o, const void *src, size_t bytes) -> hipError_t {
        return bytes ? hipMemcpy(base + o, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    HIP_TRY(up(o_dist, d->dist, nn * 8));
    HIP_TRY(up(o_open, d->open_t, n * 8));
    HIP_TRY(up(o_close, d->close_t, n * 8));
    HIP_TRY(up(o_dur, d->durations, n * 8));
    HIP_TRY(up(o_d2d, d->dist_to_depot, n * 8));
    HIP_TRY(up(o_score, d->scores, n * 8));
    HIP_TRY(up(o_durrob, durrob.data(), n * 8));
    HIP_TRY(up(o_sdv, sdv.data(), n * 8));
    HIP_TRY(up(o_cls, cls.data(), n * 4));
    HIP_TRY(up(o_pw2, pw2.data(), (size_t)K * K * 8));
    HIP_TRY(up(o_pw1, pw1.data(), K * 8));
    HIP_TRY(up(o_atp1, atp1.data(), K * 8));
    HIP_TRY(up(o_atp0, atp0.data(), K * 8));
    HIP_TRY(up(o_atq2, atq2.data(), (size_t)K * K * 16));
    HIP_TRY(up(o_atq1, atq1.data(), (size_t)K * 16));
    HIP_TRY(up(o_atpb0, atpb0.data(), (size_t)n * 8));

    DevInst &D = I->dev;
    D.n = n;
    D.m = d->m;
    D.ld = ld;
    D.K = K;
    D.tmax = d->tmax;
    auto cd = [&](size_t o) { return cdp{(cbase)base, (uint32_t)o}; };
    D.dist = cd(o_dist);
    D.distT = cd(o_distT);
    D.rk = cd(o_rk);
    D.rkT = cd(o_rkT);
    D.open = cd(o_open);
    D.close = cd(o_close);
    D.dur = cd(o_dur);
    D.d2d = cd(o_d2d);
    D.score = cd(o_score);
    D.durrob = cd(o_durrob);
    D.sdv = cd(o_sdv);
    D.dcls = cid{(cbase)base, (uint32_t)o_cls};
    D.pw2 = cd(o_pw2);
    D.pw1 = cd(o_pw1);
    D.atp1 = cd(o_atp1);
    D.atp0 = cd(o_atp0);
    D.atq2 = cd(o_atq2);
    D.atq1 = cd(o_atq1);
    D.atpb0 = cd(o_atpb0);
    D.zp = cd(o_zp);

    {
        const long long total = (long long)nn;
        int blocks = (int)std::min<long long>((total + 255) / 256, 4096);
        if (blocks < 1) blocks = 1;
        prep_tables_kernel<<<blocks, 256>>>(n, (const double *)(base + o_dist),
                                            (const double *)(base + o_score), (double *)(base + o_distT),
                                            (double *)(base + o_rk), (double *)(base + o_rkT));
        HIP_TRY(hipGetLastError());
    }
    {
        FusedTabs &F = I->tabs;
        F.ldp = (int)np;
        F.ldq = (n + 1 + 7) / 8 * 8;
        const size_t rows = (size_t)ld * F.ldp, cols = (size_t)F.ldp * F.ldq;
        char *tb = nullptr;
        HIP_TRY(hipMalloc((void **)&tb, 6 * (rows + cols) * 8));
        I->tbuf = tb;
        double *p = (double *)tb;
        double *t[12];
        for (int k = 0; k < 6; ++k) { t[k] = p; p += rows; }
        for (int k = 6; k < 12; ++k) { t[k] = p; p += cols; }
        const long long total = (long long)rows;
        int blocks = (int)std::min<long long>((total + 255) / 256, 8192);
        if (blocks < 1) blocks = 1;
        prep_thresholds_kernel<<<blocks, 256>>>(
            n, F.ldp, F.ldq, (const double *)(base + o_dist), (const double *)(base + o_open),
            (const double *)(base + o_close), (const double *)(base + o_dur),
            (const double *)(base + o_durrob), (const double *)(base + o_d2d),
            (const double *)(base + o_score), d->tmax, t[0], t[1], t[2], t[3], t[4], t[5], t[6],
            t[7], t[8], t[9], t[10], t[11]);
        HIP_TRY(hipGetLastError());
        auto ct = [&](int k) { return cdp{(cbase)tb, (uint32_t)((char *)t[k] - tb)}; };
        F.t0 = ct(0); F.t0n = ct(1); F.t2 = ct(2); F.t2n = ct(3);
        F.r = ct(4); F.d = ct(5);
        F.t0c = ct(6); F.t0nc = ct(7); F.t1c = ct(8); F.t2c = ct(9);
        F.t2nc = ct(10); F.rc = ct(11);
        F.open = D.open; F.close = D.close; F.dur = D.dur; F.d2d = D.d2d; F.score = D.score;
        F.durrob = D.durrob; F.sdv = D.sdv;
    }
    I->zp.assign(10, 0.0);
    if (d->zp) {
        std::memcpy(I->zp.data(), d->zp, 80);
        HIP_TRY(up(o_zp, d->zp, 80));
    } else {
        double *feat = (double *)(base + o_feat);
        int *okp = (int *)(base + o_ok);
        double *zpd = (double *)(base + o_zp);
        switch (nw_bucket(n)) {
        case 4: launch_zstats<4>(D, feat, okp, zpd); break;
        case 9: launch_zstats<9>(D, feat, okp, zpd); break;
        default: launch_zstats<16>(D, feat, okp, zpd); break;
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpy(I->zp.data(), zpd, 80, hipMemcpyDeviceToHost));
    }
    HIP_TRY(hipDeviceSynchronize());
    return SIGP_OK;
}

extern "C" sigp_status sigp_instance_create(const sigp_instance_desc *desc, sigp_instance **out) {
    if (!out) return fail(SIGP_INVALID_ARG, "out is NULL");
    *out = nullptr;
    sigp_status st = validate_desc(desc);
    if (st != SIGP_OK) return st;
    st = check_device(nullptr);
    if (st != SIGP_OK) return st;
    sigp_instance *I = new sigp_instance();
    st = build_instance(desc, I);
    if (st != SIGP_OK) {
        if (I->dbuf) (void)hipFree(I->dbuf);
        if (I->tbuf) (void)hipFree(I->tbuf);
        delete I;
        return st;
    }
    *out = I;
    return SIGP_OK;
}

extern "C" sigp_status sigp_instance_destroy(sigp_instance *inst) {
    if (!inst) return SIGP_OK;
    if (inst->dbuf) (void)hipFree(inst->dbuf);
    if (inst->tbuf) (void)hipFree(inst->tbuf);
    if (inst->scen_buf) (void)hipFree(inst->scen_buf);
    if (inst->out_buf) (void)hipFree(inst->out_buf);
    if (inst->mean_buf) (void)hipFree(inst->mean_buf);
    delete inst;
    return SIGP_OK;
}

extern "C" sigp_status sigp_zstats(const sigp_instance_desc *desc, double *zp_out) {
    if (!zp_out) return fail(SIGP_INVALID_ARG, "zp_out is NULL");
    sigp_instance_desc d = *desc;
    d.zp = nullptr;
    sigp_instance *I = nullptr;
    sigp_status st = sigp_instance_create(&d, &I);
    if (st != SIGP_OK) return st;
    std::memcpy(zp_out, I->zp.data(), 80);
    return sigp_instance_destroy(I);
}

// Validated rules, rewritten for the device: dead leaves become the constant
// 0, live leaves index the compact value store; per rule the read set, the
// fused computations and a relative cost estimate.
struct PreparedRules {
    int P = 0;
    std::vector<int32_t> ops, args, offs, lens;
    std::vector<uint64_t> rd, flags;
    std::vector<float> cost;
    int max_depth = 1, max_k = 1;
    bool any_rmod = false;
    uint64_t any_rd = 0;
};

static sigp_status prepare_rules(const sigp_rules *rules, const int64_t *masks, int32_t n_masks,
                                 int32_t mask_group, PreparedRules &pr) {
    const int P = rules->n_rules;
    if (P < 0) return fail(SIGP_INVALID_ARG, "n_rules < 0");
    if (mask_group < 1) return fail(SIGP_INVALID_ARG, "mask_group must be >= 1");
    if (P > 0 && (!rules->ops || !rules->args || !rules->offsets || !rules->lengths))
        return fail(SIGP_INVALID_ARG, "rule arrays must be non-NULL");
    if (P > 0 && (!masks || n_masks < 1)) return fail(SIGP_INVALID_ARG, "masks missing");
    if ((long long)(P > 0 ? (P - 1) / mask_group : 0) >= (long long)std::max(n_masks, 1))
        return fail(SIGP_INVALID_ARG, "mask row (n_rules-1)/mask_group out of range");
    pr.P = P;
    pr.offs.resize(P);
    pr.lens.resize(P);
    pr.rd.resize(P);
    pr.flags.resize(P);
    pr.cost.resize(P);
    for (int p = 0; p < P; ++p) {
        const int64_t o = rules->offsets[p], L = rules->lengths[p];
        if (o < 0 || L < 1 || o + L > rules->n_nodes)
            return fail(SIGP_INVALID_ARG, "rule " + std::to_string(p) + ": offset/length out of range");
        const int64_t *mrow = masks + (size_t)(p / mask_group) * SIGP_N_TERMS;
        uint64_t mk = 0;
        for (int s = 0; s < SIGP_N_TERMS; ++s)
            if (mrow[s]) mk |= B(s);
        uint64_t used = 0;
        int depth = 0;
        for (int64_t k = 0; k < L; ++k) {
            const int op = rules->ops[o + k];
            if (op == -1) {
                const int s = rules->args[2 * (o + k)];
                if (s < 0 || s >= SIGP_N_TERMS)
                    return fail(SIGP_INVALID_ARG, "rule " + std::to_string(p) + ": bad terminal slot");
                used |= B(s);
                ++depth;
            } else if (op == -2) {
                ++depth;
            } else if (op >= 0 && op <= 6) {
                if (op == 6) pr.any_rmod = true;
                if (depth < 2)
                    return fail(SIGP_INVALID_ARG, "rule " + std::to_string(p) + ": stack underflow");
                --depth;
            } else {
                return fail(SIGP_INVALID_ARG, "rule " + std::to_string(p) + ": bad opcode");
            }
            pr.max_depth = std::max(pr.max_depth, depth);
        }
        if (depth != 1)
            return fail(SIGP_INVALID_ARG, "rule " + std::to_string(p) + ": postfix does not reduce to one value");
        if (pr.max_depth > SIGP_MAX_STACK)
            return fail(SIGP_INVALID_ARG, "rule " + std::to_string(p) + ": stack deeper than " +
                                              std::to_string(SIGP_MAX_STACK));
        uint64_t r = 0, c = 0, f = 0;
        compute_sets(used, mk, &r, &c, &f);
        pr.max_k = std::max(pr.max_k, __builtin_popcountll(r));
        pr.offs[p] = (int32_t)pr.ops.size();
        pr.lens[p] = (int32_t)L;
        for (int64_t k = 0; k < L; ++k) {
            int op = rules->ops[o + k];
            int a0 = rules->args[2 * (o + k)], a1 = rules->args[2 * (o + k) + 1];
            if (op == -1) {
                if (!(r & B(a0))) { op = -2; a0 = 0; a1 = 0; }   // dead slot reads 0.0
                else a0 = __builtin_popcountll(r & (B(a0) - 1));    // compact value-store row
            }
            pr.ops.push_back(op);
            pr.args.push_back(a0);
            pr.args.push_back(a1);
        }
        pr.rd[p] = r;
        pr.any_rd |= r;
        pr.flags[p] = f;
        pr.cost[p] = rule_cost((int)L, f, c);
    }
    return SIGP_OK;
}

// A batch over `insts` (one instance, or several evaluating the same rules):
// the rule arrays, one InstSlot per instance and, per size class, the
// longest-first order of its (instance, rule) pairs, cost = rule cost x n^2.
static sigp_status setup_order(sigp_batch *bt, long long perms);
static constexpr long long ORDER_MAX_BYTES = 256LL << 20;

// Repeated rules: rule p repeats rule q < p when both read the same scenario
// group and their prepared forms (read set, scan flags, compiled code with
// the mask applied) are equal -- then they simulate to the same rewards, so p
// is skipped and its results copied from q after the launch.  Equality is
// checked in full; the hash only finds candidates.
static sigp_status find_repeats(sigp_batch *bt, const PreparedRules &pr, int group_size) {
    const int P = pr.P;
    std::vector<int32_t> rep(P);
    std::unordered_map<uint64_t, std::vector<int32_t>> seen;
    bool any = false;
    auto same = [&](int a, int b) {
        if (a / group_size != b / group_size || pr.rd[a] != pr.rd[b] || pr.flags[a] != pr.flags[b] ||
            pr.lens[a] != pr.lens[b])
            return false;
        const int L = pr.lens[a], oa = pr.offs[a], ob = pr.offs[b];
        for (int k = 0; k < L; ++k)
            if (pr.ops[oa + k] != pr.ops[ob + k] || pr.args[2 * (oa + k)] != pr.args[2 * (ob + k)] ||
                pr.args[2 * (oa + k) + 1] != pr.args[2 * (ob + k) + 1])
                return false;
        return true;
    };
    for (int p = 0; p < P; ++p) {
        uint64_t h = 1469598103934665603ull;
        auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
        mix((uint64_t)(p / group_size));
        mix(pr.rd[p]);
        mix(pr.flags[p]);
        for (int k = 0; k < pr.lens[p]; ++k) {
            const int o = pr.offs[p] + k;
            mix((uint64_t)(uint32_t)pr.ops[o]);
            mix((uint64_t)(uint32_t)pr.args[2 * o]);
            mix((uint64_t)(uint32_t)pr.args[2 * o + 1]);
        }
        rep[p] = p;
        std::vector<int32_t> &cands = seen[h];
        for (int32_t q : cands)
            if (same(p, q)) {
                rep[p] = q;
                any = true;
                break;
            }
        if (rep[p] == p) cands.push_back(p);
    }
    if (!any) return SIGP_OK;
    HIP_TRY(hipMalloc((void **)&bt->rep_owned, (size_t)P * 4));
    HIP_TRY(hipMemcpy(bt->rep_owned, rep.data(), (size_t)P * 4, hipMemcpyHostToDevice));
    bt->rep = bt->rep_owned;
    return SIGP_OK;
}

// After a launch: each skipped pair (k, p) takes the rewards of (k, rep[p]).
__global__ __launch_bounds__(256) void fill_repeats_kernel(const int32_t *__restrict__ rep, int P, int T,
                                                           double *__restrict__ out) {
    const int p = (int)blockIdx.x, k = (int)blockIdx.y;
    const int r = rep[p];
    if (r == p) return;
    double *base = out + (size_t)k * P * T;
    for (int s = (int)threadIdx.x; s < T; s += (int)blockDim.x) base[(size_t)p * T + s] = base[(size_t)r * T + s];
}

// Layout of a batch's device block (rules, instance slots, launch orders)
// and its host image, computed without touching the device.
struct BatchPlan {
    std::vector<InstSlot> slots;
    std::vector<int32_t> orders[3], samples[3], checks[3];
    size_t o_ops = 0, o_args = 0, o_off = 0, o_len = 0, o_rd = 0, o_fl = 0, o_slots = 0;
    size_t o_ord[3] = {0, 0, 0}, o_smp[3] = {0, 0, 0}, o_chk[3] = {0, 0, 0};
    size_t bytes = 0;
    long long perms = 0;
};

// The round-robin protocol keeps 2 * (largest m) LDS rows per wavefront.
static sigp_status set_protocol(sigp_batch *bt, int protocol) {
    if (protocol != SIGP_PROTOCOL_VEHICLES && protocol != SIGP_PROTOCOL_ROUND_ROBIN)
        return fail(SIGP_INVALID_ARG, "protocol must be SIGP_PROTOCOL_VEHICLES or SIGP_PROTOCOL_ROUND_ROBIN");
    int rr = 0;
    if (protocol == SIGP_PROTOCOL_ROUND_ROBIN) {
        rr = 1;
        for (const sigp_instance *inst : bt->insts) {
            if (inst->m > SIGP_RR_MAX_M)
                return fail(SIGP_INVALID_ARG, "round robin simulates at most SIGP_RR_MAX_M = " +
                                                  std::to_string(SIGP_RR_MAX_M) + " vehicles; an instance has m = " +
                                                  std::to_string(inst->m));
            rr = std::max(rr, inst->m);
        }
    }
    bt->rr = rr;
    return SIGP_OK;
}

// Fills bt's host fields (shape, size classes, variant flags) and the plan.
static sigp_status plan_batch(const std::vector<sigp_instance *> &insts, bool multi, const PreparedRules &pr,
                              int T, int group_size, sigp_batch *bt, BatchPlan &pl) {
    const int P = pr.P, K = (int)insts.size();
    if ((long long)K * P > 0x7fffffffLL) return fail(SIGP_INVALID_ARG, "instances x rules exceeds 2^31");
    bt->insts = insts;
    bt->multi = multi;
    bt->P = P;
    bt->T = T;
    bt->group_size = group_size;
    bt->waves_per_item = insts[0]->waves_per_item;
    bt->mapping = insts[0]->mapping;
    bt->rows_needed = P == 0 ? 0 : (multi ? (int64_t)K * T : (int64_t)((P - 1) / group_size + 1) * T);
    bt->n_out = (long long)K * P * T;
    bt->has_rmod = pr.any_rmod;
    bt->has_rare = (pr.any_rd & RARE_SLOTS) != 0;
    for (const sigp_instance *inst : insts)
        if (inst->protocol != insts[0]->protocol)
            return fail(SIGP_INVALID_ARG, "the instances of a batch must share one protocol (sigp_instance_protocol)");
    sigp_status pst = set_protocol(bt, insts[0]->protocol);
    if (pst != SIGP_OK) return pst;

    pl.slots.resize(K);
    long long samp = 0, perms = 0;
    bt->G = multi || P == 0 ? 1 : (P - 1) / group_size + 1;
    for (int k = 0; k < K; ++k) {
        pl.slots[k].dev = insts[k]->dev;
        pl.slots[k].tabs = insts[k]->tabs;
        pl.slots[k].samp_off = samp;
        pl.slots[k].out_off = (long long)k * P * T;
        pl.slots[k].perm_off = perms;
        if (multi) samp += (long long)T * insts[k]->n;
        perms += (long long)bt->G * insts[k]->n * T;
    }
    pl.perms = perms;
    const int classes[3] = {4, 9, 16};
    // One launch over every instance when all of them ask for it: every
    // instance then runs in the widest class's kernels (a narrower visited
    // set is the wider one with its top words zero).
    bool one = multi;
    int widest = 4;
    for (const sigp_instance *inst : insts) {
        one = one && inst->one_launch;
        widest = std::max(widest, nw_bucket(inst->n));
    }
    auto class_of = [&](int k) { return one ? widest : nw_bucket(insts[k]->n); };
    std::vector<int32_t> *orders = pl.orders, *samples = pl.samples, *checks = pl.checks;
    // the autotuner times each mapping on about 2^19 simulations of a class,
    // and checks a close ranking on about 2^21
    const long long tune_cap = std::max(1LL, (1LL << 19) / std::max(T, 1));
    const long long check_cap = 4 * tune_cap;
    for (int c = 0; c < 3; ++c) {
        LaunchUnit &u = bt->units[bt->n_units];
        std::vector<std::pair<float, int32_t>> items;
        for (int k = 0; k < K; ++k) {
            if (class_of(k) != classes[c]) continue;
            u.max_n = std::max(u.max_n, insts[k]->n);
            const float nn = (float)insts[k]->n * (float)insts[k]->n;
            for (int p = 0; p < P; ++p) items.push_back({pr.cost[p] * nn, (int32_t)(k * P + p)});
        }
        if (items.empty()) continue;
        std::stable_sort(items.begin(), items.end(),
                         [](const std::pair<float, int32_t> &a, const std::pair<float, int32_t> &b) {
                             return a.first > b.first;
                         });
        for (const auto &it : items) orders[bt->n_units].push_back(it.second);
        const long long stride = ((long long)items.size() + tune_cap - 1) / tune_cap;
        for (size_t i = 0; i < items.size(); i += (size_t)stride) samples[bt->n_units].push_back(items[i].second);
        const long long stride4 = ((long long)items.size() + check_cap - 1) / check_cap;
        for (size_t i = 0; i < items.size(); i += (size_t)stride4) checks[bt->n_units].push_back(items[i].second);
        u.nwb = classes[c];
        u.n_pairs = (long long)items.size();
        u.tune_pairs = (long long)samples[bt->n_units].size();
        u.check_pairs = (long long)checks[bt->n_units].size();
        ++bt->n_units;
    }

    // heaviest class first: it starts first and the others fill in beside it
    {
        int idx[3] = {0, 1, 2};
        auto work = [&](int u) {
            return (double)bt->units[u].n_pairs * bt->units[u].max_n * bt->units[u].max_n;
        };
        std::stable_sort(idx, idx + bt->n_units, [&](int a, int b) { return work(a) > work(b); });
        LaunchUnit us[3];
        std::vector<int32_t> os[3], ss[3], cs[3];
        for (int u = 0; u < bt->n_units; ++u) {
            us[u] = bt->units[idx[u]];
            os[u].swap(orders[idx[u]]);
            ss[u].swap(samples[idx[u]]);
            cs[u].swap(checks[idx[u]]);
        }
        for (int u = 0; u < bt->n_units; ++u) {
            bt->units[u] = us[u];
            orders[u].swap(os[u]);
            samples[u].swap(ss[u]);
            checks[u].swap(cs[u]);
        }
    }
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes + 8); return o; };
    pl.o_ops = take(pr.ops.size() * 4);
    pl.o_args = take(pr.args.size() * 4);
    pl.o_off = take((size_t)P * 4);
    pl.o_len = take((size_t)P * 4);
    pl.o_rd = take((size_t)P * 8);
    pl.o_fl = take((size_t)P * 8);
    pl.o_slots = take((size_t)K * sizeof(InstSlot));
    for (int u = 0; u < bt->n_units; ++u) {
        pl.o_ord[u] = take(orders[u].size() * 4);
        pl.o_smp[u] = take(samples[u].size() * 4);
        pl.o_chk[u] = take(checks[u].size() * 4);
    }
    pl.bytes = off;
    return SIGP_OK;
}

// The device block's contents, laid out as planned, in host memory.
static std::vector<char> batch_image(const sigp_batch *bt, const BatchPlan &pl, const PreparedRules &pr) {
    std::vector<char> img(pl.bytes, 0);
    auto put = [&](size_t o, const void *src, size_t bytes) {
        if (bytes) std::memcpy(img.data() + o, src, bytes);
    };
    const int P = pr.P;
    put(pl.o_ops, pr.ops.data(), pr.ops.size() * 4);
    put(pl.o_args, pr.args.data(), pr.args.size() * 4);
    put(pl.o_off, pr.offs.data(), (size_t)P * 4);
    put(pl.o_len, pr.lens.data(), (size_t)P * 4);
    put(pl.o_rd, pr.rd.data(), (size_t)P * 8);
    put(pl.o_fl, pr.flags.data(), (size_t)P * 8);
    put(pl.o_slots, pl.slots.data(), pl.slots.size() * sizeof(InstSlot));
    for (int u = 0; u < bt->n_units; ++u) {
        put(pl.o_ord[u], pl.orders[u].data(), pl.orders[u].size() * 4);
        put(pl.o_smp[u], pl.samples[u].data(), pl.samples[u].size() * 4);
        put(pl.o_chk[u], pl.checks[u].data(), pl.checks[u].size() * 4);
    }
    return img;
}

// Points bt's rule set, slots and launch orders into a device block `base`
// holding batch_image's contents.
static void bind_batch(sigp_batch *bt, char *base, const BatchPlan &pl, const PreparedRules &pr) {
    bt->rs.ops = (cip)(base + pl.o_ops);
    bt->rs.args = (cip)(base + pl.o_args);
    bt->rs.off = (cip)(base + pl.o_off);
    bt->rs.len = (cip)(base + pl.o_len);
    bt->rs.rd = (cup)(base + pl.o_rd);
    bt->rs.flags = (cup)(base + pl.o_fl);
    bt->rs.order = nullptr;
    bt->rs.P = pr.P;
    bt->rs.KL = pr.max_k;
    bt->rs.SL = std::max(1, pr.max_depth - 1);
    bt->slots = (cslot)(base + pl.o_slots);
    for (int u = 0; u < bt->n_units; ++u) {
        bt->units[u].order = (cip)(base + pl.o_ord[u]);
        bt->units[u].tune_order = (cip)(base + pl.o_smp[u]);
        bt->units[u].check_order = (cip)(base + pl.o_chk[u]);
    }
}

static sigp_status make_batch(const std::vector<sigp_instance *> &insts, bool multi,
                              const PreparedRules &pr, int T, int group_size, sigp_batch **out,
                              bool order = true) {
    sigp_batch *bt = new sigp_batch();
    BatchPlan pl;
    sigp_status st = plan_batch(insts, multi, pr, T, group_size, bt, pl);
    if (st != SIGP_OK) {
        delete bt;
        return st;
    }
    const std::vector<char> img = batch_image(bt, pl, pr);
    char *base = nullptr;
    if (hipMalloc((void **)&base, pl.bytes) != hipSuccess) {
        delete bt;
        return fail(SIGP_OOM, "batch allocation failed");
    }
    bt->dbuf = base;
    if (hipMemcpy(base, img.data(), pl.bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(base);
        delete bt;
        return fail(SIGP_DEVICE_ERROR, "batch upload failed");
    }
    bind_batch(bt, base, pl, pr);
    bool skip = pr.P > 1;
    for (const sigp_instance *inst : insts) skip = skip && inst->skip_repeats;
    if (skip) {
        st = find_repeats(bt, pr, group_size);
        if (st != SIGP_OK) {
            sigp_batch_destroy(bt);
            return st;
        }
    }
    if (order && pr.P > 0 && T >= 64 && T <= ORDER_MAX_T && pl.perms * 4 <= ORDER_MAX_BYTES) {
        st = setup_order(bt, pl.perms);
        if (st != SIGP_OK) {
            sigp_batch_destroy(bt);
            return st;
        }
    }
    *out = bt;
    return SIGP_OK;
}

extern "C" sigp_status sigp_batch_create(sigp_instance *inst, const sigp_rules *rules, int32_t T,
                                         int32_t group_size, const int64_t *masks, int32_t n_masks,
                                         int32_t mask_group, sigp_batch **out) {
    if (!out) return fail(SIGP_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!inst || !rules) return fail(SIGP_INVALID_ARG, "inst and rules must be non-NULL");
    if (T < 0) return fail(SIGP_INVALID_ARG, "T < 0");
    if (group_size < 1) return fail(SIGP_INVALID_ARG, "group_size must be >= 1");
    PreparedRules pr;
    sigp_status st = prepare_rules(rules, masks, n_masks, mask_group, pr);
    if (st != SIGP_OK) return st;
    return make_batch({inst}, false, pr, T, group_size, out);
}

extern "C" sigp_status sigp_batch_create_multi(sigp_instance *const *insts, int32_t n_inst,
                                               const sigp_rules *rules, int32_t T, const int64_t *masks,
                                               int32_t n_masks, int32_t mask_group, sigp_batch **out) {
    if (!out) return fail(SIGP_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!insts || n_inst < 1 || !rules) return fail(SIGP_INVALID_ARG, "instances and rules must be given");
    if (T < 0) return fail(SIGP_INVALID_ARG, "T < 0");
    std::vector<sigp_instance *> v(insts, insts + n_inst);
    for (sigp_instance *i : v)
        if (!i) return fail(SIGP_INVALID_ARG, "an instance is NULL");
    PreparedRules pr;
    sigp_status st = prepare_rules(rules, masks, n_masks, mask_group, pr);
    if (st != SIGP_OK) return st;
    return make_batch(v, true, pr, T, std::max(pr.P, 1), out);
}

extern "C" sigp_status sigp_batch_destroy(sigp_batch *bt) {
    if (!bt) return SIGP_OK;
    for (int i = 0; i < 2; ++i) {
        if (bt->side[i]) (void)hipStreamDestroy(bt->side[i]);
        if (bt->join[i]) (void)hipEventDestroy(bt->join[i]);
    }
    if (bt->fork) (void)hipEventDestroy(bt->fork);
    if (bt->dbuf) (void)hipFree(bt->dbuf);
    if (bt->pairs_tmp) (void)hipFree(bt->pairs_tmp);
    if (bt->first) (void)hipFree(bt->first);
    if (bt->perm) (void)hipFree(bt->perm);
    if (bt->rep_owned) (void)hipFree(bt->rep_owned);
    delete bt;
    return SIGP_OK;
}

static sigp_status check_config(const sigp_launch_config *cfg) {
    const int w = cfg->waves_per_item;
    if (w != 0 && w != 1 && w != 2 && w != 4)
        return fail(SIGP_INVALID_ARG, "waves_per_item must be 0, 1, 2 or 4");
    if (cfg->mapping < SIGP_MAP_AUTO || cfg->mapping > SIGP_MAP_CANDIDATES_2)
        return fail(SIGP_INVALID_ARG, "mapping must be SIGP_MAP_AUTO, _SCENARIOS, _CANDIDATES or _CANDIDATES_2");
    return SIGP_OK;
}

extern "C" sigp_status sigp_instance_skip_repeats(sigp_instance *inst, int32_t enable) {
    if (!inst) return fail(SIGP_INVALID_ARG, "instance is NULL");
    inst->skip_repeats = enable != 0;
    return SIGP_OK;
}

extern "C" sigp_status sigp_instance_one_launch(sigp_instance *inst, int32_t enable) {
    if (!inst) return fail(SIGP_INVALID_ARG, "instance is NULL");
    inst->one_launch = enable != 0;
    return SIGP_OK;
}

extern "C" sigp_status sigp_instance_protocol(sigp_instance *inst, int32_t protocol) {
    if (!inst) return fail(SIGP_INVALID_ARG, "instance is NULL");
    if (protocol != SIGP_PROTOCOL_VEHICLES && protocol != SIGP_PROTOCOL_ROUND_ROBIN)
        return fail(SIGP_INVALID_ARG, "protocol must be SIGP_PROTOCOL_VEHICLES or SIGP_PROTOCOL_ROUND_ROBIN");
    if (protocol == SIGP_PROTOCOL_ROUND_ROBIN && inst->m > SIGP_RR_MAX_M)
        return fail(SIGP_INVALID_ARG, "round robin simulates at most SIGP_RR_MAX_M = " +
                                          std::to_string(SIGP_RR_MAX_M) + " vehicles");
    inst->protocol = protocol;
    return SIGP_OK;
}

extern "C" sigp_status sigp_batch_protocol(sigp_batch *bt, int32_t protocol) {
    if (!bt) return fail(SIGP_INVALID_ARG, "batch is NULL");
    return set_protocol(bt, protocol);
}

extern "C" sigp_status sigp_batch_order_scenarios(sigp_batch *bt, int32_t enable, int32_t *active) {
    if (!bt) return fail(SIGP_INVALID_ARG, "batch is NULL");
    bt->order_on = enable != 0 && bt->perm != nullptr;
    if (active) *active = bt->order_on ? 1 : 0;
    return SIGP_OK;
}

extern "C" sigp_status sigp_batch_configure(sigp_batch *bt, const sigp_launch_config *cfg) {
    if (!bt || !cfg) return fail(SIGP_INVALID_ARG, "batch and cfg must be non-NULL");
    const sigp_status st = check_config(cfg);
    if (st != SIGP_OK) return st;
    bt->waves_per_item = cfg->waves_per_item;
    bt->mapping = cfg->mapping;
    for (int u = 0; u < bt->n_units; ++u) {   // drop per-class tuning
        bt->units[u].mapping = SIGP_MAP_AUTO;
        bt->units[u].waves = 0;
    }
    return SIGP_OK;
}

extern "C" sigp_status sigp_instance_configure(sigp_instance *inst, const sigp_launch_config *cfg) {
    if (!inst || !cfg) return fail(SIGP_INVALID_ARG, "inst and cfg must be non-NULL");
    const sigp_status st = check_config(cfg);
    if (st != SIGP_OK) return st;
    inst->waves_per_item = cfg->waves_per_item;
    inst->mapping = cfg->mapping;
    return SIGP_OK;
}

extern "C" sigp_status sigp_batch_shape(const sigp_batch *bt, int32_t *n_rules, int64_t *n_pairs) {
    if (!bt) return fail(SIGP_INVALID_ARG, "batch is NULL");
    if (n_rules) *n_rules = bt->P;
    if (n_pairs) *n_pairs = bt->n_out;
    return SIGP_OK;
}

// LDS of one workgroup: the merge arrays plus W wavefronts' value stores and
// stacks (and round-robin route states), kept within 64 KiB.  W = 1 always
// fits: (61 + 31 + 2 * SIGP_RR_MAX_M) rows * 512 B.
static constexpr size_t LDS_LIMIT = 65536, LDS_STATIC = 4 * 64 * (8 + 4);
static_assert(LDS_STATIC + (SIGP_N_TERMS + SIGP_MAX_STACK - 1 + 2 * SIGP_RR_MAX_M) * 512 <= LDS_LIMIT,
              "one wavefront's LDS rows must fit");
static size_t lds_per_wave(const sigp_batch *bt) {
    return (size_t)(bt->rs.KL + bt->rs.SL + 2 * bt->rr) * 64 * sizeof(double);
}
static int fit_waves(const sigp_batch *bt, int w) {
    while (w > 1 && LDS_STATIC + (size_t)w * lds_per_wave(bt) > LDS_LIMIT) w >>= 1;
    return w;
}

// A size class's tuned mapping, else the batch's; AUTO: candidate lanes below
// 64 scenarios per rule (less than one full block of scenario lanes; the GP
// loop's shapes measured 64 ms with candidate lanes against about 300 ms with
// scenario lanes at 50 scenarios) or when scenario lanes would make fewer
// than 512 workgroups (too few to fill 256 CUs, e.g. one rule scored on a
// held-out set), scenario lanes otherwise.
static int resolved_mapping(const sigp_batch *bt, const LaunchUnit &u) {
    if (u.mapping != SIGP_MAP_AUTO) return u.mapping;
    if (bt->mapping != SIGP_MAP_AUTO) return bt->mapping;
    const long long scen_items = u.n_pairs * ((bt->T + 63) / 64);
    return (bt->T < 64 || scen_items < 512) ? SIGP_MAP_CANDIDATES : SIGP_MAP_SCENARIOS;
}

// Wavefronts per workgroup: the tuned or configured count, else 4 for
// scenario lanes and 1 for candidate lanes (one wavefront making several
// passes over the candidates measured fastest: at n = 100 and at n = 288,
// 1,024 / 4,096 rules x 50 scenarios, 1 wavefront beat 2 and 4 by 15-60 %);
// then reduced until LDS fits.
static int resolved_waves(const sigp_batch *bt, const LaunchUnit &u, int mapping) {
    if (mapping == SIGP_MAP_CANDIDATES_2) return 1;
    int w = u.waves ? u.waves : bt->waves_per_item;
    if (w == 0) w = mapping == SIGP_MAP_CANDIDATES ? 1 : 4;
    return fit_waves(bt, w);
}

template <int NW, bool RMOD, bool RARE>
static void launch_variant(const sigp_batch *bt, const LaunchUnit &u, int mapping, int w,
                           unsigned items, const ScenSpec &S, double *out, hipStream_t stream) {
    RuleSet R = bt->rs;
    R.order = u.order;
    const unsigned lds = (unsigned)(w * lds_per_wave(bt));
    if constexpr (NW <= 9) {
        if (mapping == SIGP_MAP_CANDIDATES_2) {
            sim_seg_kernel<NW, RMOD, RARE, 32><<<items, 64, lds, stream>>>(bt->slots, R, S, out);
            return;
        }
    }
    if (mapping != SIGP_MAP_SCENARIOS)
        sim_cand_kernel<NW, RMOD, RARE><<<items, 64 * w, lds, stream>>>(bt->slots, R, S, out);
    else
        sim_kernel<NW, RMOD, RARE><<<items, 64 * w, lds, stream>>>(bt->slots, R, S, out);
}

template <int NW>
static void launch_class(const sigp_batch *bt, const LaunchUnit &u, int mapping, int w, unsigned items,
                         const ScenSpec &S, double *out, hipStream_t stream) {
    if (bt->has_rmod) {
        if (bt->has_rare) launch_variant<NW, true, true>(bt, u, mapping, w, items, S, out, stream);
        else launch_variant<NW, true, false>(bt, u, mapping, w, items, S, out, stream);
    } else {
        if (bt->has_rare) launch_variant<NW, false, true>(bt, u, mapping, w, items, S, out, stream);
        else launch_variant<NW, false, false>(bt, u, mapping, w, items, S, out, stream);
    }
}

// Each pair's first customer (the first decision reads no sampled duration,
// so it is the same in every scenario): candidate lanes in first-only mode,
// one workgroup per pair, at batch creation.
static sigp_status setup_order(sigp_batch *bt, long long perms) {
    const size_t pairs = (size_t)bt->insts.size() * bt->P;
    HIP_TRY(hipMalloc((void **)&bt->first, pairs * 4));
    HIP_TRY(hipMalloc((void **)&bt->perm, (size_t)perms * 4));
    for (int u = 0; u < bt->n_units; ++u) {
        const LaunchUnit &lu = bt->units[u];
        ScenSpec S{};
        S.T = 1;
        S.group_size = bt->group_size;
        S.nblk = 1;
        S.first = bt->first;
        S.first_only = 1;
        if (lu.nwb == 4) launch_class<4>(bt, lu, SIGP_MAP_CANDIDATES, 1, (unsigned)lu.n_pairs, S, nullptr, nullptr);
        else if (lu.nwb == 9) launch_class<9>(bt, lu, SIGP_MAP_CANDIDATES, 1, (unsigned)lu.n_pairs, S, nullptr, nullptr);
        else launch_class<16>(bt, lu, SIGP_MAP_CANDIDATES, 1, (unsigned)lu.n_pairs, S, nullptr, nullptr);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipDeviceSynchronize());
    bt->order_on = true;
    return SIGP_OK;
}

static bool uses_scenario_lanes(const sigp_batch *bt) {
    for (int u = 0; u < bt->n_units; ++u) {
        const int m = resolved_mapping(bt, bt->units[u]);
        if (m == SIGP_MAP_SCENARIOS || (m == SIGP_MAP_CANDIDATES_2 && bt->units[u].nwb <= 9)) return true;
    }
    return false;
}

// Per run: the order of each instance's scenario groups by each customer's
// duration (order_rows_kernel), on `stream` before the simulation.
static sigp_status order_scenarios(const sigp_batch *bt, const double *sampled, hipStream_t stream) {
    long long samp = 0, perms = 0;
    for (const sigp_instance *inst : bt->insts) {
        const int n = inst->n;
        if (n > 0) order_rows_kernel<<<(unsigned)(bt->G * n), 256, 0, stream>>>(sampled + samp, n, bt->T, bt->perm + perms);
        if (bt->multi) samp += (long long)bt->T * n;
        perms += (long long)bt->G * n * bt->T;
    }
    HIP_TRY(hipGetLastError());
    return SIGP_OK;
}

static sigp_status launch_unit(const sigp_batch *bt, const LaunchUnit &u, const double *sampled,
                               double *out, hipStream_t stream) {
    ScenSpec S{};
    S.sampled = sampled;
    S.T = bt->T;
    S.group_size = bt->group_size;
    S.nblk = (bt->T + 63) / 64;
    int mapping = resolved_mapping(bt, u);
    S.rep = bt->rep;
    S.rr = bt->rr;
    if (bt->order_on && (mapping == SIGP_MAP_SCENARIOS || mapping == SIGP_MAP_CANDIDATES_2)) {
        S.perm = bt->perm;
        S.first = bt->first;
    }
    if (u.nwb > 9 && mapping == SIGP_MAP_CANDIDATES_2)
        mapping = SIGP_MAP_CANDIDATES;   // segments are built for n <= 288 only
    const int w = resolved_waves(bt, u, mapping);
    S.n_ranks = u.n_pairs;
    const long long items =
        mapping == SIGP_MAP_CANDIDATES   ? u.n_pairs * (long long)S.T
        : mapping == SIGP_MAP_SCENARIOS  ? (u.n_pairs + N_XCD - 1) / N_XCD * N_XCD * (long long)S.nblk
                                         : u.n_pairs * (long long)((S.T + 1) / 2);
    if (items > 0xffffffffLL) return fail(SIGP_INVALID_ARG, "launch exceeds 2^32 workgroups");
    if (u.nwb == 4) launch_class<4>(bt, u, mapping, w, (unsigned)items, S, out, stream);
    else if (u.nwb == 9) launch_class<9>(bt, u, mapping, w, (unsigned)items, S, out, stream);
    else launch_class<16>(bt, u, mapping, w, (unsigned)items, S, out, stream);
    HIP_TRY(hipGetLastError());
    return SIGP_OK;
}

extern "C" sigp_status sigp_batch_run(sigp_batch *bt, const double *sampled_dev, int64_t n_rows,
                                      double *out_pairs_dev, double *out_mean_dev, void *stream) {
    if (!bt) return fail(SIGP_INVALID_ARG, "batch is NULL");
    if (bt->P == 0 || bt->T == 0) {
        if (out_mean_dev && bt->P > 0 && bt->T == 0)
            return fail(SIGP_INVALID_ARG, "mean over T = 0 scenarios");
        return SIGP_OK;
    }
    if (!sampled_dev) return fail(SIGP_INVALID_ARG, "sampled_dev is NULL");
    if (n_rows < bt->rows_needed)
        return fail(SIGP_INVALID_ARG, "sampled has " + std::to_string(n_rows) + " rows, rules need " +
                                          std::to_string(bt->rows_needed));
    hipStream_t s = (hipStream_t)stream;
    double *pairs = out_pairs_dev;
    if (!pairs) {
        if (!out_mean_dev) return SIGP_OK;
        if (!bt->pairs_tmp) HIP_TRY(hipMalloc((void **)&bt->pairs_tmp, (size_t)bt->n_out * 8));
        pairs = bt->pairs_tmp;
    }
    if (bt->order_on && uses_scenario_lanes(bt)) {
        const sigp_status st = order_scenarios(bt, sampled_dev, s);
        if (st != SIGP_OK) return st;
    }
    if (bt->n_units == 1) {
        const sigp_status st = launch_unit(bt, bt->units[0], sampled_dev, pairs, s);
        if (st != SIGP_OK) return st;
    } else {
        // Size classes run concurrently: the first on the caller's stream, the
        // others on the batch's own streams, forked from and joined back into
        // the caller's stream with events (default-priority streams).
        for (int i = 0; i + 1 < bt->n_units; ++i) {
            if (!bt->side[i]) HIP_TRY(hipStreamCreateWithFlags(&bt->side[i], hipStreamNonBlocking));
            if (!bt->join[i]) HIP_TRY(hipEventCreateWithFlags(&bt->join[i], hipEventDisableTiming));
        }
        if (!bt->fork) HIP_TRY(hipEventCreateWithFlags(&bt->fork, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(bt->fork, s));
        for (int u = 0; u < bt->n_units; ++u) {
            hipStream_t su = u == 0 ? s : bt->side[u - 1];
            if (u > 0) HIP_TRY(hipStreamWaitEvent(su, bt->fork, 0));
            const sigp_status st = launch_unit(bt, bt->units[u], sampled_dev, pairs, su);
            if (st != SIGP_OK) return st;
            if (u > 0) HIP_TRY(hipEventRecord(bt->join[u - 1], su));
        }
        for (int u = 1; u < bt->n_units; ++u) HIP_TRY(hipStreamWaitEvent(s, bt->join[u - 1], 0));
    }
    if (bt->rep_owned)
        fill_repeats_kernel<<<dim3((unsigned)bt->P, (unsigned)bt->insts.size()), 256, 0, s>>>(bt->rep_owned, bt->P,
                                                                                            bt->T, pairs);
    if (out_mean_dev) {
        const int rows = (int)(bt->n_out / bt->T);   // instances x rules
        mean_kernel<<<(rows + 255) / 256, 256, 0, s>>>(pairs, rows, bt->T, out_mean_dev);
        HIP_TRY(hipGetLastError());
    }
    return SIGP_OK;
}

// Times every mapping of one size class on its sample of pairs (twice each:
// the first run of a kernel loads its code object) and keeps the fastest.
// Repeated rules (bt->rep) return at once here as in the runs being tuned,
// so the sample's time reflects the work the batch will do.
static sigp_status tune_unit(sigp_batch *bt, LaunchUnit &u, const double *sampled, hipStream_t s,
                             hipEvent_t e0, hipEvent_t e1) {
    LaunchUnit probe = u;
    probe.order = u.tune_order;
    probe.n_pairs = u.tune_pairs;
    struct Cand { int mapping, w; };
    const Cand cands[7] = {{SIGP_MAP_SCENARIOS, 1},  {SIGP_MAP_SCENARIOS, 2},  {SIGP_MAP_SCENARIOS, 4},
                           {SIGP_MAP_CANDIDATES, 1}, {SIGP_MAP_CANDIDATES, 2}, {SIGP_MAP_CANDIDATES, 4},
                           {SIGP_MAP_CANDIDATES_2, 1}};
    const int groups = (u.max_n + 63) / 64;
    int best_w = 0, best_m = SIGP_MAP_AUTO;
    float best_ms = 0.0f;
    float map_ms[4] = {0.0f, 0.0f, 0.0f, 0.0f};   // best time of each mapping (SIGP_MAP_* 1..3)
    int map_w[4] = {0, 0, 0, 0};
    for (const Cand &c : cands) {
        if (fit_waves(bt, c.w) != c.w) continue;                           // LDS does not fit
        if (c.mapping == SIGP_MAP_CANDIDATES_2 && u.nwb > 9) continue;      // built for n <= 288
        if (c.mapping == SIGP_MAP_CANDIDATES && c.w > 1 && 32 * c.w >= groups * 64) continue;
        probe.mapping = c.mapping;
        probe.waves = c.w;
        float ms = 0.0f;
        for (int rep = 0; rep < 2; ++rep) {
            HIP_TRY(hipEventRecord(e0, s));
            const sigp_status st = launch_unit(bt, probe, sampled, bt->pairs_tmp, s);
            if (st != SIGP_OK) return st;
            HIP_TRY(hipEventRecord(e1, s));
            HIP_TRY(hipEventSynchronize(e1));
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        }
        if (best_w == 0 || ms < best_ms) { best_ms = ms; best_w = c.w; best_m = c.mapping; }
        if (map_w[c.mapping] == 0 || ms < map_ms[c.mapping]) { map_ms[c.mapping] = ms; map_w[c.mapping] = c.w; }
    }
    // The sample can rank scenario lanes and candidate lanes the wrong way
    // round when they are close: a small sample is bound by its heaviest
    // rules' latency, which candidate lanes (a workgroup per scenario) hide
    // better, a whole class by throughput.  (The 40-instance workload's
    // n <= 288 class timed candidate lanes 12 % ahead on 2^19 simulations and
    // 35 % behind on the whole batch.)  So the best of each family, within
    // 25 % of each other, are timed once more on a 4x larger sample.
    {
        const int cand = map_w[SIGP_MAP_CANDIDATES_2] &&
                                 (!map_w[SIGP_MAP_CANDIDATES] || map_ms[SIGP_MAP_CANDIDATES_2] < map_ms[SIGP_MAP_CANDIDATES])
                             ? SIGP_MAP_CANDIDATES_2 : SIGP_MAP_CANDIDATES;
        const int scen = SIGP_MAP_SCENARIOS;
        if (u.check_pairs > u.tune_pairs && map_w[scen] && map_w[cand] &&
            std::max(map_ms[scen], map_ms[cand]) < 1.25f * std::min(map_ms[scen], map_ms[cand])) {
            LaunchUnit chk = u;
            chk.order = u.check_order;
            chk.n_pairs = u.check_pairs;
            const int ms_m[2] = {scen, cand};
            float t[2] = {0.0f, 0.0f};
            for (int i = 0; i < 2; ++i) {
                chk.mapping = ms_m[i];
                chk.waves = map_w[ms_m[i]];
                HIP_TRY(hipEventRecord(e0, s));
                const sigp_status st = launch_unit(bt, chk, sampled, bt->pairs_tmp, s);
                if (st != SIGP_OK) return st;
                HIP_TRY(hipEventRecord(e1, s));
                HIP_TRY(hipEventSynchronize(e1));
                HIP_TRY(hipEventElapsedTime(&t[i], e0, e1));
            }
            const int win = t[0] <= t[1] ? scen : cand;
            best_m = win;
            best_w = map_w[win];
        }
    }
    u.mapping = best_m;
    u.waves = best_w;
    return SIGP_OK;
}

extern "C" sigp_status sigp_batch_autotune(sigp_batch *bt, const double *sampled_dev, int64_t n_rows,
                                           void *stream, sigp_launch_config *cfg_out) {
    if (!bt) return fail(SIGP_INVALID_ARG, "batch is NULL");
    if (bt->P == 0 || bt->T == 0 || bt->n_units == 0) {
        if (cfg_out) { cfg_out->waves_per_item = bt->waves_per_item; cfg_out->mapping = bt->mapping; }
        return SIGP_OK;
    }
    if (!sampled_dev) return fail(SIGP_INVALID_ARG, "sampled_dev is NULL");
    if (n_rows < bt->rows_needed) return fail(SIGP_INVALID_ARG, "sampled has too few rows");
    if (!bt->pairs_tmp) HIP_TRY(hipMalloc((void **)&bt->pairs_tmp, (size_t)bt->n_out * 8));
    hipStream_t s = (hipStream_t)stream;
    if (bt->order_on) {
        const sigp_status st = order_scenarios(bt, sampled_dev, s);
        if (st != SIGP_OK) return st;
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(SIGP_DEVICE_ERROR, "hipEventCreate failed");
    }
    sigp_status st = SIGP_OK;
    int heaviest = 0;
    double heavy_w = -1.0;
    for (int u = 0; u < bt->n_units && st == SIGP_OK; ++u) {
        st = tune_unit(bt, bt->units[u], sampled_dev, s, e0, e1);
        const double w = (double)bt->units[u].n_pairs * bt->units[u].max_n * bt->units[u].max_n;
        if (w > heavy_w) { heavy_w = w; heaviest = u; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != SIGP_OK) {
        for (int u = 0; u < bt->n_units; ++u) { bt->units[u].mapping = SIGP_MAP_AUTO; bt->units[u].waves = 0; }
        return st;
    }
    if (cfg_out) {
        cfg_out->waves_per_item = bt->units[heaviest].waves;
        cfg_out->mapping = bt->units[heaviest].mapping;
    }
    return SIGP_OK;
}

static sigp_status ensure(double **buf, size_t *cap, size_t bytes) {
    if (*cap >= bytes && *buf) return SIGP_OK;
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    HIP_TRY(hipMalloc((void **)buf, bytes ? bytes : 8));
    *cap = bytes;
    return SIGP_OK;
}

// Uploads `sbytes` of scenarios, runs the batch, downloads results; the
// buffers are the first instance's cached staging buffers.
static sigp_status run_host(sigp_batch *bt, const double *sampled, size_t sbytes, int64_t n_rows,
                            double *out_pairs, double *out_mean) {
    sigp_instance *inst = bt->insts[0];
    const size_t obytes = (size_t)bt->n_out * 8, mbytes = (size_t)(bt->T ? bt->n_out / bt->T : 0) * 8;
    sigp_status st;
    if ((st = ensure(&inst->scen_buf, &inst->scen_cap, sbytes)) != SIGP_OK ||
        (st = ensure(&inst->out_buf, &inst->out_cap, obytes)) != SIGP_OK ||
        (st = ensure(&inst->mean_buf, &inst->mean_cap, mbytes)) != SIGP_OK)
        return st;
    hipError_t e = sbytes ? hipMemcpy(inst->scen_buf, sampled, sbytes, hipMemcpyHostToDevice) : hipSuccess;
    if (e != hipSuccess) return fail(SIGP_DEVICE_ERROR, "scenario upload failed");
    st = sigp_batch_run(bt, inst->scen_buf, n_rows, inst->out_buf, out_mean ? inst->mean_buf : nullptr, nullptr);
    if (st != SIGP_OK) return st;
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(SIGP_DEVICE_ERROR, std::string("kernel failed: ") + hipGetErrorString(e));
    if (out_pairs && obytes && hipMemcpy(out_pairs, inst->out_buf, obytes, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(SIGP_DEVICE_ERROR, "result download failed");
    if (out_mean && mbytes && hipMemcpy(out_mean, inst->mean_buf, mbytes, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(SIGP_DEVICE_ERROR, "mean download failed");
    return SIGP_OK;
}

extern "C" sigp_status sigp_evaluate(sigp_instance *inst, const sigp_rules *rules, const double *sampled,
                                     int64_t n_rows, int32_t T, int32_t group_size, const int64_t *masks,
                                     int32_t n_masks, int32_t mask_group, double *out_pairs,
                                     double *out_mean) {
    if (!inst) return fail(SIGP_INVALID_ARG, "inst is NULL");
    sigp_batch *bt = nullptr;
    sigp_status st = sigp_batch_create(inst, rules, T, group_size, masks, n_masks, mask_group, &bt);
    if (st != SIGP_OK) return st;
    if (bt->P == 0 || T == 0) {
        sigp_batch_destroy(bt);
        if (out_mean && rules->n_rules > 0) return fail(SIGP_INVALID_ARG, "mean over T = 0 scenarios");
        return SIGP_OK;
    }
    if (!sampled) { sigp_batch_destroy(bt); return fail(SIGP_INVALID_ARG, "sampled is NULL"); }
    if (n_rows < bt->rows_needed) {
        sigp_batch_destroy(bt);
        return fail(SIGP_INVALID_ARG, "sampled has too few rows");
    }
    st = run_host(bt, sampled, (size_t)n_rows * inst->n * 8, n_rows, out_pairs, out_mean);
    sigp_batch_destroy(bt);
    return st;
}

extern "C" sigp_status sigp_evaluate_multi(sigp_instance *const *insts, int32_t n_inst,
                                           const sigp_rules *rules, const double *sampled, int32_t T,
                                           const int64_t *masks, int32_t n_masks, int32_t mask_group,
                                           double *out_pairs, double *out_mean) {
    sigp_batch *bt = nullptr;
    sigp_status st = sigp_batch_create_multi(insts, n_inst, rules, T, masks, n_masks, mask_group, &bt);
    if (st != SIGP_OK) return st;
    if (bt->P == 0 || T == 0) {
        sigp_batch_destroy(bt);
        if (out_mean && rules->n_rules > 0) return fail(SIGP_INVALID_ARG, "mean over T = 0 scenarios");
        return SIGP_OK;
    }
    if (!sampled) { sigp_batch_destroy(bt); return fail(SIGP_INVALID_ARG, "sampled is NULL"); }
    size_t doubles = 0;
    for (int k = 0; k < n_inst; ++k) doubles += (size_t)T * insts[k]->n;
    st = run_host(bt, sampled, doubles * 8, (int64_t)n_inst * T, out_pairs, out_mean);
    sigp_batch_destroy(bt);
    return st;
}

// Phenotype launch on `stream`: one workgroup per rule, as many wavefronts
// as LDS allows beside the kernel's 1024 static priorities.
// ---- allocation-free evaluation --------------------------------------------
// sigp_evaluate_noalloc keeps the block sigp_batch_create would hipMalloc --
// compiled rules, read sets, fused-computation flags, the instance slot and
// the launch order -- in this file-scope array instead: one copy per GPU,
// held until the process ends.  Host code never names the array when it
// needs its memory: the device address comes from hipGetSymbolAddress, once
// per device, with that device current.
static __device__ __align__(256) char scratch[SIGP_NOALLOC_MAX_BYTES];

static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static char *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

static sigp_status scratch_on_current_device(char **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(SIGP_INVALID_ARG, "device index " + std::to_string(dev) + " beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SIGP_DEVICE_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (char *)p;
    }
    *out = g_scratch_ptr[dev];
    return SIGP_OK;
}

extern "C" sigp_status sigp_evaluate_noalloc(sigp_instance *inst, const sigp_rules *rules,
                                             const double *sampled_dev, int64_t n_rows, int32_t T,
                                             int32_t group_size, const int64_t *masks, int32_t n_masks,
                                             int32_t mask_group, double *out_pairs_dev, double *out_mean_dev,
                                             void *stream) {
    if (!inst || !rules) return fail(SIGP_INVALID_ARG, "inst and rules must be non-NULL");
    if (T < 0) return fail(SIGP_INVALID_ARG, "T < 0");
    if (group_size < 1) return fail(SIGP_INVALID_ARG, "group_size must be >= 1");
    PreparedRules pr;
    sigp_status st = prepare_rules(rules, masks, n_masks, mask_group, pr);
    if (st != SIGP_OK) return st;
    if (pr.P == 0 || T == 0) {
        if (out_mean_dev && pr.P > 0) return fail(SIGP_INVALID_ARG, "mean over T = 0 scenarios");
        return SIGP_OK;
    }
    if (!sampled_dev) return fail(SIGP_INVALID_ARG, "sampled_dev is NULL");
    if (!out_pairs_dev) return fail(SIGP_INVALID_ARG, "out_pairs_dev is NULL (the rewards are the caller's buffer)");
    sigp_batch bt;   // host-side description; its device block is the scratch array
    BatchPlan pl;
    st = plan_batch({inst}, false, pr, T, group_size, &bt, pl);
    if (st != SIGP_OK) return st;
    if (pl.bytes > SIGP_NOALLOC_MAX_BYTES)
        return fail(SIGP_INVALID_ARG, "these rules need " + std::to_string(pl.bytes) +
                                          " bytes of device scratch; sigp_evaluate_noalloc holds at most " +
                                          std::to_string((long long)SIGP_NOALLOC_MAX_BYTES) +
                                          " (SIGP_NOALLOC_MAX_BYTES): use sigp_evaluate or fewer rules");
    if (n_rows < bt.rows_needed)
        return fail(SIGP_INVALID_ARG, "sampled has " + std::to_string(n_rows) + " rows, rules need " +
                                          std::to_string(bt.rows_needed));
    char *ptr = nullptr;
    st = scratch_on_current_device(&ptr);
    if (st != SIGP_OK) return st;
    const std::vector<char> img = batch_image(&bt, pl, pr);
    hipStream_t s = (hipStream_t)stream;
    // every byte the kernels read is rewritten here, so nothing left by the
    // previous call matters
    HIP_TRY(hipMemcpyAsync(ptr, img.data(), pl.bytes, hipMemcpyHostToDevice, s));
    bind_batch(&bt, ptr, pl, pr);
    st = launch_unit(&bt, bt.units[0], sampled_dev, out_pairs_dev, s);
    if (st != SIGP_OK) return st;
    if (out_mean_dev) {
        mean_kernel<<<(pr.P + 255) / 256, 256, 0, s>>>(out_pairs_dev, pr.P, T, out_mean_dev);
        HIP_TRY(hipGetLastError());
    }
    // synchronous: the host image must outlive the copy, and the scratch
    // array is free for the next call once this one returns
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(SIGP_DEVICE_ERROR, std::string("kernel failed: ") + hipGetErrorString(e));
    return SIGP_OK;
}

static void launch_pheno(const sigp_batch *bt, const PhenoSpec &Ph, int32_t *outd, hipStream_t stream) {
    int w = 4;
    while (w > 1 && 1024 * sizeof(double) + (size_t)w * lds_per_wave(bt) > LDS_LIMIT) w >>= 1;
    const unsigned lds = (unsigned)(w * lds_per_wave(bt));
    const int P = bt->P;
    const int nwb = nw_bucket(bt->insts[0]->n);
#define PHENO(NWV)                                                                                   \
    do {                                                                                             \
        if (bt->has_rmod) {                                                                          \
            if (bt->has_rare) pheno_kernel<NWV, true, true><<<P, 64 * w, lds, stream>>>(bt->slots, bt->rs, Ph, outd); \
            else pheno_kernel<NWV, true, false><<<P, 64 * w, lds, stream>>>(bt->slots, bt->rs, Ph, outd); \
        } else {                                                                                     \
            if (bt->has_rare) pheno_kernel<NWV, false, true><<<P, 64 * w, lds, stream>>>(bt->slots, bt->rs, Ph, outd); \
            else pheno_kernel<NWV, false, false><<<P, 64 * w, lds, stream>>>(bt->slots, bt->rs, Ph, outd); \
        }                                                                                            \
    } while (0)
    if (nwb == 4) PHENO(4);
    else if (nwb == 9) PHENO(9);
    else PHENO(16);
#undef PHENO
}

extern "C" sigp_status sigp_phenotype(sigp_instance *inst, const sigp_rules *rules, const int64_t *masks,
                                      int32_t n_masks, int32_t mask_group, const sigp_situations *sit,
                                      int32_t *out_ranks) {
    if (!inst || !rules || !sit || !out_ranks) return fail(SIGP_INVALID_ARG, "NULL argument");
    const int S = sit->n_sit, C = sit->n_cand, n = inst->n;
    if (S < 1 || C < 1 || (long long)S * C > 1024)
        return fail(SIGP_INVALID_ARG, "need 1 <= n_sit * n_cand <= 1024");
    if (!sit->cur || !sit->ct || !sit->remaining || !sit->visited || !sit->cand)
        return fail(SIGP_INVALID_ARG, "situation arrays must be non-NULL");
    std::vector<uint32_t> words((size_t)S * 16, 0u);
    for (int k = 0; k < S; ++k) {
        if (sit->cur[k] < 0 || sit->cur[k] > n) return fail(SIGP_INVALID_ARG, "situation cur out of range");
        if (sit->remaining[k] < 0) return fail(SIGP_INVALID_ARG, "situation remaining < 0");
        for (int i = 0; i < n; ++i)
            if (sit->visited[(size_t)k * n + i]) words[(size_t)k * 16 + (i >> 5)] |= 1u << (i & 31);
        for (int j = 0; j < C; ++j) {
            const int c = sit->cand[(size_t)k * C + j];
            if (c < 0 || c >= n) return fail(SIGP_INVALID_ARG, "candidate out of range");
        }
    }
    PreparedRules pr;
    sigp_status st = prepare_rules(rules, masks, n_masks, mask_group, pr);
    if (st != SIGP_OK) return st;
    if (pr.P == 0) return SIGP_OK;
    sigp_batch *bt = nullptr;
    if ((st = make_batch({inst}, false, pr, 1, std::max(pr.P, 1), &bt)) != SIGP_OK) return st;
    // situations + results in one allocation
    const size_t o_cur = 0, o_rem = align256(S * 4 + 8), o_cand = o_rem + align256(S * 4 + 8),
                 o_ct = o_cand + align256((size_t)S * C * 4 + 8), o_vis = o_ct + align256(S * 8 + 8),
                 o_out = o_vis + align256(words.size() * 4 + 8),
                 total = o_out + align256((size_t)pr.P * S * C * 4 + 8);
    char *d = nullptr;
    if (hipMalloc((void **)&d, total) != hipSuccess) {
        sigp_batch_destroy(bt);
        return fail(SIGP_OOM, "phenotype allocation failed");
    }
    hipError_t e = hipMemcpy(d + o_cur, sit->cur, S * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + o_rem, sit->remaining, S * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + o_cand, sit->cand, (size_t)S * C * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + o_ct, sit->ct, S * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + o_vis, words.data(), words.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        PhenoSpec Ph;
        Ph.n_sit = S;
        Ph.n_cand = C;
        Ph.cur = (const int *)(d + o_cur);
        Ph.rem = (const int *)(d + o_rem);
        Ph.cand = (const int *)(d + o_cand);
        Ph.ct = (const double *)(d + o_ct);
        Ph.vis = (const uint32_t *)(d + o_vis);
        int32_t *outd = (int32_t *)(d + o_out);
        launch_pheno(bt, Ph, outd, nullptr);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(out_ranks, outd, (size_t)pr.P * S * C * 4, hipMemcpyDeviceToHost);
    }
    (void)hipFree(d);
    sigp_batch_destroy(bt);
    if (e != hipSuccess) return fail(SIGP_DEVICE_ERROR, std::string("phenotype failed: ") + hipGetErrorString(e));
    return SIGP_OK;
}

extern "C" sigp_status sigp_device_math(int32_t which, const double *x, const double *y, int64_t count,
                                        double *out) {
    if (count < 0 || (count > 0 && (!x || !out)) || (which == 4 && count > 0 && !y))
        return fail(SIGP_INVALID_ARG, "bad arguments");
    if (count == 0) return SIGP_OK;
    sigp_status st = check_device(nullptr);
    if (st != SIGP_OK) return st;
    double *dx = nullptr, *dy = nullptr, *dout = nullptr;
    const size_t bytes = (size_t)count * 8;
    HIP_TRY(hipMalloc((void **)&dx, bytes));
    HIP_TRY(hipMalloc((void **)&dy, bytes));
    HIP_TRY(hipMalloc((void **)&dout, bytes));
    HIP_TRY(hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice));
    if (y) HIP_TRY(hipMemcpy(dy, y, bytes, hipMemcpyHostToDevice));
    else HIP_TRY(hipMemset(dy, 0, bytes));
    math_kernel<<<(unsigned)((count + 255) / 256), 256>>>(which, dx, dy, count, dout);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost));
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipFree(dout);
    return SIGP_OK;
}

// ---------------------------------------------------------------------------
// interfaces for the library's other translation units

int sigp_internal_instance_n(const sigp_instance *inst) { return inst->n; }
const double *sigp_internal_instance_dur(const sigp_instance *inst) {
    return (const double *)((uintptr_t)inst->dev.dur.base + inst->dev.dur.off);
}

sigp_status sigp_internal_batch_device(sigp_instance *inst, const sigp_device_rules &r, int T, int group_size,
                                       const sigp_launch_config &cfg, sigp_batch **out) {
    if (!inst || !out) return fail(SIGP_INVALID_ARG, "NULL argument");
    *out = nullptr;
    sigp_status st = check_config(&cfg);
    if (st != SIGP_OK) return st;
    if (r.P < 1 || T < 1 || group_size < 1) return fail(SIGP_INVALID_ARG, "empty device batch");
    if (r.KL < 1 || r.KL > SIGP_N_TERMS || r.SL < 1 || r.SL > SIGP_MAX_STACK)
        return fail(SIGP_INVALID_ARG, "value-store / stack bounds out of range");
    sigp_batch *bt = new sigp_batch();
    bt->insts = {inst};
    bt->P = r.P;
    bt->T = T;
    bt->group_size = group_size;
    bt->waves_per_item = cfg.waves_per_item;
    bt->mapping = cfg.mapping;
    bt->rows_needed = (int64_t)((r.P - 1) / group_size + 1) * T;
    bt->n_out = (long long)r.P * T;
    bt->has_rmod = r.has_rmod;
    bt->has_rare = r.has_rare;
    InstSlot slot;
    slot.dev = inst->dev;
    slot.tabs = inst->tabs;
    slot.samp_off = 0;
    slot.out_off = 0;
    slot.perm_off = 0;
    bt->rep = r.rep;
    if (hipMalloc(&bt->dbuf, sizeof(InstSlot)) != hipSuccess) {
        delete bt;
        return fail(SIGP_OOM, "batch allocation failed");
    }
    if (hipMemcpy(bt->dbuf, &slot, sizeof(InstSlot), hipMemcpyHostToDevice) != hipSuccess) {
        sigp_batch_destroy(bt);
        return fail(SIGP_DEVICE_ERROR, "batch upload failed");
    }
    bt->slots = (cslot)bt->dbuf;
    bt->rs.ops = (cip)r.ops;
    bt->rs.args = (cip)r.args;
    bt->rs.off = (cip)r.off;
    bt->rs.len = (cip)r.len;
    bt->rs.rd = (cup)r.rd;
    bt->rs.flags = (cup)r.flags;
    bt->rs.order = nullptr;
    bt->rs.P = r.P;
    bt->rs.KL = r.KL;
    bt->rs.SL = r.SL;
    LaunchUnit &u = bt->units[0];
    u.nwb = nw_bucket(inst->n);
    u.max_n = inst->n;
    u.n_pairs = r.P;
    u.order = (cip)r.order;
    u.tune_order = (cip)r.order;
    u.tune_pairs = r.P;
    bt->n_units = 1;
    *out = bt;
    return SIGP_OK;
}

sigp_status sigp_internal_phenotype_device(const sigp_batch *bt, const sigp_device_situations &sit,
                                           int32_t *out_dev, hipStream_t stream) {
    if ((long long)sit.n_sit * sit.n_cand > 1024 || sit.n_sit < 1 || sit.n_cand < 1)
        return fail(SIGP_INVALID_ARG, "need 1 <= n_sit * n_cand <= 1024");
    PhenoSpec Ph;
    Ph.n_sit = sit.n_sit;
    Ph.n_cand = sit.n_cand;
    Ph.cur = sit.cur;
    Ph.rem = sit.rem;
    Ph.cand = sit.cand;
    Ph.ct = sit.ct;
    Ph.vis = sit.vis;
    launch_pheno(bt, Ph, out_dev, stream);
    HIP_TRY(hipGetLastError());
    return SIGP_OK;
}
