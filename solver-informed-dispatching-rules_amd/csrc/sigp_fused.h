// Candidate scans: the terminal families of one candidate customer c,
// evaluated for all lanes of a wavefront (lane = scenario).
//
// Feasibility thresholds.  For a POI a and customer j the reference's test
//     arr = f + d(a,j) <= close_j  and  max(arr, open_j) + dur_j + d2d_j <= tmax
// is a monotone predicate of the departure time f (IEEE addition and max are
// monotone), so it holds exactly for f <= tau(a,j), the largest double that
// satisfies it (found per instance by bisection over the ordered doubles,
// evaluating the identical expression).  Variants: tau0 (plain), tau1 (close
// time only), tau2 (robust duration dur_j * (1 + sigma)), and tau0n / tau2n,
// which also fold in the NS family's "skip d(a,j) < 1e-10" rule.  A terminal's
// neighbour test then costs one compare; every value it adds or compares is
// still the reference's own expression (engine/sim_terms.hpp).
//
// Each family is a short loop specialised at compile time for the members
// the rule needs, reading eight neighbours per scalar-load chunk (four in the
// structural and NSTW scans and in the robust scan with REACH-K-ROB's table,
// whose table streams do not fit the scalar registers at eight) from rows
// padded to a multiple of 8 (padding has tau = -inf, so it never counts).
// Scans that start from a lane's own POI (second legs, later REACH-K
// waypoints) gather down the columns of the transposed tables, whose rows are
// padded the same way.
#pragma once
#include "sigp_terms.h"
#include "sigp_rules.h"

namespace sigp {

struct FusedTabs {
    int ldp;   // row length of [a][j] tables (>= n, multiple of 8); rows of [j][a] tables
    int ldq;   // row length of [j][a] tables (>= n+1, multiple of 8)
    cdp t0, t0n, t2, t2n, r, d;          // rows [a][j], a in [0, n]
    cdp t0c, t0nc, t1c, t2c, t2nc, rc;   // columns [j][a], j in [0, ldp)
    cdp open, close, dur, d2d, score, durrob, sdv;   // customer arrays, padded to ldp
};

// A kernel's copy of an instance's tables.  The customer arrays are the
// DevInst's own (the same allocation).  SHARED takes those pointers from
// `dev` rather than loading the copies in `tabs`, which leaves one scalar
// register pair per array instead of two.  The kernels sit at the 102-SGPR
// limit, and this moves their spills: measured on MI355X, the scenario-lane
// kernel at n <= 128 ran 5 % faster with it (plain workload 67.7 -> 64.2 ms),
// the one at n <= 288 3 % slower (1,846 -> 1,888 ms on 512 rules x 512
// scenarios of the large workload) and the candidate-lane kernels no faster,
// so only the n <= 128 scenario-lane kernel takes it.
template <bool SHARED>
__device__ __forceinline__ FusedTabs tabs_of(const FusedTabs &tabs, const DevInst &dev) {
    FusedTabs F = tabs;
    const cbase b = tabs.t0.base;   // the threshold tables' allocation
    F.t0.base = b; F.t0n.base = b; F.t2.base = b; F.t2n.base = b; F.r.base = b; F.d.base = b;
    F.t0c.base = b; F.t0nc.base = b; F.t1c.base = b; F.t2c.base = b; F.t2nc.base = b; F.rc.base = b;
    if constexpr (SHARED) {
        F.open = dev.open;
        F.close = dev.close;
        F.dur = dev.dur;
        F.d2d = dev.d2d;
        F.score = dev.score;
        F.durrob = dev.durrob;
        F.sdv = dev.sdv;
    } else {
        const cbase c = dev.dist.base;   // the customer arrays live in the instance's allocation
        F.open.base = c; F.close.base = c; F.dur.base = c; F.d2d.base = c; F.score.base = c;
        F.durrob.base = c; F.sdv.base = c;
    }
    return F;
}

__device__ __forceinline__ double sel(bool c, double a, double b) { return c ? a : b; }

template <int CH> struct DN { double v[CH]; };
template <int CH>
__device__ __forceinline__ DN<CH> ldn(cdp row, int j0) {
    // one address for the chunk, so that its reads merge into one scalar
    // load (offsets in 32 bits cannot be shown not to wrap)
    const __attribute__((address_space(4))) double *p =
        (const __attribute__((address_space(4))) double *)(row.base + (row.off + 8u * (uint32_t)j0));
    DN<CH> r;
#pragma unroll
    for (int k = 0; k < CH; ++k) r.v[k] = p[k];
    return r;
}
typedef DN<8> D8;
typedef DN<4> D4;
__device__ __forceinline__ D8 ld8(cdp row, int j0) { return ldn<8>(row, j0); }
__device__ __forceinline__ D4 ld4(cdp row, int j0) { return ldn<4>(row, j0); }

// True when every active lane holds the same value of v.
__device__ __forceinline__ bool wave_uniform(int v) {
    const int v0 = __builtin_amdgcn_readfirstlane(v);
    return __builtin_amdgcn_ballot_w64(v != v0) == 0;
}

// A per-lane bit set read in ascending chunks of CH bits from bit 0 (CH
// divides 32): the word holding the chunk is selected once per 32 customers
// rather than once per chunk.
template <int NW, int CH>
struct BitRun {
    const uint32_t (&a)[NW];
    uint32_t w;
    __device__ __forceinline__ explicit BitRun(const uint32_t (&s)[NW]) : a(s), w(0u) {}
    __device__ __forceinline__ uint32_t next(int j0) {
        if ((j0 & 31) == 0) w = wsel<NW>(a, j0 >> 5);
        return (w >> (j0 & 31)) & ((1u << CH) - 1u);
    }
};

// Candidate c's bit in the CH-bit chunk at j0 (0 outside it).  The scans add
// it to the visited bits and so skip c itself: one OR per chunk takes the
// place of a compare per neighbour.
template <int CH>
__device__ __forceinline__ uint32_t cbit(int c, int j0) {
    const uint32_t d = (uint32_t)(c - j0);
    return d < (uint32_t)CH ? 1u << d : 0u;
}

// NS (engine/sim_terms.hpp:18-34), MAXN (:76-93) and the first leg of
// REACH2 (:101-111) from candidate c at time fin, one pass.
template <int NW, bool NS, bool MX, bool R2>
__device__ __forceinline__ void scan_nsfam(const FusedTabs &F, int n, int c, double fin, const uint32_t (&vis)[NW],
                           double &ns, double &mx, double &bv, int &bj) {
    const cdp tr = F.t0n + (size_t)(c + 1) * F.ldp, rr = F.r + (size_t)(c + 1) * F.ldp;
    double s = 0.0, m = 0.0, b = 0.0;
    int bi = -1;
    BitRun<NW, 8> vr(vis);
    for (int j0 = 0; j0 < n; j0 += 8) {
        const uint32_t vb = vr.next(j0) | cbit<8>(c, j0);
        const D8 T = ld8(tr, j0), R = ld8(rr, j0);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int j = j0 + k;
            const bool ok = !((vb >> k) & 1u) && fin <= T.v[k];
            const double r = R.v[k];
            if (NS) s = sel(ok, s + r, s);
            if (MX) m = sel(ok && r > m, r, m);
            if (R2) {
                const bool take = ok && r > b;
                b = sel(take, r, b);
                bi = take ? j : bi;
            }
        }
    }
    ns = s; mx = m; bv = b; bj = bi;
}

// NSROB (engine/sim_terms.hpp:56-74), the first leg of REACH2ROB (:164-174)
// and the first step of REACH-K-ROB (:455-468).
template <int NW, bool NSR, bool R2R, bool RKR>
__device__ __forceinline__ void scan_robfam(const FusedTabs &F, int n, int c, double frob, double trob,
                            const uint32_t (&vis)[NW], double &nsr, int &bjr, int &bkr) {
    const cdp tn = F.t2n + (size_t)(c + 1) * F.ldp, tp = F.t2 + (size_t)(c + 1) * F.ldp,
              rr = F.r + (size_t)(c + 1) * F.ldp;
    double s = 0.0, b = 0.0, bk = -1.0;
    int bi = -1, ki = -1;
    // With REACH-K-ROB's table the scan reads three tables per neighbour; four
    // neighbours per chunk then keep their scalar loads from spilling.
    constexpr int CH = RKR ? 4 : 8;
    BitRun<NW, CH> vr(vis);
    for (int j0 = 0; j0 < n; j0 += CH) {
        const uint32_t vb = vr.next(j0) | cbit<CH>(c, j0);
        const DN<CH> T = ldn<CH>(tn, j0), R = ldn<CH>(rr, j0);
        DN<CH> TP;
        if (RKR) TP = ldn<CH>(tp, j0);
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int j = j0 + k;
            const bool open = !((vb >> k) & 1u);
            const bool ok = open && frob <= T.v[k];
            const double r = R.v[k];
            if (NSR) s = sel(ok, s + r, s);
            if (R2R) {
                const bool take = ok && r > b;
                b = sel(take, r, b);
                bi = take ? j : bi;
            }
            if (RKR) {
                const bool take = open && trob <= TP.v[k] && r > bk;
                bk = sel(take, r, bk);
                ki = take ? j : ki;
            }
        }
    }
    nsr = s; bjr = bi; bkr = ki;
}

// REGRET (engine/sim_terms.hpp:36-54): neighbours feasible now (fnow) that
// become infeasible after serving c.
template <int NW>
__device__ __forceinline__ double scan_regret(const FusedTabs &F, int n, int c, double fin,
                              const uint32_t (&vis)[NW], const uint32_t (&fnow)[NW]) {
    const cdp tr = F.t0 + (size_t)(c + 1) * F.ldp;
    double s = 0.0;
    BitRun<NW, 8> fr(fnow), vr(vis);
    for (int j0 = 0; j0 < n; j0 += 8) {
        const uint32_t mb = fr.next(j0) & ~(vr.next(j0) | cbit<8>(c, j0));
        const D8 T = ld8(tr, j0), S = ld8(F.score, j0);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const bool add = ((mb >> k) & 1u) && !(fin <= T.v[k]);
            s = sel(add, s + S.v[k], s);
        }
    }
    return s;
}

// FRAGCNT, FRAGSCORE, DOWNSLACK and PBUST's best neighbour
// (engine/sim_terms.hpp:544-562), after serving c at fin_c.
template <int NW>
__device__ __forceinline__ void scan_struct(const FusedTabs &F, int n, int c, double fin_c, const uint32_t (&vis)[NW],
                            const uint32_t (&fnow)[NW], double &fcnt, double &fsc, double &dsl,
                            int &bq) {
    const size_t row = (size_t)(c + 1) * F.ldp;
    const cdp tr = F.t0 + row, rr = F.r + row, dd = F.d + row;
    double cnt = 0.0, sc = 0.0, ds = 1e18, bv = -1.0;
    int bi = -1;
    // Four neighbours per chunk: five streams of eight would hold 80 scalar
    // registers and spill inside the loop.
    BitRun<NW, 4> vr(vis), fr(fnow);
    for (int j0 = 0; j0 < n; j0 += 4) {
        const uint32_t vb = vr.next(j0) | cbit<4>(c, j0);
        const uint32_t fb = fr.next(j0);
        const D4 T = ld4(tr, j0), R = ld4(rr, j0), D = ld4(dd, j0), S = ld4(F.score, j0),
                 C = ld4(F.close, j0);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int q = j0 + k;
            const bool open = !((vb >> k) & 1u);
            const bool after = fin_c <= T.v[k];
            const bool frag = open && ((fb >> k) & 1u) && !after;
            cnt = sel(frag, cnt + 1.0, cnt);
            sc = sel(frag, sc + S.v[k], sc);
            const bool aft = open && after;
            const double slack = C.v[k] - (fin_c + D.v[k]);
            ds = sel(aft && slack < ds, slack, ds);
            const bool take = aft && D.v[k] > 1e-10 && R.v[k] > bv;
            bv = sel(take, R.v[k], bv);
            bi = take ? q : bi;
        }
    }
    fcnt = cnt; fsc = sc; dsl = ds; bq = bi;
}

// First step of REACH-K (engine/sim_terms.hpp:417-431) from c at t = fin_c.
template <int NW>
__device__ __forceinline__ int scan_reach1(const FusedTabs &F, int n, int c, double t, const uint32_t (&vis)[NW]) {
    const cdp tr = F.t0 + (size_t)(c + 1) * F.ldp, rr = F.r + (size_t)(c + 1) * F.ldp;
    double bv = -1.0;
    int bi = -1;
    BitRun<NW, 8> vr(vis);
    for (int j0 = 0; j0 < n; j0 += 8) {
        const uint32_t vb = vr.next(j0) | cbit<8>(c, j0);
        const D8 T = ld8(tr, j0), R = ld8(rr, j0);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int q = j0 + k;
            const bool take = !((vb >> k) & 1u) && t <= T.v[k] && R.v[k] > bv;
            bv = sel(take, R.v[k], bv);
            bi = take ? q : bi;
        }
    }
    return bi;
}

// NSTW (engine/sim_terms.hpp:192-213).  The weight is exactly 1 unless the
// lane arrives before the window opens (with no wait, 1 - 0 / win == 1), so
// the division runs only for lanes that wait, and r * 1.0 == r is added for
// the rest.  Four neighbours per chunk: the five tables a chunk with an
// eligible neighbour reads fit the scalar registers.
template <int NW>
__device__ __forceinline__ double scan_nstw(const FusedTabs &F, int n, int c, double fin, const uint32_t (&vis)[NW]) {
    const size_t row = (size_t)(c + 1) * F.ldp;
    const cdp tr = F.t0n + row, rr = F.r + row, dd = F.d + row;
    double s = 0.0;
    BitRun<NW, 4> vr(vis);
    for (int j0 = 0; j0 < n; j0 += 4) {
        const uint32_t vb = vr.next(j0) | cbit<4>(c, j0);
        const D4 T = ld4(tr, j0);
        uint32_t okm = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool ok = !((vb >> k) & 1u) && fin <= T.v[k];
            okm |= (ok ? 1u : 0u) << k;
        }
        if (__builtin_amdgcn_ballot_w64(okm != 0) == 0) continue;
        const D4 R = ld4(rr, j0), O = ld4(F.open, j0), C = ld4(F.close, j0), D = ld4(dd, j0);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!((okm >> k) & 1u)) continue;
            double wait = O.v[k] - (fin + D.v[k]);
            if (wait < 0.0) wait = 0.0;
            const double win = C.v[k] - O.v[k];
            double w = 1.0;
            if (wait > 0.0 && win > 1e-10) w = 1.0 - wait / win;
            if (w < 0.0) w = 0.0;
            s += R.v[k] * w;
        }
    }
    return s;
}

// ATRISK (engine/sim_terms.hpp:225-257) for candidate c.  pow(var, 0.5)
// comes from the per-instance duration-class table.  One neighbour at a
// time: the body holds two erf evaluations.  A lane whose difference lies
// beyond its divisor's saturation threshold (DevInst::atp*, atq*) takes the
// value norm_cdf would return there, exactly 1 or 0, without the division
// and the erf; the erf path runs only for lanes that need it.  A lane whose
// mean arrival at j is not after open_j starts j at open_j with no
// variance, so its p_bud is j's own constant (DevInst::atpb0).
template <int NW>
__device__ __forceinline__ double scan_eloss(const DevInst &I, const FusedTabs &F, int c, double fin_c,
                             const uint32_t (&vis)[NW], const uint32_t (&cnow)[NW]) {
    const double sdc = F.sdv[c];
    const double sdc2 = sdc * sdc;
    const int kc = I.dcls[c];
    const cdp pwrow = I.pw2 + (size_t)kc * I.K;
    const cdp qrow = I.atq2 + (size_t)kc * I.K * 2;
    const double p_hi = I.atp1[kc], p_lo = I.atp0[kc];
    const cdp dd = F.d + (size_t)(c + 1) * F.ldp;
    double el = 0.0;
    uint32_t mb = 0;
#pragma unroll 1
    for (int j = 0; j < I.n; ++j) {
        if ((j & 31) == 0) {
            const uint32_t d = (uint32_t)(c - j);   // c itself is skipped
            mb = wsel<NW>(cnow, j >> 5) & ~(wsel<NW>(vis, j >> 5) | (d < 32u ? 1u << d : 0u));
        }
        const bool act = (mb >> (j & 31)) & 1u;
        if (__builtin_amdgcn_ballot_w64(act) == 0) continue;
        if (!act) continue;
        const double ma = fin_c + dd[j];
        const double cl = F.close[j];
        // Each factor is first classified without erf: saturated (exactly 0
        // or 1: DevInst::atp*, atq*, atpb0) or needing it.  1 - pw * pb is 1
        // whenever either factor is 0 (both lie in [0, 1]), so a factor's
        // erf is skipped when the other is exactly 0.
        double pw, dcs = 0.0;
        bool pw_e = false;
        if (sdc > 1e-9) {
            dcs = cl - ma;
            pw = dcs >= p_hi ? 1.0 : 0.0;
            pw_e = dcs < p_hi && dcs > p_lo;
        } else {
            pw = ma <= cl ? 1.0 : 0.0;
        }
        double pb = 1.0, es = 0.0, pdiv = 1.0;
        bool pb_e = false;
        if (pw_e || pw != 0.0) {
            const double op = F.open[j];
            if (ma > op) {   // late: the start time inherits the candidate's variance
                const double sdj = F.sdv[j];
                const double vb = sdc2 + sdj * sdj;
                const double mbd = ma + F.dur[j] + F.d2d[j];
                if (vb > 1e-18) {
                    const int kj = I.dcls[j];
                    es = I.tmax - mbd;
                    const double q_hi = qrow[2 * kj], q_lo = qrow[2 * kj + 1];
                    pb = es >= q_hi ? 1.0 : 0.0;
                    pb_e = es < q_hi && es > q_lo;
                    pdiv = pwrow[kj];
                } else {
                    pb = mbd <= I.tmax ? 1.0 : 0.0;
                }
            } else {
                pb = I.atpb0[j];   // starts at open_j: the per-customer value
            }
        }
        if (pw_e && (pb_e || pb != 0.0)) pw = norm_cdf(dcs / sdc);
        if (pb_e && pw != 0.0) pb = norm_cdf(es / pdiv);
        el += F.score[j] * (1.0 - pw * pb);
    }
    return el;
}

// Customers per chunk where each lane gathers from its own POI's column
// (REACH2's second leg, later REACH-K waypoints): eight loads in flight per
// table at n > 128, four at n <= 128 (measured on MI355X: eight made the
// n <= 288 kernel 1.7 % faster and the n <= 128 one 0.5 % slower).
template <int NW> struct GatherWidth { static constexpr int value = NW > 4 ? 8 : 4; };

// Second leg of REACH2 / REACH2ROB (engine/sim_terms.hpp:113-126 and
// :176-189) from each lane's own first neighbour bj at time fj.
template <int NW>
__device__ __forceinline__ double reach2_leg2(const DevInst &I, const FusedTabs &F, int i, int bj, double fj,
                              const uint32_t (&vis)[NW], bool rob) {
    const int pj = bj + 1;
    double bv2 = 0.0, s2 = 0.0;
    constexpr int GW = GatherWidth<NW>::value;
    if (wave_uniform(pj)) {
        // every lane continues from the same neighbour: its row, read as scalars
        const int pu = __builtin_amdgcn_readfirstlane(pj), bu = pu - 1;
        const cdp tr = (rob ? F.t2n : F.t0n) + (size_t)pu * F.ldp, rr = F.r + (size_t)pu * F.ldp;
        BitRun<NW, 8> vr(vis);
        for (int k0 = 0; k0 < I.n; k0 += 8) {
            const uint32_t vb = vr.next(k0);
            const D8 T = ld8(tr, k0), R = ld8(rr, k0), S = ld8(F.score, k0);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = k0 + u;
                const bool take = k != i && k != bu && !((vb >> u) & 1u) && fj <= T.v[u] &&
                                  R.v[u] > bv2;
                bv2 = sel(take, R.v[u], bv2);
                s2 = sel(take, S.v[u], s2);
            }
        }
        return I.score[bu] + s2;
    }
    const cdp tcol = rob ? F.t2nc : F.t0nc;
    BitRun<NW, GW> vr(vis);
    for (int k0 = 0; k0 < I.n; k0 += GW) {
        const uint32_t vb = vr.next(k0);
        double T[GW], R[GW];
#pragma unroll
        for (int u = 0; u < GW; ++u) {
            const size_t at = (size_t)(k0 + u) * F.ldq + pj;
            T[u] = tcol[at];
            R[u] = F.rc[at];
        }
#pragma unroll
        for (int u = 0; u < GW; ++u) {
            const int k = k0 + u;
            const bool take = k != i && k != bj && !((vb >> u) & 1u) && fj <= T[u] && R[u] > bv2;
            bv2 = sel(take, R[u], bv2);
            s2 = sel(take, F.score[k], s2);
        }
    }
    return I.score[bj] + s2;
}

// The REACH-K greedy path (engine/sim_terms.hpp:405-441; robust :443-479)
// given its first waypoint; later waypoints scan the lane's own row.
// REACH1/3/5 are prefixes of one path: returns the totals after
// min(K, picks) for K = 1, 3, 5.
template <int NW>
__device__ __forceinline__ void reach_path(const DevInst &I, const FusedTabs &F, int c, int best1, double t0,
                           const uint32_t (&vis)[NW], int K, bool rob, double *tot1,
                           double *tot3, double *tot5) {
    double total = 0.0;
    *tot1 = *tot3 = *tot5 = 0.0;
    if (best1 < 0) return;
    constexpr int GW = GatherWidth<NW>::value;
    const cdp tcol = rob ? F.t2c : F.t0c;
    uint32_t scr[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) scr[w] = vis[w];
    wset<NW>(scr, c);
    int cur = c + 1;
    double t = t0;
    int best = best1;
    for (int step = 0; step < K; ++step) {
        if (step > 0 && wave_uniform(cur)) {
            // every lane stands at the same waypoint: its row, read as scalars
            const int cu = __builtin_amdgcn_readfirstlane(cur);
            const cdp tr = (rob ? F.t2 : F.t0) + (size_t)cu * F.ldp, rr = F.r + (size_t)cu * F.ldp;
            best = -1;
            double bv = -1.0;
            BitRun<NW, 8> sr(scr);
            for (int q0 = 0; q0 < I.n; q0 += 8) {
                const uint32_t sb = sr.next(q0);
                const D8 T = ld8(tr, q0), R = ld8(rr, q0);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool take = !((sb >> u) & 1u) && t <= T.v[u] && R.v[u] > bv;
                    bv = sel(take, R.v[u], bv);
                    best = take ? q0 + u : best;
                }
            }
            if (best < 0) break;
        } else if (step > 0) {
            best = -1;
            double bv = -1.0;
            BitRun<NW, GW> sr(scr);
            for (int q0 = 0; q0 < I.n; q0 += GW) {
                const uint32_t sb = sr.next(q0);
                double T[GW], R[GW];
#pragma unroll
                for (int u = 0; u < GW; ++u) {
                    const size_t at = (size_t)(q0 + u) * F.ldq + cur;
                    T[u] = tcol[at];
                    R[u] = F.rc[at];
                }
#pragma unroll
                for (int u = 0; u < GW; ++u) {
                    const bool take = !((sb >> u) & 1u) && t <= T[u] && R[u] > bv;
                    bv = sel(take, R[u], bv);
                    best = take ? q0 + u : best;
                }
            }
            if (best < 0) break;
        }
        const double ta = F.d[(size_t)cur * F.ldp + best];
        const double arr = t + ta;
        const double st = arr > F.open[best] ? arr : F.open[best];
        t = rob ? st + F.durrob[best] : st + F.dur[best];
        cur = best + 1;
        total += F.score[best];
        wset<NW>(scr, best);
        // REACH-K stops after K picks or when no waypoint is left, so its
        // value is the running total after min(K, picks) steps
        if (step == 0) *tot1 = total;
        if (step <= 2) *tot3 = total;
        *tot5 = total;
    }
}

// Store terminal slot s of the rule (read set rd) in the wavefront's LDS
// value store: row popcount(rd & (bit(s) - 1)), one double per lane.
__device__ __forceinline__ void put_slot(double *tv, uint64_t rd, int s, double v) {
    if ((rd >> s) & 1ull) tv[__builtin_popcountll(rd & ((1ull << s) - 1ull)) * 64] = v;
}

// Every scan the rule needs for candidate c, each family's slots stored as
// soon as it finishes (engine/sim_core.hpp:99-198 for the slot formulas), so
// that only the four values the derived slots combine (NS, NSROB, REACH2ROB,
// REGRET) stay live across families.  fin = ct + tfv (NS family),
// st_c = max(ct + ta, open_c), fin_c = st_c + dur_c (structural, REACH-K,
// ATRISK); ta = travel time from the current POI.
template <int NW>
__device__ __forceinline__ void cand_terms(const DevInst &I, const FusedTabs &F, uint64_t fl,
                                           uint64_t rd, int kpath, int kpathr, int c, double ta,
                                           double fin, double st_c, double fin_c,
                                           const uint32_t (&vis)[NW], const uint32_t (&fnow)[NW],
                                           const uint32_t (&cnow)[NW], double *tv) {
    const int n = I.n;
    const double dur_c = F.dur[c];
    double ns = 0.0, nsrob = 0.0, r2rob = 0.0, regret = 0.0;
    if (fl & (F_NS | F_MAXN | F_REACH2)) {
        double maxn = 0.0, bv2 = 0.0, reach2 = 0.0;
        int bj2 = -1;
        switch ((int)(fl & (F_NS | F_MAXN | F_REACH2))) {
        case (int)F_NS: scan_nsfam<NW, true, false, false>(F, n, c, fin, vis, ns, maxn, bv2, bj2); break;
        case (int)F_MAXN: scan_nsfam<NW, false, true, false>(F, n, c, fin, vis, ns, maxn, bv2, bj2); break;
        case (int)F_REACH2: scan_nsfam<NW, false, false, true>(F, n, c, fin, vis, ns, maxn, bv2, bj2); break;
        case (int)(F_NS | F_MAXN): scan_nsfam<NW, true, true, false>(F, n, c, fin, vis, ns, maxn, bv2, bj2); break;
        case (int)(F_NS | F_REACH2): scan_nsfam<NW, true, false, true>(F, n, c, fin, vis, ns, maxn, bv2, bj2); break;
        case (int)(F_MAXN | F_REACH2): scan_nsfam<NW, false, true, true>(F, n, c, fin, vis, ns, maxn, bv2, bj2); break;
        default: scan_nsfam<NW, true, true, true>(F, n, c, fin, vis, ns, maxn, bv2, bj2); break;
        }
        put_slot(tv, rd, 10, ns);
        put_slot(tv, rd, 13, maxn);
        if ((fl & F_REACH2) && bj2 >= 0) {   // engine/sim_terms.hpp:110,112-126
            const double arr = fin + F.d[(size_t)(c + 1) * F.ldp + bj2];
            const double st = arr > F.open[bj2] ? arr : F.open[bj2];
            reach2 = reach2_leg2<NW>(I, F, c, bj2, st + F.dur[bj2], vis, false);
        }
        put_slot(tv, rd, 14, reach2);
        put_slot(tv, rd, 32, (reach2 - I.zp[1]) / I.zp[6]);
    }

    if (fl & (F_NSROB | F_R2ROB | F_REACHKR)) {
        const double frob = fin + dur_c * kSigma;   // engine/sim_terms.hpp:61,161
        const double trob = st_c + F.durrob[c];     // engine/sim_terms.hpp:453
        int bjr = -1, bkr = -1;
        switch ((int)(fl & (F_NSROB | F_R2ROB | F_REACHKR))) {
        case (int)F_NSROB: scan_robfam<NW, true, false, false>(F, n, c, frob, trob, vis, nsrob, bjr, bkr); break;
        case (int)F_R2ROB: scan_robfam<NW, false, true, false>(F, n, c, frob, trob, vis, nsrob, bjr, bkr); break;
        case (int)F_REACHKR: scan_robfam<NW, false, false, true>(F, n, c, frob, trob, vis, nsrob, bjr, bkr); break;
        case (int)(F_NSROB | F_R2ROB): scan_robfam<NW, true, true, false>(F, n, c, frob, trob, vis, nsrob, bjr, bkr); break;
        case (int)(F_NSROB | F_REACHKR): scan_robfam<NW, true, false, true>(F, n, c, frob, trob, vis, nsrob, bjr, bkr); break;
        case (int)(F_R2ROB | F_REACHKR): scan_robfam<NW, false, true, true>(F, n, c, frob, trob, vis, nsrob, bjr, bkr); break;
        default: scan_robfam<NW, true, true, true>(F, n, c, frob, trob, vis, nsrob, bjr, bkr); break;
        }
        put_slot(tv, rd, 12, nsrob);
        if ((fl & F_R2ROB) && bjr >= 0) {   // engine/sim_terms.hpp:173,175-189
            const double arr = frob + F.d[(size_t)(c + 1) * F.ldp + bjr];
            const double st = arr > F.open[bjr] ? arr : F.open[bjr];
            r2rob = reach2_leg2<NW>(I, F, c, bjr, st + F.durrob[bjr], vis, true);
        }
        put_slot(tv, rd, 41, r2rob);
        if (fl & F_REACHKR) {
            double t1, rkr3, rkr5;
            reach_path<NW>(I, F, c, bkr, trob, vis, kpathr, true, &t1, &rkr3, &rkr5);
            put_slot(tv, rd, 47, rkr3);
            put_slot(tv, rd, 48, rkr5);
        }
    }
    put_slot(tv, rd, 45, ns - nsrob);
    if (rd & (1ull << 46)) {
        double den = ta + F.d2d[c];
        if (den < 1.0) den = 1.0;
        put_slot(tv, rd, 46, (F.score[c] + nsrob) / den);
    }

    if (fl & F_REGRET) regret = scan_regret<NW>(F, n, c, fin, vis, fnow);
    put_slot(tv, rd, 11, regret);
    put_slot(tv, rd, 33, (regret - I.zp[2]) / I.zp[7]);
    put_slot(tv, rd, 43, r2rob - regret);

    if (fl & F_STRUCT) {
        double fcnt, fscore, dsl, pbust = 0.0;
        int bq = -1;
        scan_struct<NW>(F, n, c, fin_c, vis, fnow, fcnt, fscore, dsl, bq);
        if (bq >= 0) {   // engine/sim_terms.hpp:565-575
            const double ab = fin_c + F.d[(size_t)(c + 1) * F.ldp + bq];
            const double sd = kSigma * dur_c;
            if (sd > 1e-12) {
                const double z = (F.close[bq] - ab) / (sd * kSqrt2);
                pbust = 0.5 * sigp_erfc(z);
            } else {
                pbust = ab > F.close[bq] ? 1.0 : 0.0;
            }
        }
        put_slot(tv, rd, 17, fcnt);
        put_slot(tv, rd, 18, fscore);
        put_slot(tv, rd, 19, dsl > 1e17 ? 0.0 : dsl);
        put_slot(tv, rd, 20, pbust);
        put_slot(tv, rd, 28, fscore / (ns + 1.0));
        put_slot(tv, rd, 29, fscore - ns);
    }

    if (fl & F_REACHK) {
        double rk1, rk3, rk5;
        const int b1 = scan_reach1<NW>(F, n, c, fin_c, vis);
        reach_path<NW>(I, F, c, b1, fin_c, vis, kpath, false, &rk1, &rk3, &rk5);
        put_slot(tv, rd, 21, rk1);
        put_slot(tv, rd, 22, rk3);
        put_slot(tv, rd, 23, rk5);
    }
    if (fl & F_NSTW) put_slot(tv, rd, 15, scan_nstw<NW>(F, n, c, fin, vis));
    if (fl & F_ATRISK) {
        const double eloss = scan_eloss<NW>(I, F, c, fin_c, vis, cnow);
        put_slot(tv, rd, 30, eloss);
        put_slot(tv, rd, 34, (eloss - I.zp[3]) / I.zp[8]);
    }
}

}  // namespace sigp
