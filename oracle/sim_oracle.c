/*
 * CPU ORACLE -- test infrastructure only (see sim_oracle.h for the pins).
 *
 * Restates the reference simulator in C.  Every floating-point expression
 * keeps the reference's operand order and types; build with -O2
 * -ffp-contract=off (the reference's CPU build, Makefile:7, is g++ -O2 on
 * x86-64, which never contracts).  Transcendentals come from the system libm,
 * exactly as in the reference (erf, erfc, pow, sqrt).
 *
 * With -DORC_COUNT every FP64 operation on the simulation path is tallied
 * (add, sub, mul, div, sqrt, compare, min/max select and libm call each count
 * as one) into a per-thread counter: the algorithmic work model used for the
 * roofline.
 */
#include "sim_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORC_COUNT
static __thread unsigned long long orc_ops_ctr;
#define OP(k) (orc_ops_ctr += (unsigned long long)(k))
unsigned long long orc_ops_read(void) { return orc_ops_ctr; }
void orc_ops_reset(void) { orc_ops_ctr = 0; }
#else
#define OP(k) ((void)0)
unsigned long long orc_ops_read(void) { return 0; }
void orc_ops_reset(void) {}
#endif

#define SIGMA 0.2                     /* engine/sim_terms.hpp:15 */
#define SQRT2 1.4142135623730951      /* engine/sim_terms.hpp:16 */
#define D(a, b) (I->dist[(size_t)(a) * (size_t)(I->n + 1) + (size_t)(b)])

/* Feasibility of customer q seen from POI `at` at time t
 * (engine/sim_terms.hpp:522-531 and the repeated inline checks). */
static int feas_from(const orc_instance *I, int q, int at, double t) {
    double ta = D(at, q + 1);
    double arr = t + ta;
    OP(2);
    if (arr > I->close[q]) return 0;
    double st = arr > I->open[q] ? arr : I->open[q];
    OP(4);
    if (st + I->dur[q] + I->d2d[q] > I->tmax) return 0;
    return 1;
}

/* ---- NS family: engine/sim_terms.hpp:18-34 (t_ns) ------------------- */
static double t_ns(const orc_instance *I, int i, double fin, const unsigned char *vis) {
    int pi = i + 1;
    double total = 0.0;
    for (int j = 0; j < I->n; j++) {
        if (j == i || vis[j]) continue;
        double t = D(pi, j + 1);
        OP(1);
        if (t < 1e-10) continue;
        double arr = fin + t;
        OP(2);
        if (arr > I->close[j]) continue;
        double st = arr > I->open[j] ? arr : I->open[j];
        OP(4);
        if (st + I->dur[j] + I->d2d[j] > I->tmax) continue;
        OP(2);
        total += I->score[j] / t;
    }
    return total;
}

/* engine/sim_terms.hpp:36-54 */
static double t_regret(const orc_instance *I, int i, int cur, double ct, double fin_i,
                       const unsigned char *vis) {
    int pi = i + 1;
    double total = 0.0;
    for (int j = 0; j < I->n; j++) {
        if (j == i || vis[j]) continue;
        double an = ct + D(cur, j + 1);
        OP(2);
        if (an > I->close[j]) continue;
        double sn = an > I->open[j] ? an : I->open[j];
        OP(4);
        if (sn + I->dur[j] + I->d2d[j] > I->tmax) continue;
        double aa = fin_i + D(pi, j + 1);
        OP(2);
        if (aa > I->close[j]) { OP(1); total += I->score[j]; continue; }
        double sa = aa > I->open[j] ? aa : I->open[j];
        OP(4);
        if (sa + I->dur[j] + I->d2d[j] > I->tmax) { OP(1); total += I->score[j]; }
    }
    return total;
}

/* engine/sim_terms.hpp:56-74 (m_sig = 1) and :295-314 (t_ns_rob_m). */
static double t_ns_rob_m(const orc_instance *I, int i, double fin_i, double dur_i,
                         const unsigned char *vis, double msig, int general) {
    int pi = i + 1;
    /* t_ns_rob uses SIM_SIGMA directly; t_ns_rob_m uses ms = m * SIM_SIGMA. */
    double ms = general ? msig * SIGMA : SIGMA;
    double frob = fin_i + dur_i * ms;
    double k1 = 1.0 + ms;
    OP(general ? 4 : 3);
    double total = 0.0;
    for (int j = 0; j < I->n; j++) {
        if (j == i || vis[j]) continue;
        double t = D(pi, j + 1);
        OP(1);
        if (t < 1e-10) continue;
        double arr = frob + t;
        OP(2);
        if (arr > I->close[j]) continue;
        double st = arr > I->open[j] ? arr : I->open[j];
        OP(5);
        if (st + I->dur[j] * k1 + I->d2d[j] > I->tmax) continue;
        OP(2);
        total += I->score[j] / t;
    }
    return total;
}

/* engine/sim_terms.hpp:76-93 */
static double t_maxn(const orc_instance *I, int i, double fin_i, const unsigned char *vis) {
    int pi = i + 1;
    double best = 0.0;
    for (int j = 0; j < I->n; j++) {
        if (j == i || vis[j]) continue;
        double t = D(pi, j + 1);
        OP(1);
        if (t < 1e-10) continue;
        double arr = fin_i + t;
        OP(2);
        if (arr > I->close[j]) continue;
        double st = arr > I->open[j] ? arr : I->open[j];
        OP(4);
        if (st + I->dur[j] + I->d2d[j] > I->tmax) continue;
        double v = I->score[j] / t;
        OP(2);
        if (v > best) best = v;
    }
    return best;
}

/* engine/sim_terms.hpp:95-127 (rob = 0) and :156-190 (rob = 1) */
static double t_reach2x(const orc_instance *I, int i, double fin_i, double dur_i,
                        const unsigned char *vis, int rob) {
    int pi = i + 1;
    double k1 = 1.0 + SIGMA;
    double f0 = fin_i;
    if (rob) { f0 = fin_i + dur_i * SIGMA; OP(3); }
    double best_v = 0.0, fj = 0.0;
    int bj = -1;
    for (int j = 0; j < I->n; j++) {
        if (j == i || vis[j]) continue;
        double t = D(pi, j + 1);
        OP(1);
        if (t < 1e-10) continue;
        double arr = f0 + t;
        OP(2);
        if (arr > I->close[j]) continue;
        double st = arr > I->open[j] ? arr : I->open[j];
        double lim;
        if (rob) { lim = st + I->dur[j] * k1 + I->d2d[j]; OP(5); }
        else { lim = st + I->dur[j] + I->d2d[j]; OP(4); }
        if (lim > I->tmax) continue;
        double v = I->score[j] / t;
        OP(2);
        if (v > best_v) {
            best_v = v; bj = j;
            fj = rob ? st + I->dur[j] * k1 : st + I->dur[j];
            OP(rob ? 2 : 1);
        }
    }
    if (bj < 0) return 0.0;
    int pj = bj + 1;
    double best_v2 = 0.0, s2 = 0.0;
    for (int k = 0; k < I->n; k++) {
        if (k == i || k == bj || vis[k]) continue;
        double t = D(pj, k + 1);
        OP(1);
        if (t < 1e-10) continue;
        double arr = fj + t;
        OP(2);
        if (arr > I->close[k]) continue;
        double st = arr > I->open[k] ? arr : I->open[k];
        double lim;
        if (rob) { lim = st + I->dur[k] * k1 + I->d2d[k]; OP(5); }
        else { lim = st + I->dur[k] + I->d2d[k]; OP(4); }
        if (lim > I->tmax) continue;
        double v = I->score[k] / t;
        OP(2);
        if (v > best_v2) { best_v2 = v; s2 = I->score[k]; }
    }
    OP(1);
    return I->score[bj] + s2;
}

/* engine/sim_terms.hpp:129-154 */
static double t_skiploss(const orc_instance *I, int i, int cur, double ct,
                         const unsigned char *vis) {
    int pi = i + 1;
    double best_v = 0.0, fq = 0.0;
    int bq = -1;
    for (int q = 0; q < I->n; q++) {
        if (q == i || vis[q]) continue;
        double t = D(cur, q + 1);
        OP(1);
        if (t < 1e-10) continue;
        double arr = ct + t;
        OP(2);
        if (arr > I->close[q]) continue;
        double st = arr > I->open[q] ? arr : I->open[q];
        OP(4);
        if (st + I->dur[q] + I->d2d[q] > I->tmax) continue;
        double v = I->score[q] / t;
        OP(2);
        if (v > best_v) { best_v = v; bq = q; fq = st + I->dur[q]; OP(1); }
    }
    if (bq < 0) return 0.0;
    double tqi = D(bq + 1, pi);
    double ai = fq + tqi;
    OP(2);
    if (ai > I->close[i]) return I->score[i];
    double si = ai > I->open[i] ? ai : I->open[i];
    OP(4);
    if (si + I->dur[i] + I->d2d[i] > I->tmax) return I->score[i];
    return 0.0;
}

/* engine/sim_terms.hpp:192-213 */
static double t_nstw(const orc_instance *I, int i, double fin_i, const unsigned char *vis) {
    int pi = i + 1;
    double total = 0.0;
    for (int j = 0; j < I->n; j++) {
        if (j == i || vis[j]) continue;
        double t = D(pi, j + 1);
        OP(1);
        if (t < 1e-10) continue;
        double arr = fin_i + t;
        OP(2);
        if (arr > I->close[j]) continue;
        double st = arr > I->open[j] ? arr : I->open[j];
        OP(4);
        if (st + I->dur[j] + I->d2d[j] > I->tmax) continue;
        double wait = I->open[j] - arr;
        OP(2);
        if (wait < 0.0) wait = 0.0;
        double win = I->close[j] - I->open[j];
        double w;
        OP(2);
        if (win > 1e-10) { w = 1.0 - wait / win; OP(2); }
        else w = 1.0;
        OP(1);
        if (w < 0.0) w = 0.0;
        OP(3);
        total += (I->score[j] / t) * w;
    }
    return total;
}

/* engine/sim_terms.hpp:215-219 */
static double t_rand(int i) {
    long long x = (((long long)i + 1) * 1103515245LL + 12345LL) & 0x7FFFFFFFLL;
    x = (x * 1103515245LL + 12345LL) & 0x7FFFFFFFLL;
    OP(1);
    return (double)(x % 100000LL) / 1000.0;
}

/* engine/sim_terms.hpp:221-223 */
static double ncdf(double z) {
    OP(4);
    return 0.5 * (1.0 + erf(z * 0.7071067811865476));
}

/* engine/sim_terms.hpp:225-257 */
static double t_eloss(const orc_instance *I, int c, int cur, double ct,
                      const unsigned char *vis) {
    const double SIG = 0.2;
    int pc = c + 1;
    double ac = ct + D(cur, pc);
    double sc = ac > I->open[c] ? ac : I->open[c];
    double mfc = sc + I->dur[c];
    double sdc = SIG * I->dur[c];
    OP(4);
    double eloss = 0.0;
    for (int j = 0; j < I->n; j++) {
        if (vis[j] || j == c) continue;
        int pj = j + 1;
        double an = ct + D(cur, pj);
        OP(2);
        if (an > I->close[j]) continue;
        double ma = mfc + D(pc, pj);
        double pw;
        OP(2);
        if (sdc > 1e-9) { OP(2); pw = ncdf((I->close[j] - ma) / sdc); }
        else { OP(1); pw = ma <= I->close[j] ? 1.0 : 0.0; }
        double ms, vs;
        OP(1);
        if (ma > I->open[j]) { ms = ma; vs = sdc * sdc; OP(1); }
        else { ms = I->open[j]; vs = 0.0; }
        double sdj = SIG * I->dur[j];
        double vb = vs + sdj * sdj;
        double mb = ms + I->dur[j] + I->d2d[j];
        double pb;
        OP(6);
        if (vb > 1e-18) { OP(3); pb = ncdf((I->tmax - mb) / pow(vb, 0.5)); }
        else { OP(1); pb = mb <= I->tmax ? 1.0 : 0.0; }
        OP(4);
        eloss += I->score[j] * (1.0 - pw * pb);
    }
    return eloss;
}

/* engine/sim_terms.hpp:259-275 (alpha = 2) and :277-293 (alpha = 0.5) */
static double t_ns_alpha(const orc_instance *I, int i, double fin, const unsigned char *vis,
                         int half) {
    int pi = i + 1;
    double total = 0.0;
    for (int j = 0; j < I->n; j++) {
        if (j == i || vis[j]) continue;
        double t = D(pi, j + 1);
        OP(1);
        if (t < 1e-10) continue;
        double arr = fin + t;
        OP(2);
        if (arr > I->close[j]) continue;
        double st = arr > I->open[j] ? arr : I->open[j];
        OP(4);
        if (st + I->dur[j] + I->d2d[j] > I->tmax) continue;
        OP(3);
        if (half) total += I->score[j] / sqrt(t);
        else total += I->score[j] / (t * t);
    }
    return total;
}

/* engine/sim_terms.hpp:316-334 */
static double t_nnfwd(const orc_instance *I, int i, double fin_i, const unsigned char *vis) {
    int pi = i + 1;
    double dmin = -1.0;
    for (int j = 0; j < I->n; j++) {
        if (j == i || vis[j]) continue;
        double t = D(pi, j + 1);
        OP(1);
        if (t < 1e-10) continue;
        double arr = fin_i + t;
        OP(2);
        if (arr > I->close[j]) continue;
        double st = arr > I->open[j] ? arr : I->open[j];
        OP(4);
        if (st + I->dur[j] + I->d2d[j] > I->tmax) continue;
        OP(2);
        if (dmin < 0.0 || t < dmin) dmin = t;
    }
    if (dmin < 0.0) return 0.0;
    OP(2);
    return 1.0 / (1.0 + dmin);
}

/* engine/sim_terms.hpp:336-359 */
static double t_nnk3(const orc_instance *I, int i, double fin_i, const unsigned char *vis) {
    int pi = i + 1;
    double d1 = -1.0, d2 = -1.0, d3 = -1.0;
    for (int j = 0; j < I->n; j++) {
        if (j == i || vis[j]) continue;
        double t = D(pi, j + 1);
        OP(1);
        if (t < 1e-10) continue;
        double arr = fin_i + t;
        OP(2);
        if (arr > I->close[j]) continue;
        double st = arr > I->open[j] ? arr : I->open[j];
        OP(4);
        if (st + I->dur[j] + I->d2d[j] > I->tmax) continue;
        OP(2);
        if (d1 < 0.0 || t < d1) { d3 = d2; d2 = d1; d1 = t; }
        else if (OP(2), d2 < 0.0 || t < d2) { d3 = d2; d2 = t; }
        else if (OP(2), d3 < 0.0 || t < d3) { d3 = t; }
    }
    if (d1 < 0.0) return 0.0;
    double tot = d1, k = 1.0;
    if (d2 >= 0.0) { tot += d2; k += 1.0; OP(2); }
    if (d3 >= 0.0) { tot += d3; k += 1.0; OP(2); }
    OP(3);
    return 1.0 / (1.0 + tot / k);
}

/* engine/sim_terms.hpp:361-381 */
static double t_detour(const orc_instance *I, int i, int cur, double fin_i, double ta_i,
                       const unsigned char *vis) {
    int pi = i + 1;
    double total = 0.0, cnt = 0.0;
    for (int j = 0; j < I->n; j++) {
        if (j == i || vis[j]) continue;
        double t = D(pi, j + 1);
        OP(1);
        if (t < 1e-10) continue;
        double arr = fin_i + t;
        OP(2);
        if (arr > I->close[j]) continue;
        double st = arr > I->open[j] ? arr : I->open[j];
        OP(4);
        if (st + I->dur[j] + I->d2d[j] > I->tmax) continue;
        OP(4);
        total += ta_i + t - D(cur, j + 1);
        cnt += 1.0;
    }
    if (cnt == 0.0) return 0.0;
    OP(1);
    return total / cnt;
}

/* engine/sim_terms.hpp:383-401 */
static double t_scorerank(const orc_instance *I, int i, double fin_i, const unsigned char *vis) {
    int pi = i + 1;
    double below = 0.0, tot = 0.0;
    for (int j = 0; j < I->n; j++) {
        if (j == i || vis[j]) continue;
        double t = D(pi, j + 1);
        OP(1);
        if (t < 1e-10) continue;
        double arr = fin_i + t;
        OP(2);
        if (arr > I->close[j]) continue;
        double st = arr > I->open[j] ? arr : I->open[j];
        OP(4);
        if (st + I->dur[j] + I->d2d[j] > I->tmax) continue;
        OP(2);
        tot += 1.0;
        if (I->score[j] < I->score[i]) { below += 1.0; OP(1); }
    }
    if (tot == 0.0) return 1.0;
    OP(1);
    return below / tot;
}

/* engine/sim_terms.hpp:405-441 (rob = 0) and :443-479 (rob = 1): K greedy
 * waypoints by score/distance from candidate c. */
static double t_reach_k(const orc_instance *I, int c, int cur0, double ct,
                        const unsigned char *vis, int K, int rob, unsigned char *scr) {
    for (int j = 0; j < I->n; j++) scr[j] = vis[j];
    scr[c] = 1;
    double k1 = 1.0 + SIGMA;
    double tac = D(cur0, c + 1);
    double ac = ct + tac;
    double sc = ac > I->open[c] ? ac : I->open[c];
    int cur = c + 1;
    double t = rob ? sc + I->dur[c] * k1 : sc + I->dur[c];
    OP(rob ? 4 : 3);
    double total = 0.0;
    for (int step = 0; step < K; step++) {
        int best = -1;
        double bv = -1.0;
        for (int q = 0; q < I->n; q++) {
            if (scr[q]) continue;
            double ta = D(cur, q + 1);
            double arr = t + ta;
            OP(2);
            if (arr > I->close[q]) continue;
            double st = arr > I->open[q] ? arr : I->open[q];
            double lim;
            if (rob) { lim = st + I->dur[q] * k1 + I->d2d[q]; OP(5); }
            else { lim = st + I->dur[q] + I->d2d[q]; OP(4); }
            if (lim > I->tmax) continue;
            double tt = ta > 1e-10 ? ta : 1e-10;
            double v = I->score[q] / tt;
            OP(4);
            if (v > bv) { bv = v; best = q; }
        }
        if (best < 0) break;
        double ta = D(cur, best + 1);
        double arr = t + ta;
        double st = arr > I->open[best] ? arr : I->open[best];
        t = rob ? st + I->dur[best] * k1 : st + I->dur[best];
        cur = best + 1;
        total += I->score[best];
        scr[best] = 1;
        OP(rob ? 5 : 4);
    }
    return total;
}

/* engine/sim_terms.hpp:481-519 */
static double t_nnchain(const orc_instance *I, int c, int cur0, double ct,
                        const unsigned char *vis, int K, unsigned char *scr) {
    for (int j = 0; j < I->n; j++) scr[j] = vis[j];
    scr[c] = 1;
    double tac = D(cur0, c + 1);
    double ac = ct + tac;
    double sc = ac > I->open[c] ? ac : I->open[c];
    int cur = c + 1;
    double t = sc + I->dur[c];
    OP(3);
    double total_d = 0.0;
    int done = 0;
    for (int step = 0; step < K; step++) {
        int best = -1;
        double bd = -1.0;
        for (int q = 0; q < I->n; q++) {
            if (scr[q]) continue;
            double ta = D(cur, q + 1);
            OP(1);
            if (ta < 1e-10) continue;
            double arr = t + ta;
            OP(2);
            if (arr > I->close[q]) continue;
            double st = arr > I->open[q] ? arr : I->open[q];
            OP(4);
            if (st + I->dur[q] + I->d2d[q] > I->tmax) continue;
            OP(2);
            if (bd < 0.0 || ta < bd) { bd = ta; best = q; }
        }
        if (best < 0) break;
        double arr = t + bd;
        double st = arr > I->open[best] ? arr : I->open[best];
        t = st + I->dur[best];
        cur = best + 1;
        total_d += bd;
        done++;
        scr[best] = 1;
        OP(4);
    }
    if (done == 0) return 0.0;
    OP(3);
    return 1.0 / (1.0 + total_d / done);
}

/* engine/sim_terms.hpp:533-580: FRAGCNT, FRAGSCORE, DOWNSLACK, PBUST. */
static void struct_feats(const orc_instance *I, int c, int cur, double ct, double sigma,
                         const unsigned char *vis, double out[4]) {
    double tac = D(cur, c + 1);
    double ac = ct + tac;
    double sc = ac > I->open[c] ? ac : I->open[c];
    double fc = sc + I->dur[c];
    int cp = c + 1;
    OP(3);
    double fcnt = 0.0, fscore = 0.0, dslack = 1e18;
    double bv = -1.0;
    int bq = -1;
    for (int q = 0; q < I->n; q++) {
        if (vis[q] || q == c) continue;
        int now = feas_from(I, q, cur, ct);
        int after = feas_from(I, q, cp, fc);
        if (now && !after) { fcnt += 1.0; fscore += I->score[q]; OP(2); }
        if (after) {
            double aq = fc + D(cp, q + 1);
            double slack = I->close[q] - aq;
            OP(3);
            if (slack < dslack) dslack = slack;
            double t = D(cp, q + 1);
            OP(1);
            if (t > 1e-10) {
                double v = I->score[q] / t;
                OP(2);
                if (v > bv) { bv = v; bq = q; }
            }
        }
    }
    OP(1);
    if (dslack > 1e17) dslack = 0.0;
    double pbust = 0.0;
    if (bq >= 0) {
        double ab = fc + D(cp, bq + 1);
        double sd = sigma * I->dur[c];
        OP(3);
        if (sd > 1e-12) {
            double z = (I->close[bq] - ab) / (sd * SQRT2);
            pbust = 0.5 * erfc(z);
            OP(5);
        } else {
            OP(1);
            pbust = ab > I->close[bq] ? 1.0 : 0.0;
        }
    }
    out[0] = fcnt;
    out[1] = fscore;
    out[2] = dslack;
    out[3] = pbust;
}

/* engine/sim_core.hpp:11-41.  The constant of a -2 node is args/1000.0 at
 * evaluation time; '/' is protected below |b| < 1e-10; op 6 is rmod. */
double orc_eval_postfix(const int32_t *ops, const int32_t *args, int len,
                        const double *terms, double *stack) {
    int sp = 0;
    for (int k = 0; k < len; k++) {
        int o = ops[k];
        if (o == -1) {
            stack[sp++] = terms[args[2 * k]];
        } else if (o == -2) {
            OP(1);
            stack[sp++] = args[2 * k] / 1000.0;
        } else {
            double b = stack[sp - 1];
            double a = stack[sp - 2];
            sp -= 2;
            double r;
            switch (o) {
            case 0: r = a + b; OP(1); break;
            case 1: r = a - b; OP(1); break;
            case 2: r = a * b; OP(1); break;
            case 3: OP(2); r = fabs(b) < 1e-10 ? 1.0 : a / b; break;
            case 4: OP(1); r = a < b ? a : b; break;
            case 5: OP(1); r = a > b ? a : b; break;
            default: {
                double base = a > 0.0 ? a : 0.0;
                double e = b;
                if (e < 0.0) e = 0.0;
                else if (e > 1.0) e = 1.0;
                OP(5);
                r = pow(base, 1.0 + e);
            }
            }
            stack[sp++] = r;
        }
    }
    return stack[0];
}

typedef struct {
    unsigned char *scr;
    double terms[ORC_N_TERMS];
    double stack[ORC_MAX_STACK];
    double sbuf[4];
    int cap;            /* candidate whose slot values are copied out, -1 = none */
    double *cap_out;
    int cap_hit;
} orc_ws;

static void ws_init(orc_ws *w, int n) {
    w->scr = (unsigned char *)calloc((size_t)n + 1, 1);
    memset(w->terms, 0, sizeof w->terms);
    memset(w->sbuf, 0, sizeof w->sbuf);
    w->cap = -1;
    w->cap_out = NULL;
    w->cap_hit = 0;
}
static void ws_free(orc_ws *w) { free(w->scr); }

#define NEED(s) (need == NULL || need[s])

/* engine/sim_core.hpp:43-204: one decision over every unvisited candidate. */
static int pick_best_ws(const orc_instance *I, int cur, double ct, int rem,
                        const int32_t *ops, const int32_t *args, int len,
                        const unsigned char *vis, const int64_t *mk, const double *zp,
                        const unsigned char *need, double *pri_out, orc_ws *w) {
    const int n = I->n;
    const double tmax = I->tmax;
    double *T = w->terms;
    int need_ns = (mk[10] || mk[28] || mk[29]) && (NEED(10) || NEED(28) || NEED(29));
    int need_st = (mk[17] || mk[18] || mk[19] || mk[20] || mk[28] || mk[29]) &&
                  (NEED(17) || NEED(18) || NEED(19) || NEED(20) || NEED(28) || NEED(29));

    OP(2);
    double sat = tmax > 0.0 ? ct / tmax : 0.0;
    double vfrac = 0.0, nfeas = 0.0, fsc = 0.0, nfr = 0.0;
    if ((mk[25] || mk[26] || mk[51] || mk[52]) && (NEED(25) || NEED(26) || NEED(51) || NEED(52))) {
        double nvis = 0.0;
        for (int j = 0; j < n; j++) {
            if (vis[j]) { nvis += 1.0; OP(1); continue; }
            double ta = D(cur, j + 1);
            double a = ct + ta;
            OP(2);
            if (a > I->close[j]) continue;
            double st = a > I->open[j] ? a : I->open[j];
            OP(4);
            if (st + I->dur[j] + I->d2d[j] > tmax) continue;
            nfeas += 1.0;
            fsc += I->score[j];
            OP(7);
            if (st + I->dur[j] * (1.0 + SIGMA) + I->d2d[j] <= tmax) { nfr += 1.0; OP(1); }
        }
        OP(1);
        vfrac = nvis / n;
    }

    int best = -1;
    double best_pri = -1e300;
    for (int i = 0; i < n; i++) {
        if (pri_out) pri_out[i] = -INFINITY;
        if (vis[i]) continue;
        int pi = i + 1;
        double ta = D(cur, pi);
        double arr = ct + ta;
        OP(2);
        if (arr > I->close[i]) continue;
        double st = arr > I->open[i] ? arr : I->open[i];
        OP(4);
        if (st + I->dur[i] + I->d2d[i] > tmax) continue;

        double to = I->open[i] - ct;
        double tc = I->close[i] - ct;
        double tsv = to > ta ? to : ta;
        double tfv = tsv + I->dur[i];
        OP(4);
        T[0] = I->score[i];
        T[1] = I->dur[i];
        T[2] = to;
        T[3] = tc;
        T[4] = ta;
        T[5] = I->d2d[i];
        T[6] = tsv;
        T[7] = tfv;
        T[8] = tc - ta;
        T[9] = rem * tmax + (tmax - ct);
        OP(4);

        double fin = ct + tfv;
        OP(1);
        double nsv = 0.0;
        if (need_ns) nsv = t_ns(I, i, fin, vis);
        T[10] = mk[10] ? nsv : 0.0;
        T[11] = ((mk[11] || mk[33]) && (NEED(11) || NEED(33))) ? t_regret(I, i, cur, ct, fin, vis) : 0.0;
        T[12] = (mk[12] && NEED(12)) ? t_ns_rob_m(I, i, fin, I->dur[i], vis, 1.0, 0) : 0.0;
        T[13] = (mk[13] && NEED(13)) ? t_maxn(I, i, fin, vis) : 0.0;
        T[14] = ((mk[14] || mk[32]) && (NEED(14) || NEED(32))) ? t_reach2x(I, i, fin, 0.0, vis, 0) : 0.0;
        T[15] = (mk[15] && NEED(15)) ? t_nstw(I, i, fin, vis) : 0.0;
        T[16] = (mk[16] && NEED(16)) ? t_rand(i) : 0.0;

        double frag = 0.0;
        if (need_st) {
            struct_feats(I, i, cur, ct, SIGMA, vis, w->sbuf);
            frag = w->sbuf[1];
        }
        T[17] = mk[17] ? w->sbuf[0] : 0.0;
        T[18] = mk[18] ? w->sbuf[1] : 0.0;
        T[19] = mk[19] ? w->sbuf[2] : 0.0;
        T[20] = mk[20] ? w->sbuf[3] : 0.0;

        T[21] = (mk[21] && NEED(21)) ? t_reach_k(I, i, cur, ct, vis, 1, 0, w->scr) : 0.0;
        T[22] = (mk[22] && NEED(22)) ? t_reach_k(I, i, cur, ct, vis, 3, 0, w->scr) : 0.0;
        T[23] = (mk[23] && NEED(23)) ? t_reach_k(I, i, cur, ct, vis, 5, 0, w->scr) : 0.0;

        T[24] = mk[24] ? sat : 0.0;
        T[25] = mk[25] ? vfrac : 0.0;
        T[26] = mk[26] ? nfeas : 0.0;
        T[27] = mk[27] ? (double)rem : 0.0;
        if (mk[28] && NEED(28)) { OP(2); T[28] = frag / (nsv + 1.0); } else T[28] = 0.0;
        if (mk[29] && NEED(29)) { OP(1); T[29] = frag - nsv; } else T[29] = 0.0;
        T[30] = ((mk[30] || mk[34]) && (NEED(30) || NEED(34))) ? t_eloss(I, i, cur, ct, vis) : 0.0;
        if (mk[31] && NEED(31)) { OP(2); T[31] = (T[3] - zp[0]) / zp[5]; } else T[31] = 0.0;
        if (mk[32] && NEED(32)) { OP(2); T[32] = (T[14] - zp[1]) / zp[6]; } else T[32] = 0.0;
        if (mk[33] && NEED(33)) { OP(2); T[33] = (T[11] - zp[2]) / zp[7]; } else T[33] = 0.0;
        if (mk[34] && NEED(34)) { OP(2); T[34] = (T[30] - zp[3]) / zp[8]; } else T[34] = 0.0;
        if (mk[35] && NEED(35)) { OP(2); T[35] = (T[5] - zp[4]) / zp[9]; } else T[35] = 0.0;
        T[36] = mk[36] ? 0.1 : 0.0;
        T[37] = mk[37] ? 0.2 : 0.0;
        T[38] = mk[38] ? 0.3 : 0.0;
        T[39] = mk[39] ? 0.4 : 0.0;
        T[40] = mk[40] ? 0.5 : 0.0;
        T[41] = (mk[41] && NEED(41)) ? t_reach2x(I, i, fin, I->dur[i], vis, 1) : 0.0;
        {
            double den = tmax - ct;
            OP(2);
            if (den < 1.0) den = 1.0;
            if (mk[42] && NEED(42)) { OP(2); T[42] = (ta + I->dur[i]) / den; } else T[42] = 0.0;
        }
        if (mk[43] && NEED(43)) {
            OP(1);
            T[43] = t_reach2x(I, i, fin, I->dur[i], vis, 1) - t_regret(I, i, cur, ct, fin, vis);
        } else T[43] = 0.0;
        T[44] = (mk[44] && NEED(44)) ? t_skiploss(I, i, cur, ct, vis) : 0.0;
        if (mk[45] && NEED(45)) {
            OP(1);
            T[45] = t_ns(I, i, fin, vis) - t_ns_rob_m(I, i, fin, I->dur[i], vis, 1.0, 0);
        } else T[45] = 0.0;
        {
            double den = ta + I->d2d[i];
            OP(2);
            if (den < 1.0) den = 1.0;
            if (mk[46] && NEED(46)) {
                OP(2);
                T[46] = (I->score[i] + t_ns_rob_m(I, i, fin, I->dur[i], vis, 1.0, 0)) / den;
            } else T[46] = 0.0;
        }
        T[47] = (mk[47] && NEED(47)) ? t_reach_k(I, i, cur, ct, vis, 3, 1, w->scr) : 0.0;
        T[48] = (mk[48] && NEED(48)) ? t_reach_k(I, i, cur, ct, vis, 5, 1, w->scr) : 0.0;
        T[49] = (mk[49] && NEED(49)) ? t_ns_alpha(I, i, fin, vis, 0) : 0.0;
        T[50] = (mk[50] && NEED(50)) ? t_ns_alpha(I, i, fin, vis, 1) : 0.0;
        T[51] = mk[51] ? fsc : 0.0;
        T[52] = mk[52] ? nfr : 0.0;
        T[53] = (mk[53] && NEED(53)) ? t_ns_rob_m(I, i, fin, I->dur[i], vis, 0.5, 1) : 0.0;
        T[54] = (mk[54] && NEED(54)) ? t_ns_rob_m(I, i, fin, I->dur[i], vis, 2.0, 1) : 0.0;
        T[55] = (mk[55] && NEED(55)) ? t_nnfwd(I, i, fin, vis) : 0.0;
        T[56] = (mk[56] && NEED(56)) ? t_nnk3(I, i, fin, vis) : 0.0;
        T[57] = (mk[57] && NEED(57)) ? t_nnchain(I, i, cur, ct, vis, 3, w->scr) : 0.0;
        if (mk[58] && NEED(58)) { OP(2); T[58] = D(cur, 0) + I->d2d[i] - ta; } else T[58] = 0.0;
        T[59] = (mk[59] && NEED(59)) ? t_detour(I, i, cur, fin, ta, vis) : 0.0;
        T[60] = (mk[60] && NEED(60)) ? t_scorerank(I, i, fin, vis) : 0.0;

        if (i == w->cap || w->cap == -2) {
            memcpy(w->cap_out + (w->cap == -2 ? (size_t)i * ORC_N_TERMS : 0), T,
                   sizeof(double) * ORC_N_TERMS);
            w->cap_hit++;
        }
        double pri = orc_eval_postfix(ops, args, len, T, w->stack);
        OP(2);
        if (pri != pri) pri = -1e300;
        if (pri_out) pri_out[i] = pri;
        if (pri > best_pri) { best_pri = pri; best = i; }
    }
    return best;
}

int orc_pick_best(const orc_instance *I, int current, double ctime, int remaining,
                  const int32_t *ops, const int32_t *args, int len,
                  const unsigned char *visited, const int64_t *mask, const double *zp,
                  const unsigned char *need, double *pri_out) {
    orc_ws w;
    ws_init(&w, I->n);
    int b = pick_best_ws(I, current, ctime, remaining, ops, args, len, visited, mask, zp, need,
                         pri_out, &w);
    ws_free(&w);
    return b;
}

/* engine/sim_core.hpp:207-238: vehicle v closes its route before v + 1;
 * the sampled duration advances time, feasibility uses the expected one. */
static double sim_one_ws(const orc_instance *I, const int32_t *ops, const int32_t *args, int len,
                         const double *sampled, const int64_t *mk, const double *zp,
                         const unsigned char *need, unsigned char *vis, orc_ws *w) {
    for (int i = 0; i < I->n; i++) vis[i] = 0;
    double total = 0.0;
    for (int v = 0; v < I->m; v++) {
        int rem = I->m - v - 1;
        int cur = 0;
        double ct = 0.0;
        for (;;) {
            int b = pick_best_ws(I, cur, ct, rem, ops, args, len, vis, mk, zp, need, NULL, w);
            if (b < 0) break;
            int pb = b + 1;
            double ta = D(cur, pb);
            double arr = ct + ta;
            double st = arr > I->open[b] ? arr : I->open[b];
            double fin = st + sampled[b];
            OP(5);
            cur = pb;
            ct = fin;
            total += I->score[b];
            vis[b] = 1;
        }
    }
    return total;
}

double orc_sim_one(const orc_instance *I, const int32_t *ops, const int32_t *args, int len,
                   const double *sampled, const int64_t *mask, const double *zp,
                   const unsigned char *need) {
    orc_ws w;
    ws_init(&w, I->n);
    unsigned char *vis = (unsigned char *)calloc((size_t)I->n + 1, 1);
    double r = sim_one_ws(I, ops, args, len, sampled, mask, zp, need, vis, &w);
    free(vis);
    ws_free(&w);
    return r;
}

/* engine/sim_core.hpp:241-274: the m vehicles advance one POI per round. */
double orc_sim_one_rr(const orc_instance *I, const int32_t *ops, const int32_t *args, int len,
                      const double *sampled, const int64_t *mk, const double *zp) {
    orc_ws w;
    ws_init(&w, I->n);
    unsigned char *vis = (unsigned char *)calloc((size_t)I->n + 1, 1);
    int *cv = (int *)calloc((size_t)I->m + 1, sizeof(int));
    double *tv = (double *)calloc((size_t)I->m + 1, sizeof(double));
    unsigned char *act = (unsigned char *)calloc((size_t)I->m + 1, 1);
    for (int v = 0; v < I->m; v++) { cv[v] = 0; tv[v] = 0.0; act[v] = 1; }
    double total = 0.0;
    int n_act = I->m;
    while (n_act > 0) {
        for (int v = 0; v < I->m; v++) {
            if (!act[v]) continue;
            int rem = I->m - v - 1;
            int b = pick_best_ws(I, cv[v], tv[v], rem, ops, args, len, vis, mk, zp, NULL, NULL, &w);
            if (b < 0) { act[v] = 0; n_act--; continue; }
            int pb = b + 1;
            double ta = D(cv[v], pb);
            double arr = tv[v] + ta;
            double st = arr > I->open[b] ? arr : I->open[b];
            double fin = st + sampled[b];
            cv[v] = pb;
            tv[v] = fin;
            total += I->score[b];
            vis[b] = 1;
        }
    }
    free(act); free(tv); free(cv); free(vis);
    ws_free(&w);
    return total;
}

/* engine/sim_core.hpp:277-312 with the zp layout of engine/sim_host.hpp:96-105. */
void orc_zstats(const orc_instance *I, double zp[10]) {
    double zs[5] = {0, 0, 0, 0, 0}, zq[5] = {0, 0, 0, 0, 0};
    double cnt = 0.0;
    unsigned char *vis = (unsigned char *)calloc((size_t)I->n + 1, 1);
    for (int i = 0; i < I->n; i++) {
        double ta = D(0, i + 1);
        double arr = ta;
        if (arr > I->close[i]) continue;
        double st = arr > I->open[i] ? arr : I->open[i];
        if (st + I->dur[i] + I->d2d[i] > I->tmax) continue;
        double to = I->open[i];
        double tsv = to > ta ? to : ta;
        double fin = tsv + I->dur[i];
        double f[5];
        f[0] = I->close[i];
        f[1] = t_reach2x(I, i, fin, 0.0, vis, 0);
        f[2] = t_regret(I, i, 0, 0.0, fin, vis);
        f[3] = t_eloss(I, i, 0, 0.0, vis);
        f[4] = I->d2d[i];
        for (int k = 0; k < 5; k++) { zs[k] += f[k]; zq[k] += f[k] * f[k]; }
        cnt += 1.0;
    }
    free(vis);
    if (cnt < 1.0) cnt = 1.0;
    for (int k = 0; k < 5; k++) {
        double mean = zs[k] / cnt;
        double var = zq[k] / cnt - mean * mean;
        if (var < 0.0) var = 0.0;
        zp[k] = mean;
        zp[5 + k] = sqrt(var) + 1e-9;
    }
}

void orc_sim_batch(const orc_instance *I, const int32_t *ops, const int32_t *args, int len,
                   const double *all_sampled, int n_samples, const int64_t *mask,
                   const double *zp, double *out) {
    orc_ws w;
    ws_init(&w, I->n);
    unsigned char *vis = (unsigned char *)calloc((size_t)I->n + 1, 1);
    for (int s = 0; s < n_samples; s++)
        out[s] = sim_one_ws(I, ops, args, len, all_sampled + (size_t)s * I->n, mask, zp, NULL,
                            vis, &w);
    free(vis);
    ws_free(&w);
}

void orc_evaluate(const orc_instance *I, const int32_t *ops, const int32_t *args,
                  const int32_t *offsets, const int32_t *lengths, int P,
                  const double *sampled, int N, const int64_t *scen_base,
                  const int64_t *masks, const int32_t *mask_row, const double *zp,
                  const unsigned char *need_rows, int n_threads, double *out) {
    long long total = (long long)P * N;
#ifdef _OPENMP
    if (n_threads < 1) n_threads = 1;
#pragma omp parallel num_threads(n_threads)
#endif
    {
        orc_ws w;
        ws_init(&w, I->n);
        unsigned char *vis = (unsigned char *)calloc((size_t)I->n + 1, 1);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (long long idx = 0; idx < total; idx++) {
            int p = (int)(idx / N), s = (int)(idx % N);
            const int32_t *o = ops + offsets[p];
            const int32_t *a = args + 2 * (size_t)offsets[p];
            const double *row = sampled + (size_t)(scen_base[p] + s) * (size_t)I->n;
            const int64_t *mk = masks + (size_t)mask_row[p] * ORC_N_TERMS;
            const unsigned char *nd = need_rows ? need_rows + (size_t)p * ORC_N_TERMS : NULL;
            out[idx] = sim_one_ws(I, o, a, lengths[p], row, mk, zp, nd, vis, &w);
        }
        free(vis);
        ws_free(&w);
    }
}

int orc_terms_at(const orc_instance *I, int cand, int current, double ctime, int remaining,
                 const unsigned char *visited, const int64_t *mask, const double *zp,
                 double *out) {
    orc_ws w;
    ws_init(&w, I->n);
    w.cap = cand;
    w.cap_out = out;
    int32_t ops[1] = {-2};
    int32_t args[2] = {0, 0};
    pick_best_ws(I, current, ctime, remaining, ops, args, 1, visited, mask, zp, NULL, NULL, &w);
    ws_free(&w);
    return w.cap_hit;
}

int orc_terms_all(const orc_instance *I, int current, double ctime, int remaining,
                  const unsigned char *visited, const int64_t *mask, const double *zp,
                  double *out, double *pri_out) {
    orc_ws w;
    ws_init(&w, I->n);
    w.cap = -2;
    w.cap_out = out;
    int32_t ops[1] = {-2};
    int32_t args[2] = {0, 0};
    pick_best_ws(I, current, ctime, remaining, ops, args, 1, visited, mask, zp, NULL, pri_out, &w);
    ws_free(&w);
    return w.cap_hit;
}

double orc_terminal(const orc_instance *I, int slot, int cand, int current, double ctime,
                    int remaining, const unsigned char *visited, const double *zp) {
    int64_t mk[ORC_N_TERMS];
    for (int k = 0; k < ORC_N_TERMS; k++) mk[k] = (k == slot);
    /* a one-node rule that returns the slot value; pri_out of the candidate */
    int32_t ops[1] = {-1};
    int32_t args[2] = {slot, 0};
    double *pri = (double *)malloc(sizeof(double) * (size_t)I->n);
    orc_pick_best(I, current, ctime, remaining, ops, args, 1, visited, mk, zp, NULL, pri);
    double v = pri[cand];
    free(pri);
    return v;
}
