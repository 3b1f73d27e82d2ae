// CPU ORACLE -- test infrastructure only.
// C-ABI binding of the reference's own CPU simulator, random streams and rule
// compiler, compiled from the reference sources in place (oracle/Makefile
// target `ref`, output oracle/_ref/libsigpref.so).  Used to pin the C
// restatement (sim_oracle.c) on inputs beyond the golden fixtures and as the
// "reference" CPU baseline in bench.py.  Nothing here is product code.
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include "engine/sim_host.hpp"
#include "engine/gp.hpp"

static SimData make_sd(int n, int m, double tmax, const double* sc, const double* du,
                       const double* op, const double* cl, const double* dist, const double* d2) {
    SimData sd;
    sd.n = n; sd.m = m; sd.ld = n + 1; sd.tmax = tmax; sd.sigma = 0.2;
    sd.scores.assign(sc, sc + n); sd.durations.assign(du, du + n);
    sd.open_t.assign(op, op + n); sd.close_t.assign(cl, cl + n);
    sd.dist.assign(dist, dist + (size_t)(n + 1) * (n + 1)); sd.dist_to_depot.assign(d2, d2 + n);
    return sd;
}

extern "C" {

void ref_zp_of(int n, int m, double tmax, const double* sc, const double* du, const double* op,
               const double* cl, const double* dist, const double* d2, double* zp_out) {
    SimData sd = make_sd(n, m, tmax, sc, du, op, cl, dist, d2);
    std::vector<double> zp = zp_of(sd);
    std::memcpy(zp_out, zp.data(), 10 * sizeof(double));
}

// out[p*N+s] = sim_one(rule p, scenario row scen_base[p]+s, mask row mask_row[p]);
// OpenMP over (p, s) pairs with per-thread buffers (the reference's sim_batch
// is single-threaded; this loop only distributes its sim_one calls).
void ref_evaluate(int n, int m, double tmax, const double* sc, const double* du, const double* op,
                  const double* cl, const double* dist, const double* d2,
                  const int32_t* ops, const int32_t* args, const int32_t* offsets,
                  const int32_t* lengths, int P, const double* sampled, int N,
                  const int64_t* scen_base, const int64_t* masks, const int32_t* mask_row,
                  const double* zp, int n_threads, double* out) {
    SimData sd = make_sd(n, m, tmax, sc, du, op, cl, dist, d2);
    long long total = (long long)P * N;
#pragma omp parallel num_threads(n_threads > 0 ? n_threads : 1)
    {
        SimBuffers buf(sd.n);
#pragma omp for schedule(dynamic, 4)
        for (long long idx = 0; idx < total; idx++) {
            int p = (int)(idx / N), s = (int)(idx % N);
            out[idx] = sim_one(ops + offsets[p], args + 2 * (size_t)offsets[p], lengths[p],
                               sd.scores.data(), sd.durations.data(), sd.open_t.data(),
                               sd.close_t.data(), sd.dist.data(), sd.ld, sd.dist_to_depot.data(),
                               sd.tmax, sampled + (size_t)(scen_base[p] + s) * sd.n, sd.m, sd.n,
                               masks + (size_t)mask_row[p] * N_TERMS, zp,
                               (bool*)buf.visited.data(), buf.terms.data(), buf.stack.data(),
                               buf.struct_buf.data(), (bool*)buf.scratch.data());
        }
    }
}

// The reference's own pick_best (engine/sim_core.hpp:43-204) on one decision
// state, with the one-node rule "constant 0".  Its terms buffer is rewritten
// for every feasible candidate in index order before eval_tree reads it, so
// after the call terms_out holds all 61 slot values of the LAST feasible
// candidate, exactly as the reference computed them under `mask`.  Returns
// pick_best's choice (the first feasible candidate, or -1; terms_out is then
// left as the caller filled it).
int ref_pick_best_terms(int n, int m, double tmax, const double* sc, const double* du,
                        const double* op, const double* cl, const double* dist, const double* d2,
                        int current, double ctime, int remaining, const uint8_t* visited,
                        const int64_t* mask, const double* zp, double* terms_out) {
    SimData sd = make_sd(n, m, tmax, sc, du, op, cl, dist, d2);
    SimBuffers buf(sd.n);
    bool* vis = (bool*)buf.visited.data();
    for (int i = 0; i < n; i++) vis[i] = visited[i] != 0;
    for (int k = 0; k < N_TERMS; k++) buf.terms[k] = terms_out[k];
    const int32_t ops[1] = {-2};
    const int32_t args[2] = {0, 0};
    int best = pick_best(current, ctime, remaining, ops, args, 1, sd.scores.data(),
                         sd.durations.data(), sd.open_t.data(), sd.close_t.data(), sd.dist.data(),
                         sd.ld, sd.dist_to_depot.data(), sd.tmax, vis, sd.n, mask, zp,
                         buf.terms.data(), buf.stack.data(), buf.struct_buf.data(),
                         (bool*)buf.scratch.data());
    for (int k = 0; k < N_TERMS; k++) terms_out[k] = buf.terms[k];
    return best;
}

// One scan terminal of engine/sim_terms.hpp, called directly with the
// arguments pick_best passes it (engine/sim_core.hpp:99-197), for each of
// the n_cand candidates cand[k]: out[k].  fin is the candidate's finish
// time at the expected duration, current_time + max(open - current_time,
// travel) + duration (engine/sim_core.hpp:86-91,98).  Slots 17-20 take
// struct_feats_one's four outputs.  Returns -1 for a slot that no single
// terminal function computes.
int ref_scan_terminal(int n, int m, double tmax, const double* sc, const double* du,
                      const double* op, const double* cl, const double* dist, const double* d2,
                      int slot, int current, double ct, const uint8_t* visited,
                      const int32_t* cand, int n_cand, double* out) {
    SimData sd = make_sd(n, m, tmax, sc, du, op, cl, dist, d2);
    SimBuffers buf(sd.n);
    bool* vis = (bool*)buf.visited.data();
    bool* scr = (bool*)buf.scratch.data();
    for (int i = 0; i < n; i++) vis[i] = visited[i] != 0;
    const double *S = sd.scores.data(), *Du = sd.durations.data(), *O = sd.open_t.data();
    const double *Cl = sd.close_t.data(), *Di = sd.dist.data(), *D2 = sd.dist_to_depot.data();
    const int ld = sd.ld;
    for (int k = 0; k < n_cand; k++) {
        int i = cand[k];
        double ta = Di[current * ld + i + 1];
        double to = O[i] - ct;
        double tsv = to > ta ? to : ta;
        double fin = ct + (tsv + Du[i]);
        double v;
        switch (slot) {
        case 10: v = t_ns(i, fin, S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 11: v = t_regret(i, current, ct, fin, S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 12: v = t_ns_rob(i, fin, Du[i], S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 13: v = t_maxn(i, fin, S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 14: v = t_reach2(i, fin, S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 15: v = t_nstw(i, fin, S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 16: v = t_rand(i); break;
        case 17: case 18: case 19: case 20: {
            double sb[4] = {0, 0, 0, 0};
            struct_feats_one(i, current, ct, SIM_SIGMA, S, Du, O, Cl, Di, ld, D2, tmax, vis, n, sb);
            v = sb[slot - 17];
            break;
        }
        case 21: v = t_reach_k(i, current, ct, S, Du, O, Cl, Di, ld, D2, tmax, vis, n, 1, scr); break;
        case 22: v = t_reach_k(i, current, ct, S, Du, O, Cl, Di, ld, D2, tmax, vis, n, 3, scr); break;
        case 23: v = t_reach_k(i, current, ct, S, Du, O, Cl, Di, ld, D2, tmax, vis, n, 5, scr); break;
        case 30: v = t_eloss_chance(i, current, ct, S, Du, O, Cl, Di, ld, D2, tmax, vis, n); break;
        case 41: v = t_reach2_rob(i, fin, Du[i], S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 44: v = t_skiploss(i, current, ct, S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 47: v = t_reach_k_rob(i, current, ct, S, Du, O, Cl, Di, ld, D2, tmax, vis, n, 3, scr); break;
        case 48: v = t_reach_k_rob(i, current, ct, S, Du, O, Cl, Di, ld, D2, tmax, vis, n, 5, scr); break;
        case 49: v = t_ns_a2(i, fin, S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 50: v = t_ns_a05(i, fin, S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 53: v = t_ns_rob_m(i, fin, Du[i], S, Du, O, Cl, Di, ld, D2, vis, tmax, n, 0.5); break;
        case 54: v = t_ns_rob_m(i, fin, Du[i], S, Du, O, Cl, Di, ld, D2, vis, tmax, n, 2.0); break;
        case 55: v = t_nnfwd(i, fin, S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 56: v = t_nnk3(i, fin, S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 57: v = t_nnchain(i, current, ct, S, Du, O, Cl, Di, ld, D2, tmax, vis, n, 3, scr); break;
        case 59: v = t_detour(i, current, fin, ta, S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        case 60: v = t_scorerank(i, fin, S, Du, O, Cl, Di, ld, D2, vis, tmax, n); break;
        default: return -1;
        }
        out[k] = v;
    }
    return 0;
}

// The reference's round-robin rollout (engine/sim_core.hpp:241-274) for one
// rule on one scenario row.
double ref_sim_one_rr(int n, int m, double tmax, const double* sc, const double* du,
                      const double* op, const double* cl, const double* dist, const double* d2,
                      const int32_t* ops, const int32_t* args, int len, const double* sampled,
                      const int64_t* mask, const double* zp) {
    SimData sd = make_sd(n, m, tmax, sc, du, op, cl, dist, d2);
    SimBuffers buf(sd.n);
    std::vector<int> cur_v(m + 1);
    std::vector<double> time_v(m + 1);
    std::vector<char> active(m + 1);
    return sim_one_rr(ops, args, len, sd.scores.data(), sd.durations.data(), sd.open_t.data(),
                      sd.close_t.data(), sd.dist.data(), sd.ld, sd.dist_to_depot.data(), sd.tmax,
                      sampled, sd.m, sd.n, mask, zp, (bool*)buf.visited.data(), buf.terms.data(),
                      buf.stack.data(), buf.struct_buf.data(), (bool*)buf.scratch.data(),
                      cur_v.data(), time_v.data(), (bool*)active.data());
}

// gen_durations (engine/sim_host.hpp:50-58) with explicit expected durations.
void ref_gen_durations(int n, const double* expected, double sigma, int n_samples,
                       long long base_seed, double* out) {
    for (int s = 0; s < n_samples; s++) {
        Mt19937 rng((uint64_t)(base_seed + s));
        for (int i = 0; i < n; i++)
            out[(size_t)s * n + i] = sample_duration(expected[i], sigma, rng);
    }
}

// compile_rmod(parse_prefix(s)) into caller buffers; returns node count or -1.
int ref_compile(const char* prefix, int32_t* ops, int32_t* args, int cap) {
    Compiled c = compile_rmod(parse_prefix(prefix));
    if ((int)c.ops.size() > cap) return -1;
    for (size_t k = 0; k < c.ops.size(); k++) {
        ops[k] = c.ops[k]; args[2 * k] = c.args[k][0]; args[2 * k + 1] = c.args[k][1];
    }
    return (int)c.ops.size();
}
}
