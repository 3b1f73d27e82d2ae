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
