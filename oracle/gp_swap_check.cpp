// CPU ORACLE side -- test infrastructure only.  The reference's generation
// loop, run with the reference's own operators (engine/gp.hpp:
// ramped_half_and_half, gp_step; Mt19937 of engine/rng.hpp) and evaluated
// through FactoryContext, against evolve_one_device of include/sigp_gp.hpp,
// which runs the whole loop on the device.  Prints how many of the final
// population's trees are identical (prefix strings), and exits 1 unless all.
//     gp_swap_check <instance> [--terms A,B] [--seed s] [--gens g] [--samples t]
// Built by oracle/Makefile (target gpcheck) against the reference's engine
// headers, with engine/factory.cuh -> include/sigp_factory.hpp.
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include "factory.cuh"
#include "gp.hpp"
#include "sigp_gp.hpp"

static std::vector<NodeP> host_loop(const SimData &sd, const EvolutionConfig &cfg, int seed, FactoryContext &ctx,
                                    const std::vector<int64_t> &tmask, const TreeConfig &tc) {
    Mt19937 rng((uint64_t)seed);
    std::vector<NodeP> pop = ramped_half_and_half(tc, rng, cfg.population_size);
    const int T = cfg.train_samples;
    for (int gen = 0; gen < cfg.generations; gen++) {
        const std::vector<double> rows = gen_durations(sd, T, (long long)seed * 10000 + gen);
        const std::vector<double> pairs = ctx.evaluate(pack_trees(pop), rows, T, tmask);
        std::vector<double> fit(pop.size());
        for (size_t p = 0; p < pop.size(); p++) {
            double acc = 0.0;
            for (int s = 0; s < T; s++) acc += pairs[p * T + s];
            fit[p] = acc / T;
        }
        if (gen + 1 < cfg.generations) pop = gp_step(pop, fit, rng, tc, cfg);
    }
    return pop;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: gp_swap_check <instance> [--terms A,B] [--seed s] [--gens g] [--samples t]\n");
        return 2;
    }
    std::string inst = argv[1], terms;
    int seed = 0;
    EvolutionConfig cfg;
    for (int i = 2; i + 1 < argc; i += 2) {
        const std::string a = argv[i];
        if (a == "--terms") terms = argv[i + 1];
        else if (a == "--seed") seed = std::atoi(argv[i + 1]);
        else if (a == "--gens") cfg.generations = std::atoi(argv[i + 1]);
        else if (a == "--samples") cfg.train_samples = std::atoi(argv[i + 1]);
    }
    std::vector<std::string> vocab = NS_TERMINALS;
    std::istringstream ss(terms);
    for (std::string w; std::getline(ss, w, ',');)
        if (!w.empty()) vocab.push_back(w);
    SimData sd(load_instance(inst), 3, 0.2);
    FactoryContext ctx(sd, zp_of(sd));
    const std::vector<int64_t> tmask = term_mask_from(vocab);
    TreeConfig tc;
    tc.terminals = vocab;
    const std::vector<NodeP> a = host_loop(sd, cfg, seed, ctx, tmask, tc);
    const std::vector<NodeP> b = evolve_one_device(sd, cfg, seed, ctx, tmask, tc);
    int same = 0;
    for (size_t i = 0; i < a.size() && i < b.size(); i++) same += to_prefix(a[i]) == to_prefix(b[i]);
    std::printf("%s seed %d, %d generations x %d rules x %d scenarios: %d / %zu trees identical\n", inst.c_str(),
                seed, cfg.generations, cfg.population_size, cfg.train_samples, same, a.size());
    return same == (int)a.size() && a.size() == b.size() ? 0 : 1;
}
