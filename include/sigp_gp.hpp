// sigp_gp.hpp -- the reference's generation loop on the device, in the
// reference's own types.
//
// engine/gp_engine.cu:66-84 (evolve_one) packs the population, evaluates it
// on the factory, averages and calls gp_step (engine/gp.hpp:170-194) on the
// host every generation.  DeviceGP runs that loop on the device through the
// C ABI (sigp_gp_* of sigp.h): the same TreeConfig / EvolutionConfig in, the
// same populations out as NodeP trees -- every random choice is the
// reference's Mt19937(seed) draw in the reference's order -- with one HIP
// graph replay per generation.  evolve_one_device is evolve_one's drop-in.
//
// Include after the reference's engine headers (gp.hpp, sim_host.hpp,
// vocab.hpp) and this repository's sigp_factory.hpp, which provides
// FactoryContext.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "gp.hpp"             // TreeConfig, EvolutionConfig, gen_durations' callers (engine/gp.hpp)
#include "sigp_factory.hpp"   // FactoryContext, SIGP_CHECK
#include "sigp.h"

class DeviceGP {
public:
    // seeds: one Mt19937 seed per run (runs evolve side by side, as
    // exhaust_batched's seeds do); train_samples scenarios per run and
    // generation; tmask: the evaluation's terminal mask.
    DeviceGP(FactoryContext &ctx, const TreeConfig &tc, const EvolutionConfig &cfg,
             const std::vector<uint64_t> &seeds, int train_samples, const std::vector<int64_t> &tmask)
        : tc_(tc), runs_((int)seeds.size()), pop_(cfg.population_size), T_(train_samples) {
        for (const std::string &t : tc.terminals) slots_.push_back(t == "ERC" ? -1 : term_index(t));
        for (const std::string &f : tc.functions) ops_.push_back(op_code(f));
        sigp_gp_config c{};
        c.n_runs = runs_;
        c.population = pop_;
        c.offspring = 0;
        c.n_terminals = (int32_t)slots_.size();
        c.terminal_slots = slots_.data();
        c.n_functions = (int32_t)ops_.size();
        c.function_ops = ops_.data();
        c.min_depth = tc.min_depth;
        c.max_depth = tc.max_depth;
        c.depth_limit = cfg.max_depth;
        c.tournament_size = cfg.tournament_size;
        c.elitism = cfg.elitism;
        c.crossover_rate = cfg.crossover_rate;
        c.mutation_rate = cfg.mutation_rate;
        c.T = T_;
        c.launch.waves_per_item = 0;
        c.launch.mapping = SIGP_MAP_AUTO;
        SIGP_CHECK(sigp_gp_create(ctx.inst, &c, seeds.data(), tmask.data(), &gp_));
        SIGP_CHECK(sigp_gp_init(gp_));   // ramped_half_and_half (engine/gp.hpp:67-77)
    }
    DeviceGP(const DeviceGP &) = delete;
    DeviceGP &operator=(const DeviceGP &) = delete;
    ~DeviceGP() { sigp_gp_destroy(gp_); }

    // `generations` evaluations, each followed by gp_step except the last
    // (unless step_last).  rows: generation g's rows for run r start at row
    // (g * runs + r) * train_samples.  Returns the fitness history
    // (generations x runs x population).
    std::vector<double> run(const std::vector<double> &rows, int generations, bool step_last = false) {
        std::vector<double> hist((size_t)generations * runs_ * pop_);
        const int32_t flags = SIGP_GP_GRAPH | (step_last ? SIGP_GP_STEP_LAST : 0);
        SIGP_CHECK(sigp_gp_run(gp_, rows.data(), (int64_t)generations * runs_ * T_, generations, flags,
                               hist.data(), nullptr, nullptr, nullptr));
        return hist;
    }

    // The populations as the reference's trees, one vector per run.
    std::vector<std::vector<NodeP>> population() const {
        std::vector<int16_t> tok((size_t)runs_ * pop_ * SIGP_GP_MAX_NODES);
        std::vector<int32_t> len((size_t)runs_ * pop_);
        SIGP_CHECK(sigp_gp_population(gp_, tok.data(), len.data()));
        std::vector<std::vector<NodeP>> out(runs_);
        for (int r = 0; r < runs_; ++r)
            for (int i = 0; i < pop_; ++i) {
                const int16_t *t = tok.data() + ((size_t)r * pop_ + i) * SIGP_GP_MAX_NODES;
                int at = 0;
                out[r].push_back(tree(t, at));
            }
        return out;
    }

private:
    // tokens (prefix order) -> NodeP: terminal index, SIGP_GP_TOK_FUNC | f,
    // or SIGP_GP_TOK_CONST | k for the ERC label "%.3f" of k / 1000
    // (engine/gp.hpp:37-45 labels it so)
    NodeP tree(const int16_t *t, int &at) const {
        const int v = t[at++];
        if (v & SIGP_GP_TOK_FUNC) {
            NodeP a = tree(t, at);
            NodeP b = tree(t, at);
            return mk_node(tc_.functions[v & 0xff], {a, b});
        }
        if (v & SIGP_GP_TOK_CONST) {
            char buf[16];
            std::snprintf(buf, sizeof buf, "%.3f", (v & (SIGP_GP_TOK_CONST - 1)) / 1000.0);
            return mk_node(buf);
        }
        return mk_node(tc_.terminals[v]);
    }

    TreeConfig tc_;
    int runs_, pop_, T_;
    std::vector<int32_t> slots_, ops_;
    sigp_gp *gp_ = nullptr;
};

// evolve_one (engine/gp_engine.cu:66-84) with the loop on the device: the
// same training draws (gen_durations seeded seed * 10000 + gen), the same
// population after cfg.generations generations.
inline std::vector<NodeP> evolve_one_device(const SimData &sd, const EvolutionConfig &cfg, int seed,
                                            FactoryContext &ctx, const std::vector<int64_t> &tmask,
                                            const TreeConfig &tc) {
    std::vector<double> rows;
    for (int gen = 0; gen < cfg.generations; gen++) {
        const std::vector<double> one = gen_durations(sd, cfg.train_samples, (long long)seed * 10000 + gen);
        rows.insert(rows.end(), one.begin(), one.end());
    }
    DeviceGP gp(ctx, tc, cfg, {(uint64_t)seed}, cfg.train_samples, tmask);
    gp.run(rows, cfg.generations);
    return gp.population()[0];
}
