// sigp_factory.hpp -- C++ drop-in for the reference's engine/factory.cuh.
//
// Put this file where engine/factory.cuh was (or include it instead of it),
// next to the reference's own engine headers; callers compile unchanged as
// plain C++ (g++ -x c++ -std=c++20, or hipcc) and link libsigp.so.  It
// re-exposes the same names with the same meaning over the C ABI of sigp.h:
//   PackedTrees, pack_trees        engine/factory.cuh:20-40
//   FactoryContext                 engine/factory.cuh:124-217
//     FactoryContext(sd, zp)       :129-137  (holds a reference to sd, as there)
//     evaluate                     :144-164
//     evaluate_grouped             :167-189
//     evaluate_grouped_masks       :193-216
// out[p*N + s] comes back bit-identical to the reference's CPU simulator.
// Errors print "sigp error ..." to stderr and exit(2), as CUDA_CHECK does
// (engine/factory.cuh:16-18).  The `threads` arguments are accepted and
// ignored: the library chooses its own launch mapping.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "sim_host.hpp"   // SimData (engine/sim_host.hpp:12-40)
#include "vocab.hpp"      // Compiled, compile_rmod (engine/vocab.hpp:96-125)
#include "sigp.h"

#define SIGP_CHECK(x)                                                                  \
    do {                                                                               \
        sigp_status st_ = (x);                                                         \
        if (st_ != SIGP_OK) {                                                          \
            std::fprintf(stderr, "sigp error %d (%s) at %s:%d\n", (int)st_,            \
                         sigp_last_error(), __FILE__, __LINE__);                       \
            std::exit(2);                                                              \
        }                                                                              \
    } while (0)

struct PackedTrees {
    std::vector<int32_t> ops, args;   // args: 2 per node
    std::vector<int32_t> offsets, lengths;
};

inline PackedTrees pack_trees(const std::vector<NodeP> &trees) {
    PackedTrees pk;
    int32_t at = 0;
    for (const NodeP &t : trees) {
        const Compiled c = compile_rmod(t);
        pk.offsets.push_back(at);
        pk.lengths.push_back((int32_t)c.ops.size());
        for (size_t k = 0; k < c.ops.size(); ++k) {
            pk.ops.push_back(c.ops[k]);
            pk.args.push_back(c.args[k][0]);
            pk.args.push_back(c.args[k][1]);
        }
        at += (int32_t)c.ops.size();
    }
    return pk;
}

struct FactoryContext {
    const SimData &sd;
    std::vector<double> zp_host;
    sigp_instance *inst = nullptr;

    FactoryContext(const SimData &sd_, const std::vector<double> &zp) : sd(sd_), zp_host(zp) {
        sigp_instance_desc d{};
        d.n = sd.n;
        d.m = sd.m;
        d.tmax = sd.tmax;
        d.scores = sd.scores.data();
        d.durations = sd.durations.data();
        d.open_t = sd.open_t.data();
        d.close_t = sd.close_t.data();
        d.dist = sd.dist.data();
        d.dist_to_depot = sd.dist_to_depot.data();
        d.zp = zp_host.size() >= 10 ? zp_host.data() : nullptr;
        SIGP_CHECK(sigp_instance_create(&d, &inst));
    }
    ~FactoryContext() { sigp_instance_destroy(inst); }
    FactoryContext(const FactoryContext &) = delete;
    FactoryContext &operator=(const FactoryContext &) = delete;

    // out[p*N+s]; all_sampled = N scenarios x n durations.
    std::vector<double> evaluate(const PackedTrees &pk, const std::vector<double> &all_sampled, int N,
                                 const std::vector<int64_t> &term_mask, int threads = 128) {
        (void)threads;
        const int P = (int)pk.offsets.size();
        return run(pk, all_sampled, N, P > 0 ? P : 1, term_mask, P > 0 ? P : 1);
    }

    // grouped_sampled = G groups x T scenarios x n; tree p -> group p/group_size.
    std::vector<double> evaluate_grouped(const PackedTrees &pk, const std::vector<double> &grouped_sampled,
                                         int G, int T, int group_size,
                                         const std::vector<int64_t> &term_mask, int threads = 128) {
        (void)G;
        (void)threads;
        const int P = (int)pk.offsets.size();
        return run(pk, grouped_sampled, T, group_size, term_mask, P > 0 ? P : 1);
    }

    // masks: one N_TERMS row per block of mask_group trees.
    std::vector<double> evaluate_grouped_masks(const PackedTrees &pk,
                                               const std::vector<double> &grouped_sampled, int T,
                                               int group_size, int mask_group,
                                               const std::vector<int64_t> &masks, int threads = 128) {
        (void)threads;
        return run(pk, grouped_sampled, T, group_size, masks, mask_group);
    }

  private:
    std::vector<double> run(const PackedTrees &pk, const std::vector<double> &sampled, int T,
                            int group_size, const std::vector<int64_t> &masks, int mask_group) {
        sigp_rules r{};
        r.ops = pk.ops.data();
        r.args = pk.args.data();
        r.offsets = pk.offsets.data();
        r.lengths = pk.lengths.data();
        r.n_rules = (int32_t)pk.offsets.size();
        r.n_nodes = (int64_t)pk.ops.size();
        const int64_t n_rows = sd.n > 0 ? (int64_t)(sampled.size() / (size_t)sd.n) : 0;
        std::vector<double> out((size_t)r.n_rules * (size_t)(T > 0 ? T : 0));
        SIGP_CHECK(sigp_evaluate(inst, &r, sampled.data(), n_rows, T, group_size, masks.data(),
                                 (int32_t)(masks.size() / SIGP_N_TERMS), mask_group, out.data(),
                                 nullptr));
        return out;
    }
};
