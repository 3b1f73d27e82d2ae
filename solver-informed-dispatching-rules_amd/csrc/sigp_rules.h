// Per-rule analysis shared by the host batch builder and the device GP loop:
// which terminal slots a rule's value depends on under a mask, which fused
// scan families produce them, and a relative cost estimate for the launch
// order.  Host and device must agree exactly, so both include this file.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sigp {

constexpr int kTerms = 61;   // terminal slots (engine/vocab.hpp:16)

__host__ __device__ constexpr uint64_t bit(int s) { return 1ull << s; }

// Per-rule computation flags (a second 64-bit word beside the slot bits).
enum : uint64_t {
    F_NS = 1ull << 0,
    F_REGRET = 1ull << 1,
    F_NSROB = 1ull << 2,
    F_MAXN = 1ull << 3,
    F_REACH2 = 1ull << 4,
    F_NSTW = 1ull << 5,
    F_STRUCT = 1ull << 6,
    F_REACHK = 1ull << 7,    // greedy path REACH1/3/5 (length in bits 40-42)
    F_ATRISK = 1ull << 8,
    F_R2ROB = 1ull << 9,
    F_REACHKR = 1ull << 10,  // robust path REACH3ROB/5ROB (length in bits 43-45)
    F_FNOWROB = 1ull << 11,  // per-decision robust bits (NFEASROB)
    F_FSCORE = 1ull << 12,   // per-decision FEASCORE sum
    F_CNOW = 1ull << 13,     // per-decision close-only bits
};

// Slots computed by their own per-candidate routines rather than the fused
// scan families (the kernels' RARE variant).
constexpr uint64_t RARE_SLOTS = bit(44) | bit(49) | bit(50) | bit(53) | bit(54) | bit(55) | bit(56) |
                                bit(57) | bit(59) | bit(60);

// A slot's value is live in the reference when (engine/sim_core.hpp:88-197):
// slots 0-9 always; 11/14/30 when their own or their Z partner's mask bit is
// set (33/32/34); any other slot when its mask bit is set.  A leaf whose slot
// is not live reads 0.0.
__host__ __device__ inline bool slot_live(int s, uint64_t mk) {
    if (s <= 9) return true;
    if (s == 11) return (mk & (bit(11) | bit(33))) != 0;
    if (s == 14) return (mk & (bit(14) | bit(32))) != 0;
    if (s == 30) return (mk & (bit(30) | bit(34))) != 0;
    return (mk & bit(s)) != 0;
}

// Slots whose value the rule reads (rd), the same closed under the Z slots'
// sources (comp), and the fused computations that produce them (fl).
__host__ __device__ inline void compute_sets(uint64_t used, uint64_t mk, uint64_t *rd_out,
                                             uint64_t *comp_out, uint64_t *fl_out) {
    uint64_t c = 0;
    for (int s = 0; s < kTerms; ++s)
        if ((used & bit(s)) && slot_live(s, mk)) c |= bit(s);
    *rd_out = c;
    if (c & bit(32)) c |= bit(14);
    if (c & bit(33)) c |= bit(11);
    if (c & bit(34)) c |= bit(30);
    uint64_t f = 0;
    if (c & (bit(10) | bit(28) | bit(29) | bit(45))) f |= F_NS;
    if (c & (bit(11) | bit(43))) f |= F_REGRET;
    if (c & (bit(12) | bit(45) | bit(46))) f |= F_NSROB;
    if (c & bit(13)) f |= F_MAXN;
    if (c & bit(14)) f |= F_REACH2;
    if (c & bit(15)) f |= F_NSTW;
    if (c & (bit(17) | bit(18) | bit(19) | bit(20) | bit(28) | bit(29))) f |= F_STRUCT;
    if (c & (bit(21) | bit(22) | bit(23))) {
        const uint64_t k = (c & bit(23)) ? 5 : ((c & bit(22)) ? 3 : 1);
        f |= F_REACHK | (k << 40);
    }
    if (c & bit(30)) f |= F_ATRISK | F_CNOW;
    if (c & (bit(41) | bit(43))) f |= F_R2ROB;
    if (c & (bit(47) | bit(48))) {
        const uint64_t k = (c & bit(48)) ? 5 : 3;
        f |= F_REACHKR | (k << 43);
    }
    if (c & bit(51)) f |= F_FSCORE;
    if (c & bit(52)) f |= F_FNOWROB;
    *comp_out = c;
    *fl_out = f;
}

// Relative cost per candidate of a rule of L nodes from the scans its flags
// select (the per-neighbour weights were measured on MI355X); the launch
// order puts expensive rules first so that they do not form its tail.
__host__ __device__ inline float rule_cost(int L, uint64_t f, uint64_t comp) {
    float w = 1.0f + 0.05f * (float)L;
    if (f & (F_NS | F_MAXN | F_REACH2)) w += 1.0f;
    if (f & F_REACH2) w += 1.0f;
    if (f & (F_NSROB | F_R2ROB | F_REACHKR)) w += 1.0f;
    if (f & F_R2ROB) w += 1.0f;
    if (f & F_REGRET) w += 0.5f;
    if (f & F_STRUCT) w += 2.0f;
    if (f & F_NSTW) w += 1.0f;
    if (f & F_ATRISK) w += 8.0f;
    if (f & F_REACHK) w += 1.5f * (float)((f >> 40) & 7);
    if (f & F_REACHKR) w += 1.5f * (float)((f >> 43) & 7);
    if (comp & RARE_SLOTS) w += 4.0f;
    return w;
}

}  // namespace sigp
