// The GP loop on the device (engine/gp.hpp, engine/gp_engine.cu:66-84 and
// :131-183): populations, breeding and evaluation stay in HBM, and one
// generation is a fixed sequence of kernels that a HIP graph captures once.
//
// Trees.  Each tree is a prefix-order array of 16-bit tokens in a slot of
// SIGP_GP_MAX_NODES (a binary tree of depth <= 8 has at most 511 nodes):
// a terminal is its index in the run's terminal list, a function is
// TOK_FUNC | its index in the function list, an ERC constant is
// TOK_CONST | k for the value k / 1000.  Every generation the parents are
// analysed by one wavefront per tree: node depth, subtree size and height,
// prefix/suffix maxima of the depth, the pre-order positions of the internal
// nodes and of the leaves -- which makes every question breeding asks of a
// parent (Koza point choice, the subtree it roots, the depth of the child)
// an O(1) lookup -- and compiled to the evaluator's postfix arrays by a
// parallel scatter: node k of pre-order goes to post-order position
// k - depth(k) + size(k) - 1.
//
// Breeding.  The reference draws every random choice of a generation from
// one CPython-compatible Mersenne Twister per run, in a fixed sequential
// order.  One wavefront per run executes that sequence: all 64 lanes run the
// same scalar program in lockstep (identical values, no divergence), lane 0
// stores the results, and when the generator's 624 words are used up the
// lanes regenerate them together in LDS.  The wavefront writes one recipe per
// offspring (copy / crossover of two parents at two points / mutation with a
// freshly grown subtree); a second kernel assembles all offspring in
// parallel.  Draws and outcomes equal engine/gp.hpp step for step, so the
// populations are identical to the reference's (verify/golden/gp.txt).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sigp.h"
#include "sigp_internal.h"
#include "sigp_rules.h"
#include "sigp_mt.h"

using namespace sigp;
using namespace sigp_mt;

#define GP_TRY(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            return sigp_internal_fail(e_ == hipErrorOutOfMemory ? SIGP_OOM : SIGP_DEVICE_ERROR,     \
                                      std::string(#x) + ": " + hipGetErrorString(e_));              \
    } while (0)

namespace {

constexpr int MAXL = SIGP_GP_MAX_NODES;
constexpr int MAX_DEPTH = 8;          // trees of depth <= 8 fit MAXL
constexpr int MAX_TERMS = 128;
constexpr int MAX_FUNCS = 8;
constexpr int MAX_POOL = 512;         // sample()'s pool method: population <= setsize
constexpr int MAX_TOURN = 64;
constexpr int MAX_POP = 4096;         // the breeding wavefront keeps fitness and tree words in LDS
constexpr int TOK_FUNC = SIGP_GP_TOK_FUNC;
constexpr int TOK_CONST = SIGP_GP_TOK_CONST;

struct GpParams {
    int nt, nf;
    int tslot[MAX_TERMS];   // slot of each terminal, -1 = ERC
    int fop[MAX_FUNCS];     // opcode of each function
    int min_depth, max_depth, depth_limit;
    int tourn, elitism;
    int pop;                // parents per run
    int n_off;              // offspring per run and step: subs x pop
    int subs;               // sub-steps per run and step, each a gp_step of pop offspring on its own stream
    int setsize;            // CPython sample(): pool method when pop <= setsize
    double p_cx, p_cxmut;   // crossover_rate, crossover_rate + mutation_rate
    double p_term;          // n_terminals / (n_terminals + n_functions)
};

enum { R_COPY = 0, R_XOVER = 1, R_MUT = 2 };
// Offspring o of a run: parent a (within the run); for crossover the subtree
// of sa nodes at a's point pa is replaced by b's subtree of sb nodes at pb,
// for mutation by the sb tokens grown into the offspring's grow slot.  L is
// the offspring's length.
struct __align__(16) Recipe {
    int kind, a, pa, sa, b, pb, sb, L;
};

// Breeding tables of the parents.  Per tree one word: length | internal
// nodes << 16.  Per tree MAXL point words, the internal nodes first, then
// the leaves, each in pre-order: position << 21 | node word, where the node
// word holds what a crossover or mutation at that node needs --
// subtree size (bits 0-8), depth (9-12), subtree height (13-16) and
// 1 + the depth of the tree outside the subtree (17-20; 0 = nothing outside).
struct Parents {
    const uint32_t *tw;
    const uint32_t *pts;
};
__device__ __forceinline__ int pw_pos(uint32_t w) { return (int)(w >> 21); }
__device__ __forceinline__ int pw_size(uint32_t w) { return (int)(w & 511u); }
__device__ __forceinline__ int pw_depth(uint32_t w) { return (int)((w >> 9) & 15u); }
__device__ __forceinline__ int pw_height(uint32_t w) { return (int)((w >> 13) & 15u); }
__device__ __forceinline__ int pw_outside(uint32_t w) { return (int)((w >> 17) & 15u) - 1; }

// A set of trees and what their analysis writes.  Parents need the breeding
// tables (tw, pts); every evaluated set needs the compiled arrays.
struct TreeSet {
    const int16_t *tok;
    const int32_t *len;
    uint32_t *tw, *pts;                       // breeding tables (null: not a parent set)
    int32_t *raw_ops, *raw_args;              // pack_trees format (engine/factory.cuh:20-40)
    int32_t *ops, *args, *off, *plen;         // prepared for the kernels
    uint64_t *rd, *flags;
    float *cost;
};

// The ERC label "%.3f" % random() read back by compile (llround(v * 1000),
// engine/vocab.hpp:101-104) is k = m * 1000 / 2^53 rounded to nearest, ties
// to even (how printf rounds an exactly representable tie) -- computed
// exactly in 64-bit integers.
__device__ __forceinline__ int erc_thousandths(uint64_t m) {
    const uint64_t x = m * 1000ull;               // < 2^63
    uint64_t q = x >> 53;
    const uint64_t r = x & ((1ull << 53) - 1);
    const uint64_t half = 1ull << 52;
    if (r > half || (r == half && (q & 1))) ++q;
    return (int)q;
}

// engine/gp.hpp:37-45
__device__ __forceinline__ int pick_terminal(const GpParams &P, Rng &g) {
    const int t = (int)g.below((uint32_t)P.nt);
    if (P.tslot[t] < 0) return TOK_CONST | erc_thousandths(g.bits53());
    return t;
}

// engine/gp.hpp:47-64 (generate_full, generate_grow): the recursion as a
// stack of pending depths (lane s holds entry s), which emits nodes and
// makes draws in the same pre-order.  Lane 0 stores the tokens; returns the
// count and the height.
__device__ __forceinline__ int generate(const GpParams &P, Rng &g, int depth, bool full, int16_t *out, int &height) {
    const int lane = (int)threadIdx.x;
    int stk = depth, sp = 1, cnt = 0;
    height = 0;
    while (sp > 0) {
        --sp;
        const int d = __builtin_amdgcn_readlane(stk, sp);
        bool term = d <= 0;
        if (!term && !full) term = g.random() < P.p_term;
        int tk;
        if (term) {
            tk = pick_terminal(P, g);
            height = max(height, depth - d);
        } else {
            tk = TOK_FUNC | (int)g.below((uint32_t)P.nf);
            if (lane == sp || lane == sp + 1) stk = d - 1;
            sp += 2;
        }
        if (lane == 0) out[cnt] = (int16_t)tk;
        ++cnt;
    }
    return cnt;
}

// engine/gp.hpp:67-77: depths ramp min..max, alternating full/grow blocks.
// One wavefront per (run, stream): wavefront r * nsub + s draws pop trees
// from run r's stream s into slots r * per_run + s * pop + i.  The
// population: nsub = 1 (the run's first stream), per_run = pop.  The
// screened initial pool: nsub = subs, per_run = n_off (sigp_gp_init_screened).
__global__ __launch_bounds__(64) void init_kernel(GpParams P, uint32_t *state, int16_t *tok, int32_t *len,
                                                  int nsub, int per_run) {
    __shared__ uint32_t mt[624];
    const int r = (int)blockIdx.x / nsub, sub = (int)blockIdx.x - r * nsub;
    uint32_t *st = state + ((size_t)r * P.subs + sub) * MT_WORDS;
    Rng g;
    load_rng(g, mt, st);
    const int nd = P.max_depth - P.min_depth + 1;
    for (int i = 0; i < P.pop; ++i) {
        const int d = P.min_depth + i % nd;
        const bool full = (i / nd) % 2 == 0;
        int h;
        const size_t gi = (size_t)r * per_run + (size_t)sub * P.pop + i;
        const int n = generate(P, g, d, full, tok + gi * MAXL, h);
        if (threadIdx.x == 0) len[gi] = n;
    }
    store_rng(g, st);
}

// ---------------------------------------------------------------------------
// breeding (engine/gp.hpp:118-194)

// engine/gp.hpp:120-131: Koza 90/10; the draw of the class happens only
// when the tree has internal nodes (it always has leaves).  Returns the
// index into the tree's point list.
__device__ __forceinline__ int select_point(uint32_t tw, Rng &g) {
    const int L = (int)(tw & 0xffffu), ni = (int)(tw >> 16), nl = L - ni;
    if (ni > 0 && (nl == 0 || g.random() < 0.9)) return (int)g.below((uint32_t)ni);
    if (nl > 0) return ni + (int)g.below((uint32_t)nl);
    return (int)g.below((uint32_t)ni);
}

// engine/gp.hpp:157-166 with engine/rng.hpp sample(): k distinct
// contenders, the first maximum in contender order wins.  Contender i is
// held by lane i.
__device__ __forceinline__ int tournament(const GpParams &P, Rng &g, const double *fit, int *pool) {
    const int lane = (int)threadIdx.x;
    const int n = P.pop, k = min(P.tourn, n);
    int c = 0;
    if (n <= P.setsize) {
        for (int j = lane; j < n; j += 64) pool[j] = j;
        __syncthreads();
        for (int i = 0; i < k; ++i) {
            const int j = (int)g.below((uint32_t)(n - i));
            const int pj = pool[j];
            const int last = pool[n - i - 1];
            __syncthreads();
            if (lane == i) c = pj;
            if (lane == 0) pool[j] = last;
            __syncthreads();
        }
    } else {
        for (int i = 0; i < k; ++i) {
            int j = (int)g.below((uint32_t)n);
            while (__ballot(lane < i && c == j) != 0) j = (int)g.below((uint32_t)n);
            if (lane == i) c = j;
        }
    }
    // first maximum in contender order: largest fitness, then lowest lane
    double f = lane < k ? fit[c] : 0.0;
    int who = lane < k ? lane : 64;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double of = __shfl_xor(f, off);
        const int ow = __shfl_xor(who, off);
        if (ow < 64 && (who == 64 || of > f || (of == f && ow < who))) {
            f = of;
            who = ow;
        }
    }
    return __builtin_amdgcn_readlane(c, who);
}

// Lane 0 stores the recipe as two 16-byte vectors (a struct assembled in
// place would go through private scratch memory).
__device__ __forceinline__ void put_recipe(Recipe *rec, int o, int kind, int a, int pa, int sa, int b, int pb,
                                           int sb, int L) {
    if (threadIdx.x == 0) {
        int4 *q = reinterpret_cast<int4 *>(rec + o);
        q[0] = make_int4(kind, a, pa, sa);
        q[1] = make_int4(b, pb, sb, L);
    }
}

// One generation step of every run (engine/gp.hpp:170-194): the elites in
// the order of the reference's stable descending sort, then crossover (two
// children, the second only while room remains), mutation or reproduction by
// rate until pop offspring exist.  One wavefront per (run, sub-step); a
// screened step breeds subs x pop offspring as subs such steps on their own
// generator streams, sub-step 0 being the reference's.
__global__ __launch_bounds__(64) void plan_kernel(GpParams P, uint32_t *state, const double *fitness,
                                                  Parents T, Recipe *recipes, int16_t *grow) {
    __shared__ uint32_t mt[624];
    __shared__ int pool[MAX_POOL];
    extern __shared__ double sh[];
    double *fit = sh;                                        // pop
    uint32_t *tws = (uint32_t *)(sh + P.pop);                // pop: the parents' tree words
    unsigned char *taken = (unsigned char *)(tws + P.pop);   // pop
    const int r = (int)blockIdx.x / P.subs, sub = (int)blockIdx.x - r * P.subs;
    const int lane = (int)threadIdx.x;
    const size_t base = (size_t)r * P.pop;
    const size_t first = (size_t)r * P.n_off + (size_t)sub * P.pop;   // this sub-step's offspring slots
    Recipe *rec = recipes + first;
    for (int j = lane; j < P.pop; j += 64) {
        fit[j] = fitness[base + j];
        tws[j] = T.tw[base + j];
        taken[j] = 0;
    }
    Rng g;
    load_rng(g, mt, state + (size_t)blockIdx.x * MT_WORDS);

    int o = 0;
    const int E = min(P.elitism, P.pop);
    for (int e = 0; e < E; ++e) {
        double bf = 0.0;
        int bi = -1;
        for (int j = lane; j < P.pop; j += 64)
            if (!taken[j] && (bi < 0 || fit[j] > bf)) {
                bf = fit[j];
                bi = j;
            }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double of = __shfl_xor(bf, off);
            const int oi = __shfl_xor(bi, off);
            if (oi >= 0 && (bi < 0 || of > bf || (of == bf && oi < bi))) {
                bf = of;
                bi = oi;
            }
        }
        __syncthreads();
        if (lane == 0) taken[bi] = 1;
        __syncthreads();
        put_recipe(rec, o++, R_COPY, bi, 0, 0, 0, 0, 0, (int)(tws[bi] & 0xffffu));
    }
    while (o < P.pop) {
        const double u = g.random();
        if (u < P.p_cx) {
            const int a = tournament(P, g, fit, pool);
            const int b = tournament(P, g, fit, pool);
            const size_t ga = base + a, gb = base + b;
            const uint32_t ta = tws[a], tb = tws[b];
            const uint32_t wa = T.pts[ga * MAXL + select_point(ta, g)];
            const uint32_t wb = T.pts[gb * MAXL + select_point(tb, g)];
            const int La = (int)(ta & 0xffffu), Lb = (int)(tb & 0xffffu);
            const int pa = pw_pos(wa), pb = pw_pos(wb), sa = pw_size(wa), sb = pw_size(wb);
            const int d1 = max(pw_outside(wa), pw_depth(wa) + pw_height(wb));
            const int d2 = max(pw_outside(wb), pw_depth(wb) + pw_height(wa));
            if (d1 > P.depth_limit) put_recipe(rec, o, R_COPY, a, 0, 0, 0, 0, 0, La);
            else put_recipe(rec, o, R_XOVER, a, pa, sa, b, pb, sb, La - sa + sb);
            ++o;
            if (o < P.pop) {
                if (d2 > P.depth_limit) put_recipe(rec, o, R_COPY, b, 0, 0, 0, 0, 0, Lb);
                else put_recipe(rec, o, R_XOVER, b, pb, sb, a, pa, sa, Lb - sb + sa);
                ++o;
            }
        } else if (u < P.p_cxmut) {
            // engine/gp.hpp:146-153
            const int a = tournament(P, g, fit, pool);
            const size_t ga = base + a;
            const uint32_t ta = tws[a];
            const uint32_t wa = T.pts[ga * MAXL + select_point(ta, g)];
            const int nd = P.min_depth + (int)g.below((uint32_t)(P.max_depth - P.min_depth + 1));
            int h;
            const int glen = generate(P, g, nd, false, grow + (first + o) * MAXL, h);
            const int La = (int)(ta & 0xffffu), pa = pw_pos(wa), sa = pw_size(wa);
            const int d = max(pw_outside(wa), pw_depth(wa) + h);
            if (d > P.max_depth) put_recipe(rec, o, R_COPY, a, 0, 0, 0, 0, 0, La);
            else put_recipe(rec, o, R_MUT, a, pa, sa, 0, 0, glen, La - sa + glen);
            ++o;
        } else {
            const int a = tournament(P, g, fit, pool);
            put_recipe(rec, o, R_COPY, a, 0, 0, 0, 0, 0, (int)(tws[a] & 0xffffu));
            ++o;
        }
    }
    store_rng(g, state + (size_t)blockIdx.x * MT_WORDS);
}

// Offspring from recipes, one wavefront each: the prefix of parent a up to
// the point, the donor subtree, the rest of a.
__global__ __launch_bounds__(64) void assemble_kernel(GpParams P, const Recipe *recipes, const int16_t *ptok,
                                                      const int16_t *grow, int16_t *otok, int32_t *olen) {
    const int r = (int)blockIdx.y, o = (int)blockIdx.x;
    const size_t slot = (size_t)r * P.n_off + o;
    const Recipe q = recipes[slot];
    const int16_t *A = ptok + ((size_t)r * P.pop + q.a) * MAXL;
    int16_t *out = otok + slot * MAXL;
    const int L = q.L;   // <= 511: the depth checks kept every child within depth 8
    int pa = L, sa = 0, pb = 0, sb = 0;
    const int16_t *Bt = A;
    if (q.kind == R_XOVER) {
        pa = q.pa;
        sa = q.sa;
        pb = q.pb;
        sb = q.sb;
        Bt = ptok + ((size_t)r * P.pop + q.b) * MAXL;
    } else if (q.kind == R_MUT) {
        pa = q.pa;
        sa = q.sa;
        sb = q.sb;
        Bt = grow + slot * MAXL;
    }
    for (int k = (int)threadIdx.x; k < L; k += 64)
        out[k] = k < pa ? A[k] : (k < pa + sb ? Bt[pb + k - pa] : A[k - sb + sa]);
    if (threadIdx.x == 0) olen[slot] = L;
}

// The next parents: tree i of run r is offspring sel[r * pop + i] of the
// run (identity when sel is null).
__global__ __launch_bounds__(64) void gather_kernel(GpParams P, const int32_t *sel, const int16_t *otok,
                                                    const int32_t *olen, int16_t *ptok, int32_t *plen) {
    const int r = (int)blockIdx.y, i = (int)blockIdx.x;
    const size_t dst = (size_t)r * P.pop + i;
    const size_t src = (size_t)r * P.n_off + (sel ? sel[dst] : i);
    const int L = olen[src];
    for (int k = (int)threadIdx.x; k < L; k += 64) ptok[dst * MAXL + k] = otok[src * MAXL + k];
    if (threadIdx.x == 0) plen[dst] = L;
}

// Phenotype hash of each offspring: FNV-1a (64-bit) over its ranks, one
// byte per rank in situation-major order.  One wavefront per offspring.
__global__ __launch_bounds__(64) void phash_kernel(const int32_t *ranks, int n_feat, uint64_t *hash) {
    const size_t q = blockIdx.x;
    if (threadIdx.x != 0) return;
    uint64_t h = 0xcbf29ce484222325ull;
    for (int k = 0; k < n_feat; ++k) {
        h ^= (uint64_t)(uint8_t)ranks[q * n_feat + k];
        h *= 0x100000001b3ull;
    }
    hash[q] = h;
}

// An offspring is a phenotypic duplicate when an earlier slot of its run
// (any sub-step, elites included) has the same phenotype hash.
__global__ __launch_bounds__(256) void dup_kernel(GpParams P, const uint64_t *hash, uint8_t *dup) {
    __shared__ uint64_t tile[1024];
    const int r = (int)blockIdx.y;
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const uint64_t *hr = hash + (size_t)r * P.n_off;
    const uint64_t mine = j < P.n_off ? hr[j] : 0;
    const int last = min(P.n_off, (int)((blockIdx.x + 1) * blockDim.x));   // slots before this block's last
    bool d = false;
    for (int b = 0; b < last; b += 1024) {
        for (int q = (int)threadIdx.x; q < 1024; q += (int)blockDim.x) tile[q] = b + q < P.n_off ? hr[b + q] : 0;
        __syncthreads();
        const int m = min(1024, j - b);
        for (int q = 0; q < m; ++q) d |= tile[q] == mine;
        __syncthreads();
    }
    if (j < P.n_off) dup[(size_t)r * P.n_off + j] = d ? 1 : 0;
}

// The next parents of a screened step: run r keeps its E elites (the first
// slots of sub-step 0), then pop - E of the offspring outside the elites'
// slots, in the order: phenotypes not seen at an earlier slot first, then
// higher predicted reward (NaN as -infinity), then lower slot.  One thread
// per slot counts the candidates ordered before it.
__global__ __launch_bounds__(256) void select_kernel(GpParams P, const float *pred, const uint8_t *dup,
                                                     int32_t *sel) {
    __shared__ float tv[1024];
    __shared__ uint8_t td[1024];
    const int r = (int)blockIdx.y;
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int E = min(P.elitism, P.pop), K = P.pop - E;
    const float *pr = pred + (size_t)r * P.n_off;
    const uint8_t *dr = dup + (size_t)r * P.n_off;
    // td: 0 distinct, 1 duplicate, 2 not a candidate (an elite's slot)
    auto cls = [&](int q) { return q % P.pop < E ? (uint8_t)2 : dr[q]; };
    auto key = [&](int q) {
        const float v = pr[q];
        return v != v ? -INFINITY : v;
    };
    const bool cand = j < P.n_off && j % P.pop >= E;
    const uint8_t mc = cand ? cls(j) : 2;
    const float mv = cand ? key(j) : 0.0f;
    int rank = 0;
    for (int b = 0; b < P.n_off; b += 1024) {
        for (int q = (int)threadIdx.x; q < 1024; q += (int)blockDim.x) {
            td[q] = b + q < P.n_off ? cls(b + q) : 2;
            tv[q] = b + q < P.n_off ? key(b + q) : 0.0f;
        }
        __syncthreads();
        const int m = min(1024, P.n_off - b);
        for (int q = 0; q < m; ++q) {
            const uint8_t c = td[q];
            const float v = tv[q];
            rank += (c < mc || (c == mc && (v > mv || (v == mv && b + q < j)))) ? 1 : 0;
        }
        __syncthreads();
    }
    if (cand && rank < K) sel[(size_t)r * P.pop + E + rank] = j;
    if (blockIdx.x == 0)
        for (int e = (int)threadIdx.x; e < E; e += (int)blockDim.x) sel[(size_t)r * P.pop + e] = e;
}

// Each run's best rule of this generation's evaluation (the first maximum
// of the fitness, as engine/gp_engine.cu:107-109 picks it): its tokens and
// length into the history at the generation counter.
__global__ __launch_bounds__(64) void best_kernel(GpParams P, const double *fit, const int16_t *ptok,
                                                  const int32_t *plen, const int *gen, int16_t *best_tok,
                                                  int32_t *best_len) {
    const int r = (int)blockIdx.x, lane = (int)threadIdx.x;
    const size_t base = (size_t)r * P.pop;
    double bf = 0.0;
    int bi = -1;
    for (int j = lane; j < P.pop; j += 64)
        if (bi < 0 || fit[base + j] > bf) {
            bf = fit[base + j];
            bi = j;
        }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double of = __shfl_xor(bf, off);
        const int oi = __shfl_xor(bi, off);
        if (oi >= 0 && (bi < 0 || of > bf || (of == bf && oi < bi))) {
            bf = of;
            bi = oi;
        }
    }
    const size_t slot = (size_t)(*gen) * gridDim.x + r;
    const int L = plen[base + bi];
    for (int k = lane; k < L; k += 64) best_tok[slot * MAXL + k] = ptok[(base + bi) * MAXL + k];
    if (lane == 0) best_len[slot] = L;
}

// Per tree, one wavefront: the breeding tables (when T.tw is set) and the
// compiled postfix arrays (engine/vocab.hpp:96-125 compile_rmod, and the
// batch preparation of the evaluator: dead leaves become the constant 0,
// live ones index the compact value store).
__global__ __launch_bounds__(64) void analyze_kernel(GpParams P, TreeSet T, uint64_t mask_bits) {
    __shared__ int16_t stok[MAXL], ssize[MAXL];
    __shared__ int8_t sdep[MAXL], shei[MAXL], spmax[MAXL], ssmax[MAXL];
    __shared__ int stk_pos[MAX_DEPTH + 4], stk_rem[MAX_DEPTH + 4], stk_h[MAX_DEPTH + 4];
    __shared__ uint64_t s_used;
    __shared__ int s_nint;
    const size_t gi = blockIdx.x;
    const int lane = (int)threadIdx.x;
    const int L = T.len[gi];
    for (int k = lane; k < L; k += 64) stok[k] = T.tok[gi * MAXL + k];
    __syncthreads();
    if (lane == 0) {
        int sp = 0, nint = 0;
        uint64_t used = 0;
        for (int k = 0; k < L; ++k) {
            const int t = stok[k];
            sdep[k] = (int8_t)sp;
            if (t & TOK_FUNC) {
                stk_pos[sp] = k;
                stk_rem[sp] = 2;
                stk_h[sp] = 0;
                ++sp;
                ++nint;
                continue;
            }
            if (!(t & TOK_CONST)) used |= bit(P.tslot[t]);
            ssize[k] = 1;
            shei[k] = 0;
            int h = 0;
            while (sp > 0) {
                const int top = sp - 1;
                stk_h[top] = max(stk_h[top], h + 1);
                if (--stk_rem[top] > 0) break;
                const int q = stk_pos[top];
                ssize[q] = (int16_t)(k - q + 1);
                shei[q] = (int8_t)stk_h[top];
                h = stk_h[top];
                --sp;
            }
        }
        if (T.tw) {
            int m = -1;
            for (int k = 0; k < L; ++k) {
                m = max(m, (int)sdep[k]);
                spmax[k] = (int8_t)m;
            }
            m = -1;
            for (int k = L - 1; k >= 0; --k) {
                m = max(m, (int)sdep[k]);
                ssmax[k] = (int8_t)m;
            }
            T.tw[gi] = (uint32_t)L | ((uint32_t)nint << 16);
        }
        s_used = used;
        s_nint = nint;
    }
    __syncthreads();
    const int nint = s_nint;
    if (T.tw) {
        // point words in pre-order, internal nodes first, then leaves
        int ci = 0, cl = 0;
        const uint64_t below_me = (1ull << lane) - 1;
        for (int b = 0; b < L; b += 64) {
            const int k = b + lane;
            const bool in = k < L;
            const bool isf = in && (stok[k] & TOK_FUNC);
            const uint64_t mf = __ballot(isf), ml = __ballot(in && !isf);
            if (in) {
                const int e = k + ssize[k];
                int out = k > 0 ? spmax[k - 1] : -1;   // depth outside the subtree at k
                if (e < L) out = max(out, (int)ssmax[e]);
                const uint32_t w = (uint32_t)ssize[k] | ((uint32_t)sdep[k] << 9) | ((uint32_t)shei[k] << 13) |
                                   ((uint32_t)(out + 1) << 17) | ((uint32_t)k << 21);
                const int slot = isf ? ci + __popcll(mf & below_me) : nint + cl + __popcll(ml & below_me);
                T.pts[gi * MAXL + slot] = w;
            }
            ci += __popcll(mf);
            cl += __popcll(ml);
        }
    }
    uint64_t rd, comp, fl;
    compute_sets(s_used, mask_bits, &rd, &comp, &fl);
    const size_t o = gi * MAXL;
    for (int k = lane; k < L; k += 64) {
        const int t = stok[k];
        const int post = k - sdep[k] + ssize[k] - 1;
        int rop, ra0, ra1, pop_, pa0;
        if (t & TOK_FUNC) {
            rop = P.fop[t & 0xff];
            ra0 = k - sdep[k] + ssize[k + 1] - 1;   // left child k + 1 in post-order
            ra1 = post - 1;                          // right child ends just before its parent
            pop_ = rop;
            pa0 = ra0;
        } else if (t & TOK_CONST) {
            rop = -2;
            ra0 = t & (TOK_CONST - 1);
            ra1 = 0;
            pop_ = -2;
            pa0 = ra0;
        } else {
            const int s = P.tslot[t];
            rop = -1;
            ra0 = s;
            ra1 = 0;
            if (rd & bit(s)) {
                pop_ = -1;
                pa0 = __popcll(rd & (bit(s) - 1));
            } else {
                pop_ = -2;   // a dead slot reads 0.0
                pa0 = 0;
            }
        }
        if (T.raw_ops) {
            T.raw_ops[o + post] = rop;
            T.raw_args[2 * (o + post)] = ra0;
            T.raw_args[2 * (o + post) + 1] = ra1;
        }
        T.ops[o + post] = pop_;
        T.args[2 * (o + post)] = pa0;
        T.args[2 * (o + post) + 1] = (t & TOK_FUNC) ? ra1 : 0;
    }
    if (lane == 0) {
        T.off[gi] = (int32_t)o;
        T.plen[gi] = L;
        T.rd[gi] = rd;
        T.flags[gi] = fl;
        T.cost[gi] = rule_cost(L, fl, comp);
    }
}

// Launch order: rules by cost, descending, ties by index (a stable sort),
// by counting for each rule the rules that precede it.
__global__ __launch_bounds__(256) void order_kernel(const float *cost, int P, int32_t *order) {
    __shared__ float tile[1024];
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const float ci = i < P ? cost[i] : 0.0f;
    int rank = 0;
    for (int b = 0; b < P; b += 1024) {
        for (int j = (int)threadIdx.x; j < 1024; j += (int)blockDim.x) tile[j] = b + j < P ? cost[b + j] : 0.0f;
        __syncthreads();
        const int m = min(1024, P - b);
        for (int j = 0; j < m; ++j) {
            const float cj = tile[j];
            rank += (cj > ci || (cj == ci && b + j < i)) ? 1 : 0;
        }
        __syncthreads();
    }
    if (i < P) order[rank] = i;
}

// Repeated trees.  Trees of one run with the same tokens simulate to the
// same rewards on the run's scenarios (the simulation is deterministic), so
// each is simulated once: tree_hash_kernel hashes every parent (FNV-1a of
// its length and tokens), tree_rep_kernel gives each the first tree of its
// run with the same tokens (compared in full, so a hash collision cannot
// merge two trees) and a launch-order cost of -1 if that is another tree
// (repeats go last in the order and their workgroups exit at once), and
// dup_fill_kernel copies the rewards and fitness of that first tree after
// the evaluation.
__global__ __launch_bounds__(256) void tree_hash_kernel(const int16_t *tok, const int32_t *len, int P,
                                                        uint64_t *h) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= P) return;
    const int L = len[i];
    const int16_t *t = tok + (size_t)i * MAXL;
    uint64_t x = 1469598103934665603ull;
    x = (x ^ (uint64_t)(uint32_t)L) * 1099511628211ull;
    for (int k = 0; k < L; ++k) x = (x ^ (uint64_t)(uint16_t)t[k]) * 1099511628211ull;
    h[i] = x;
}

__global__ __launch_bounds__(256) void tree_rep_kernel(const int16_t *tok, const int32_t *len, const uint64_t *h,
                                                       const float *cost, int P, int pop, int32_t *rep,
                                                       float *dcost) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= P) return;
    const int L = len[i];
    const uint64_t hi = h[i];
    const int16_t *a = tok + (size_t)i * MAXL;
    int r = i;
    for (int j = i - i % pop; j < i && r == i; ++j) {
        if (h[j] != hi || len[j] != L) continue;
        const int16_t *b = tok + (size_t)j * MAXL;
        int k = 0;
        while (k < L && a[k] == b[k]) ++k;
        if (k == L) r = j;
    }
    rep[i] = r;
    dcost[i] = r == i ? cost[i] : -1.0f;
}

__global__ __launch_bounds__(256) void dup_fill_kernel(const int32_t *rep, int T, double *fit, double *pairs) {
    const int i = (int)blockIdx.x;
    const int r = rep[i];
    if (r == i) return;
    for (int s = (int)threadIdx.x; s < T; s += (int)blockDim.x) pairs[(size_t)i * T + s] = pairs[(size_t)r * T + s];
    if (threadIdx.x == 0) fit[i] = fit[r];
}

// Generation counter of a captured loop and the per-generation copies that
// read it (a graph replays with fixed arguments).
__global__ void rows_kernel(const double *all, const int *gen, long long per_gen, double *dst) {
    const long long g = *gen;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < per_gen;
         k += (long long)gridDim.x * blockDim.x)
        dst[k] = all[g * per_gen + k];
}

__global__ void record_kernel(const double *fit, const int *gen, int P, double *hist) {
    const long long g = *gen;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < P) hist[g * P + i] = fit[i];
}

__global__ void tick_kernel(int *gen) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *gen += 1;
}

__global__ void set_kernel(int *gen, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *gen = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// host object

struct sigp_gp {
    sigp_instance *inst = nullptr;
    GpParams prm{};
    int runs = 0, T = 0, n = 0;
    uint64_t mask = 0;
    int P = 0;                        // runs x pop
    char *dbuf = nullptr;             // every device array below
    uint32_t *mt = nullptr;           // runs x subs x 625
    int16_t *ptok = nullptr, *otok = nullptr, *grow = nullptr;
    int32_t *plen = nullptr, *olen = nullptr;
    Recipe *rec = nullptr;
    TreeSet pset{};                   // the parents: tables + compiled arrays
    TreeSet oset{};                   // screened steps: the offspring's compiled arrays
    int32_t *order = nullptr;
    uint64_t *thash = nullptr;        // repeated trees: token hashes,
    int32_t *rep = nullptr;           // the first tree with the same tokens,
    float *dcost = nullptr;           // and the launch-order cost (-1: a repeat)
    double *fit = nullptr, *pairs = nullptr, *scen = nullptr;
    int *gen = nullptr;
    int32_t *ranks = nullptr, *sel = nullptr;   // screened steps: phenotypes, next parents
    uint64_t *phash = nullptr;                  // screened steps: phenotype hashes
    uint8_t *dup = nullptr;                     // screened steps: phenotypic duplicates
    float *pred = nullptr;                      // screened steps: predicted rewards
    sigp_batch *bt = nullptr;         // the parents' evaluation
    sigp_batch *obt = nullptr;        // the offspring's phenotype
    sigp_batch *pbt = nullptr;        // the parents' phenotype (sigp_gp_phenotype; made on first use)
    int KL = 1, SL = 1;               // the batches' value-store and stack rows
    bool rmod = false, rare = false;  // and their kernel variants
    sigp_mlp *net = nullptr;          // the screen (not owned)
    char *sbuf = nullptr;             // the decision situations
    sigp_device_situations sit{};
    bool sampling = false;            // rows sampled on the device (sigp_gp_set_sampler)
    sigp_sampler smp{};
    long long *run_seed = nullptr;    // device: n_runs
    long long gen_stride = 0, gen_offset = 0;
    hipStream_t stream = nullptr;
    bool ready = false;               // a population exists
};

static sigp_status gp_fail(const std::string &m) { return sigp_internal_fail(SIGP_INVALID_ARG, m); }

static Parents parents_of(const sigp_gp *g) {
    Parents T;
    T.tw = g->pset.tw;
    T.pts = g->pset.pts;
    return T;
}

// Analysis, compilation and launch order of the parents.
static void launch_rebuild(sigp_gp *g) {
    analyze_kernel<<<g->P, 64, 0, g->stream>>>(g->prm, g->pset, g->mask);
    const unsigned nb = (unsigned)((g->P + 255) / 256);
    tree_hash_kernel<<<nb, 256, 0, g->stream>>>(g->ptok, g->plen, g->P, g->thash);
    tree_rep_kernel<<<nb, 256, 0, g->stream>>>(g->ptok, g->plen, g->thash, g->pset.cost, g->P, g->prm.pop, g->rep,
                                               g->dcost);
    order_kernel<<<nb, 256, 0, g->stream>>>(g->dcost, g->P, g->order);
}

// The parents' evaluation on the generation's rows: every distinct tree
// simulated, then its results copied to its repeats.
static sigp_status launch_evaluation(sigp_gp *g) {
    const sigp_status st = sigp_batch_run(g->bt, g->scen, (int64_t)g->runs * g->T, g->pairs, g->fit, g->stream);
    if (st != SIGP_OK) return st;
    dup_fill_kernel<<<g->P, 64, 0, g->stream>>>(g->rep, g->T, g->fit, g->pairs);
    GP_TRY(hipGetLastError());
    return SIGP_OK;
}

// One step of every run: breeding, then (screened) the offspring's
// compilation, phenotypes, predicted rewards and the choice of the next
// parents, then the new parents' analysis.  A loop created for screening
// with no network attached takes sub-step 0's offspring, which is the plain
// gp_step on the run's own stream.
// The screen over the offspring slots: compilation, phenotypes, predicted
// rewards, phenotypic duplicates and the choice of the next parents (the
// first min(elitism, pop) slots of sub-step 0 kept as elites), into g->sel.
static sigp_status launch_screen(sigp_gp *g, const GpParams &p) {
    analyze_kernel<<<g->runs * p.n_off, 64, 0, g->stream>>>(p, g->oset, g->mask);
    sigp_status st = sigp_internal_phenotype_device(g->obt, g->sit, g->ranks, g->stream);
    if (st != SIGP_OK) return st;
    st = sigp_internal_mlp_predict_ranks(g->net, g->ranks, (int64_t)g->runs * p.n_off, g->sit.n_cand, g->pred,
                                         g->stream);
    if (st != SIGP_OK) return st;
    phash_kernel<<<g->runs * p.n_off, 64, 0, g->stream>>>(g->ranks, SIGP_MLP_IN, g->phash);
    dup_kernel<<<dim3((p.n_off + 255) / 256, g->runs), 256, 0, g->stream>>>(p, g->phash, g->dup);
    select_kernel<<<dim3((p.n_off + 255) / 256, g->runs), 256, 0, g->stream>>>(p, g->pred, g->dup, g->sel);
    GP_TRY(hipGetLastError());
    return SIGP_OK;
}

static sigp_status launch_step(sigp_gp *g) {
    const GpParams &p = g->prm;
    const size_t lds = (size_t)p.pop * (sizeof(double) + 4 + 1);
    plan_kernel<<<g->runs * p.subs, 64, lds, g->stream>>>(p, g->mt, g->fit, parents_of(g), g->rec, g->grow);
    assemble_kernel<<<dim3(p.n_off, g->runs), 64, 0, g->stream>>>(p, g->rec, g->ptok, g->grow, g->otok, g->olen);
    const int32_t *sel = nullptr;
    if (p.subs > 1 && g->net) {
        const sigp_status st = launch_screen(g, p);
        if (st != SIGP_OK) return st;
        sel = g->sel;
    }
    gather_kernel<<<dim3(p.pop, g->runs), 64, 0, g->stream>>>(p, sel, g->otok, g->olen, g->ptok, g->plen);
    launch_rebuild(g);
    GP_TRY(hipGetLastError());
    return SIGP_OK;
}

static sigp_status make_rules_batch(sigp_gp *g, const TreeSet &S, int P, const int32_t *order, int KL, int SL,
                                    bool rmod, bool rare, int T, int group_size, const sigp_launch_config &cfg,
                                    sigp_batch **out, const int32_t *rep = nullptr) {
    sigp_device_rules R;
    R.ops = S.ops;
    R.args = S.args;
    R.off = S.off;
    R.len = S.plen;
    R.rd = S.rd;
    R.flags = S.flags;
    R.order = order;
    R.rep = rep;
    R.P = P;
    R.KL = KL;
    R.SL = SL;
    R.has_rmod = rmod;
    R.has_rare = rare;
    return sigp_internal_batch_device(g->inst, R, T, group_size, cfg, out);
}

extern "C" sigp_status sigp_gp_create(sigp_instance *inst, const sigp_gp_config *c, const uint64_t *seeds,
                                      const int64_t *term_mask, sigp_gp **out) {
    if (!out) return gp_fail("out is NULL");
    *out = nullptr;
    if (!inst || !c || !seeds || !term_mask) return gp_fail("NULL argument");
    if (c->n_runs < 1 || c->n_runs > 65535) return gp_fail("n_runs must be in 1..65535");
    if (c->population < 1 || c->population > MAX_POP)
        return gp_fail("population must be in 1.." + std::to_string(MAX_POP));
    const int n_off = c->offspring == 0 ? c->population : c->offspring;
    if (n_off < c->population || n_off % c->population != 0 || n_off / c->population > 64)
        return gp_fail("offspring must be 0 or a multiple (1..64) of the population");
    if (c->n_terminals < 1 || c->n_terminals > MAX_TERMS || !c->terminal_slots)
        return gp_fail("need 1.." + std::to_string(MAX_TERMS) + " terminals");
    if (c->n_functions < 1 || c->n_functions > MAX_FUNCS || !c->function_ops)
        return gp_fail("need 1.." + std::to_string(MAX_FUNCS) + " functions");
    if (c->min_depth < 0 || c->min_depth > c->max_depth || c->max_depth > MAX_DEPTH || c->depth_limit < 0 ||
        c->depth_limit > MAX_DEPTH)
        return gp_fail("need 0 <= min_depth <= max_depth <= 8 and 0 <= depth_limit <= 8");
    if (c->tournament_size < 1 || c->tournament_size > MAX_TOURN) return gp_fail("tournament_size must be in 1..64");
    if (c->elitism < 0) return gp_fail("elitism < 0");
    if (c->T < 1) return gp_fail("T must be >= 1");
    if ((long long)c->n_runs * n_off > (1 << 20)) return gp_fail("n_runs x offspring exceeds 2^20");

    GpParams p{};
    p.nt = c->n_terminals;
    p.nf = c->n_functions;
    uint64_t vocab = 0;
    for (int t = 0; t < p.nt; ++t) {
        const int s = c->terminal_slots[t];
        if (s < -1 || s >= SIGP_N_TERMS) return gp_fail("terminal slot out of range");
        p.tslot[t] = s;
        if (s >= 0) vocab |= bit(s);
    }
    bool rmod = false;
    for (int f = 0; f < p.nf; ++f) {
        const int op = c->function_ops[f];
        if (op < 0 || op > 6) return gp_fail("function opcode out of range");
        p.fop[f] = op;
        rmod |= op == 6;
    }
    p.min_depth = c->min_depth;
    p.max_depth = c->max_depth;
    p.depth_limit = c->depth_limit;
    p.tourn = c->tournament_size;
    p.elitism = c->elitism;
    p.pop = c->population;
    p.n_off = n_off;
    p.subs = n_off / c->population;
    // CPython sample(): setsize = 21 (+ 4 ** ceil(log(3k, 4)) for k > 5)
    const int k = std::min(c->tournament_size, c->population);
    int setsize = 21;
    if (k > 5) setsize += (int)std::pow(4.0, std::ceil(std::log((double)(k * 3)) / std::log(4.0)));
    if (setsize > MAX_POOL) return gp_fail("tournament too large for the sampling pool");
    p.setsize = setsize;
    p.p_cx = c->crossover_rate;
    p.p_cxmut = c->crossover_rate + c->mutation_rate;
    p.p_term = (double)p.nt / ((double)p.nt + (double)p.nf);

    uint64_t mk = 0;
    for (int s = 0; s < SIGP_N_TERMS; ++s)
        if (term_mask[s]) mk |= bit(s);
    uint64_t rd_all, comp_all, fl_all;
    compute_sets(vocab, mk, &rd_all, &comp_all, &fl_all);

    sigp_gp *g = new sigp_gp();
    g->inst = inst;
    g->prm = p;
    g->runs = c->n_runs;
    g->T = c->T;
    g->n = sigp_internal_instance_n(inst);
    g->mask = mk;
    g->P = c->n_runs * c->population;
    const bool screened = p.subs > 1;
    const size_t P = (size_t)g->P, O = (size_t)c->n_runs * n_off, NL = P * MAXL, OL = O * MAXL;
    const size_t OC = screened ? O : 0, OCL = OC * MAXL;   // the offspring's compiled arrays
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 8 + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_mt = take((size_t)g->runs * p.subs * MT_WORDS * 4), o_ptok = take(NL * 2),
                 o_otok = take(OL * 2), o_grow = take(OL * 2), o_plen = take(P * 4), o_olen = take(O * 4),
                 o_rec = take(O * sizeof(Recipe)), o_tw = take(P * 4), o_pts = take(NL * 4),
                 o_rops = take(NL * 4), o_rargs = take(NL * 8), o_ops = take(NL * 4), o_args = take(NL * 8),
                 o_off = take(P * 4), o_len = take(P * 4), o_rd = take(P * 8), o_fl = take(P * 8),
                 o_cost = take(P * 4), o_ord = take(P * 4), o_fit = take(P * 8),
                 o_pairs = take(P * (size_t)g->T * 8), o_scen = take((size_t)g->runs * g->T * g->n * 8),
                 o_gen = take(4), o_oops = take(OCL * 4), o_oargs = take(OCL * 8), o_ooff = take(OC * 4),
                 o_olen2 = take(OC * 4), o_ord2 = take(OC * 8), o_ofl = take(OC * 8), o_ocost = take(OC * 4),
                 o_ranks = take(OC * SIGP_MLP_IN * 4), o_pred = take(OC * 4), o_sel = take(screened ? P * 4 : 0),
                 o_hash = take(OC * 8), o_dup = take(OC), o_thash = take(P * 8), o_rep = take(P * 4),
                 o_dcost = take(P * 4);
    if (hipMalloc((void **)&g->dbuf, off) != hipSuccess) {
        delete g;
        return sigp_internal_fail(SIGP_OOM, "GP allocation failed");
    }
    char *b = g->dbuf;
    g->mt = (uint32_t *)(b + o_mt);
    g->ptok = (int16_t *)(b + o_ptok);
    g->otok = (int16_t *)(b + o_otok);
    g->grow = (int16_t *)(b + o_grow);
    g->plen = (int32_t *)(b + o_plen);
    g->olen = (int32_t *)(b + o_olen);
    g->rec = (Recipe *)(b + o_rec);
    TreeSet &S = g->pset;
    S.tok = g->ptok;
    S.len = g->plen;
    S.tw = (uint32_t *)(b + o_tw);
    S.pts = (uint32_t *)(b + o_pts);
    S.raw_ops = (int32_t *)(b + o_rops);
    S.raw_args = (int32_t *)(b + o_rargs);
    S.ops = (int32_t *)(b + o_ops);
    S.args = (int32_t *)(b + o_args);
    S.off = (int32_t *)(b + o_off);
    S.plen = (int32_t *)(b + o_len);
    S.rd = (uint64_t *)(b + o_rd);
    S.flags = (uint64_t *)(b + o_fl);
    S.cost = (float *)(b + o_cost);
    TreeSet &Q = g->oset;
    Q.tok = g->otok;
    Q.len = g->olen;
    Q.ops = (int32_t *)(b + o_oops);
    Q.args = (int32_t *)(b + o_oargs);
    Q.off = (int32_t *)(b + o_ooff);
    Q.plen = (int32_t *)(b + o_olen2);
    Q.rd = (uint64_t *)(b + o_ord2);
    Q.flags = (uint64_t *)(b + o_ofl);
    Q.cost = (float *)(b + o_ocost);
    g->order = (int32_t *)(b + o_ord);
    g->thash = (uint64_t *)(b + o_thash);
    g->rep = (int32_t *)(b + o_rep);
    g->dcost = (float *)(b + o_dcost);
    g->fit = (double *)(b + o_fit);
    g->pairs = (double *)(b + o_pairs);
    g->scen = (double *)(b + o_scen);
    g->gen = (int *)(b + o_gen);
    g->ranks = (int32_t *)(b + o_ranks);
    g->pred = (float *)(b + o_pred);
    g->sel = (int32_t *)(b + o_sel);
    g->phash = (uint64_t *)(b + o_hash);
    g->dup = (uint8_t *)(b + o_dup);

    // run r's sub-step s draws from Mt19937(seeds[r] + s * 2^32)
    std::vector<uint32_t> st((size_t)g->runs * p.subs * MT_WORDS);
    for (int r = 0; r < g->runs; ++r)
        for (int q = 0; q < p.subs; ++q)
            mt_seed(seeds[r] + ((uint64_t)q << 32), st.data() + ((size_t)r * p.subs + q) * MT_WORDS);
    sigp_status status = SIGP_OK;
    if (hipMemcpy(g->mt, st.data(), st.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
        status = sigp_internal_fail(SIGP_DEVICE_ERROR, "GP set-up failed");
    const int KL = std::max(1, __builtin_popcountll(rd_all));
    const int SL = std::max(1, std::max(c->max_depth, c->depth_limit));
    const bool rare = (rd_all & RARE_SLOTS) != 0;
    // AUTO for the population's evaluation: two simulations per wavefront up
    // to n = 128 with two or more scenarios per rule (evolved populations at
    // n = 100 measured 6-11 % faster over a whole loop than one per
    // wavefront), the batch's own choice otherwise
    sigp_launch_config lc = c->launch;
    if (lc.mapping == SIGP_MAP_AUTO && g->n <= 128 && g->T >= 2) {
        lc.mapping = SIGP_MAP_CANDIDATES_2;
        lc.waves_per_item = 0;
    }
    g->KL = KL;
    g->SL = SL;
    g->rmod = rmod;
    g->rare = rare;
    if (status == SIGP_OK)
        status = make_rules_batch(g, S, g->P, g->order, KL, SL, rmod, rare, g->T, c->population, lc, &g->bt, g->rep);
    if (status == SIGP_OK && screened)
        status = make_rules_batch(g, Q, (int)O, nullptr, KL, SL, rmod, rare, 1, (int)O, c->launch, &g->obt);
    if (status != SIGP_OK) {
        sigp_gp_destroy(g);
        return status;
    }
    *out = g;
    return SIGP_OK;
}

// The decision situations in one device allocation (*buf, freed by the
// caller), checked against the instance.
static sigp_status upload_situations(const sigp_gp *g, const sigp_situations *sit, char **buf,
                                     sigp_device_situations *ds) {
    const int S = sit->n_sit, C = sit->n_cand, n = g->n;
    if (S < 1 || C < 2 || (long long)S * C != SIGP_MLP_IN)
        return gp_fail("need n_sit x n_cand = " + std::to_string(SIGP_MLP_IN) + " and n_cand >= 2");
    if (!sit->cur || !sit->ct || !sit->remaining || !sit->visited || !sit->cand)
        return gp_fail("situation arrays must be non-NULL");
    std::vector<uint32_t> words((size_t)S * 16, 0u);
    for (int q = 0; q < S; ++q) {
        if (sit->cur[q] < 0 || sit->cur[q] > n) return gp_fail("situation cur out of range");
        if (sit->remaining[q] < 0) return gp_fail("situation remaining < 0");
        for (int i = 0; i < n; ++i)
            if (sit->visited[(size_t)q * n + i]) words[(size_t)q * 16 + (i >> 5)] |= 1u << (i & 31);
        for (int j = 0; j < C; ++j)
            if (sit->cand[(size_t)q * C + j] < 0 || sit->cand[(size_t)q * C + j] >= n)
                return gp_fail("candidate out of range");
    }
    const size_t o_cur = 0, o_rem = 256 * ((S * 4 + 255) / 256), o_cand = o_rem + 256 * ((S * 4 + 255) / 256),
                 o_ct = o_cand + 256 * (((size_t)S * C * 4 + 255) / 256), o_vis = o_ct + 256 * ((S * 8 + 255) / 256),
                 total = o_vis + words.size() * 4;
    char *b = nullptr;
    GP_TRY(hipMalloc((void **)&b, total));
    *buf = b;
    GP_TRY(hipMemcpy(b + o_cur, sit->cur, S * 4, hipMemcpyHostToDevice));
    GP_TRY(hipMemcpy(b + o_rem, sit->remaining, S * 4, hipMemcpyHostToDevice));
    GP_TRY(hipMemcpy(b + o_cand, sit->cand, (size_t)S * C * 4, hipMemcpyHostToDevice));
    GP_TRY(hipMemcpy(b + o_ct, sit->ct, S * 8, hipMemcpyHostToDevice));
    GP_TRY(hipMemcpy(b + o_vis, words.data(), words.size() * 4, hipMemcpyHostToDevice));
    ds->n_sit = S;
    ds->n_cand = C;
    ds->cur = (const int32_t *)(b + o_cur);
    ds->rem = (const int32_t *)(b + o_rem);
    ds->cand = (const int32_t *)(b + o_cand);
    ds->ct = (const double *)(b + o_ct);
    ds->vis = (const uint32_t *)(b + o_vis);
    return SIGP_OK;
}

extern "C" sigp_status sigp_gp_set_surrogate(sigp_gp *g, sigp_mlp *net, const sigp_situations *sit) {
    if (!g) return gp_fail("NULL argument");
    if (g->prm.subs < 2) return gp_fail("a screen needs offspring > population");
    GP_TRY(hipStreamSynchronize(g->stream));
    if (!net) {                       // detach: the next steps are plain gp_steps
        g->net = nullptr;
        return SIGP_OK;
    }
    if (!sit) return gp_fail("NULL argument");
    if (g->sbuf) {
        (void)hipFree(g->sbuf);
        g->sbuf = nullptr;
    }
    g->net = nullptr;
    const sigp_status st = upload_situations(g, sit, &g->sbuf, &g->sit);
    if (st != SIGP_OK) return st;
    g->net = net;
    return SIGP_OK;
}

extern "C" sigp_status sigp_gp_phenotype(sigp_gp *g, const sigp_situations *sit, int32_t *ranks) {
    if (!g || !sit || !ranks) return gp_fail("NULL argument");
    if (!g->ready) return gp_fail("no population: call sigp_gp_init first");
    if ((long long)sit->n_sit * sit->n_cand != SIGP_MLP_IN)
        return gp_fail("need n_sit x n_cand = " + std::to_string(SIGP_MLP_IN));
    sigp_status st = SIGP_OK;
    if (!g->pbt &&
        (st = make_rules_batch(g, g->pset, g->P, nullptr, g->KL, g->SL, g->rmod, g->rare, 1, g->P,
                               sigp_launch_config{}, &g->pbt)) != SIGP_OK)
        return st;
    char *sb = nullptr, *out = nullptr;
    sigp_device_situations ds{};
    const size_t bytes = (size_t)g->P * SIGP_MLP_IN * 4;
    st = upload_situations(g, sit, &sb, &ds);
    if (st == SIGP_OK && hipMalloc((void **)&out, bytes) != hipSuccess)
        st = sigp_internal_fail(SIGP_OOM, "phenotype output allocation failed");
    if (st == SIGP_OK) st = sigp_internal_phenotype_device(g->pbt, ds, (int32_t *)out, g->stream);
    if (st == SIGP_OK && (hipMemcpyAsync(ranks, out, bytes, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
                          hipStreamSynchronize(g->stream) != hipSuccess))
        st = sigp_internal_fail(SIGP_DEVICE_ERROR, "phenotype failed");
    (void)hipStreamSynchronize(g->stream);
    if (out) (void)hipFree(out);
    if (sb) (void)hipFree(sb);
    return st;
}

extern "C" sigp_status sigp_gp_destroy(sigp_gp *g) {
    if (!g) return SIGP_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->bt) sigp_batch_destroy(g->bt);
    if (g->obt) sigp_batch_destroy(g->obt);
    if (g->pbt) sigp_batch_destroy(g->pbt);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    if (g->dbuf) (void)hipFree(g->dbuf);
    if (g->sbuf) (void)hipFree(g->sbuf);
    if (g->run_seed) (void)hipFree(g->run_seed);
    delete g;
    return SIGP_OK;
}

extern "C" sigp_status sigp_gp_init(sigp_gp *g) {
    if (!g) return gp_fail("gp is NULL");
    // ramped half-and-half from each run's first stream
    init_kernel<<<g->runs, 64, 0, g->stream>>>(g->prm, g->mt, g->ptok, g->plen, 1, g->prm.pop);
    launch_rebuild(g);
    GP_TRY(hipGetLastError());
    GP_TRY(hipStreamSynchronize(g->stream));
    g->ready = true;
    return SIGP_OK;
}

extern "C" sigp_status sigp_gp_init_screened(sigp_gp *g) {
    if (!g) return gp_fail("gp is NULL");
    if (g->prm.subs < 2 || !g->net) return gp_fail("a screened initial population needs a surrogate (sigp_gp_set_surrogate)");
    // subs x pop ramped half-and-half trees per run, pop from each of its
    // streams (stream 0's are the plain initial population), then the
    // screen of a step with no elites
    init_kernel<<<g->runs * g->prm.subs, 64, 0, g->stream>>>(g->prm, g->mt, g->otok, g->olen, g->prm.subs,
                                                             g->prm.n_off);
    GpParams p = g->prm;
    p.elitism = 0;
    sigp_status st = launch_screen(g, p);
    if (st != SIGP_OK) return st;
    gather_kernel<<<dim3(p.pop, g->runs), 64, 0, g->stream>>>(p, g->sel, g->otok, g->olen, g->ptok, g->plen);
    launch_rebuild(g);
    GP_TRY(hipGetLastError());
    GP_TRY(hipStreamSynchronize(g->stream));
    g->ready = true;
    return SIGP_OK;
}

static sigp_status check_step(const sigp_gp *g) {
    if (!g->ready) return gp_fail("no population: call sigp_gp_init first");
    return SIGP_OK;
}

extern "C" sigp_status sigp_gp_evaluate(sigp_gp *g, const double *sampled, int64_t n_rows, double *fitness) {
    if (!g || !sampled) return gp_fail("NULL argument");
    if (!g->ready) return gp_fail("no population: call sigp_gp_init first");
    if (n_rows != (int64_t)g->runs * g->T)
        return gp_fail("need n_runs x T = " + std::to_string((long long)g->runs * g->T) + " scenario rows");
    GP_TRY(hipMemcpyAsync(g->scen, sampled, (size_t)n_rows * g->n * 8, hipMemcpyHostToDevice, g->stream));
    sigp_status st = launch_evaluation(g);
    if (st != SIGP_OK) return st;
    if (fitness) GP_TRY(hipMemcpyAsync(fitness, g->fit, (size_t)g->P * 8, hipMemcpyDeviceToHost, g->stream));
    GP_TRY(hipStreamSynchronize(g->stream));
    return SIGP_OK;
}

extern "C" sigp_status sigp_gp_step(sigp_gp *g, const double *fitness) {
    if (!g) return gp_fail("gp is NULL");
    sigp_status st = check_step(g);
    if (st != SIGP_OK) return st;
    if (fitness) GP_TRY(hipMemcpyAsync(g->fit, fitness, (size_t)g->P * 8, hipMemcpyHostToDevice, g->stream));
    if ((st = launch_step(g)) != SIGP_OK) return st;
    GP_TRY(hipStreamSynchronize(g->stream));
    return SIGP_OK;
}

// One generation of the loop on the stream: this generation's rows, the
// evaluation, its records, and (unless last) the step to the next.
static sigp_status launch_generation(sigp_gp *g, const double *all, double *hist, int16_t *best_tok,
                                     int32_t *best_len, bool step) {
    const long long per_gen = (long long)g->runs * g->T * g->n;
    sigp_status st = SIGP_OK;
    if (all)
        rows_kernel<<<(unsigned)std::min<long long>((per_gen + 255) / 256, 1024), 256, 0, g->stream>>>(
            all, g->gen, per_gen, g->scen);
    else if ((st = sigp_internal_sample(g->inst, &g->smp, g->run_seed, 0, g->T, g->gen_stride, g->gen_offset, g->gen,
                                        g->runs * g->T, g->scen, g->stream)) != SIGP_OK)
        return st;
    st = launch_evaluation(g);
    if (st != SIGP_OK) return st;
    record_kernel<<<(g->P + 255) / 256, 256, 0, g->stream>>>(g->fit, g->gen, g->P, hist);
    best_kernel<<<g->runs, 64, 0, g->stream>>>(g->prm, g->fit, g->ptok, g->plen, g->gen, best_tok, best_len);
    if (step && (st = launch_step(g)) != SIGP_OK) return st;
    tick_kernel<<<1, 64, 0, g->stream>>>(g->gen);
    GP_TRY(hipGetLastError());
    return SIGP_OK;
}

extern "C" sigp_status sigp_gp_run(sigp_gp *g, const double *sampled, int64_t n_rows, int32_t generations,
                                   int32_t flags, double *fitness_history, int16_t *best_tokens,
                                   int32_t *best_lengths, double *generation_ms) {
    const bool use_graph = (flags & SIGP_GP_GRAPH) != 0, step_last = (flags & SIGP_GP_STEP_LAST) != 0;
    if (!g) return gp_fail("NULL argument");
    sigp_status st0 = check_step(g);
    if (st0 != SIGP_OK) return st0;
    if (generations < 1) return gp_fail("generations must be >= 1");
    const long long rows_per_gen = (long long)g->runs * g->T;
    if (!sampled) {
        if (!g->sampling) return gp_fail("sampled is NULL and no device sampler is set");
        if (n_rows != 0) return gp_fail("n_rows must be 0 with the device sampler");
    } else if (n_rows != rows_per_gen * generations) {
        return gp_fail("need generations x n_runs x T = " + std::to_string(rows_per_gen * generations) + " rows");
    }
    const size_t sbytes = (size_t)n_rows * g->n * 8, hbytes = (size_t)generations * g->P * 8,
                 nbest = (size_t)generations * g->runs;
    char *dev = nullptr;
    const size_t o_hist = (sbytes + 255) & ~(size_t)255, o_btok = o_hist + ((hbytes + 255) & ~(size_t)255),
                 o_blen = o_btok + ((nbest * MAXL * 2 + 255) & ~(size_t)255), total = o_blen + nbest * 4;
    GP_TRY(hipMalloc((void **)&dev, total));
    double *all = sampled ? (double *)dev : nullptr, *hist = (double *)(dev + o_hist);
    int16_t *btok = (int16_t *)(dev + o_btok);
    int32_t *blen = (int32_t *)(dev + o_blen);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> ev(generations + 1, nullptr);
    sigp_status st = SIGP_OK;
    auto check = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && st == SIGP_OK)
            st = sigp_internal_fail(SIGP_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
        return st == SIGP_OK;
    };
    for (auto &e : ev)
        if (!check(hipEventCreate(&e), "event")) break;
    if (st == SIGP_OK && (!sampled || check(hipMemcpyAsync(all, sampled, sbytes, hipMemcpyHostToDevice, g->stream),
                                            "upload"))) {
        set_kernel<<<1, 64, 0, g->stream>>>(g->gen, 0);
        if (use_graph && (generations > 1 || step_last)) {
            // capture one stepping generation; the counter it reads advances on the device
            if (check(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal), "begin capture")) {
                const sigp_status s2 = launch_generation(g, all, hist, btok, blen, true);
                const hipError_t ec = hipStreamEndCapture(g->stream, &graph);
                if (s2 != SIGP_OK && st == SIGP_OK) st = s2;
                if (check(ec, "end capture"))
                    check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
            }
        }
        if (st == SIGP_OK) check(hipEventRecord(ev[0], g->stream), "event");
        for (int gi = 0; st == SIGP_OK && gi < generations; ++gi) {
            const bool step = gi + 1 < generations || step_last;
            if (exec && step) check(hipGraphLaunch(exec, g->stream), "graph launch");
            else {
                const sigp_status s2 = launch_generation(g, all, hist, btok, blen, step);
                if (s2 != SIGP_OK) st = s2;
            }
            if (st == SIGP_OK) check(hipEventRecord(ev[gi + 1], g->stream), "event");
        }
        if (st == SIGP_OK && check(hipStreamSynchronize(g->stream), "run") && generation_ms) {
            for (int gi = 0; gi < generations; ++gi) {
                float ms = 0.0f;
                if (!check(hipEventElapsedTime(&ms, ev[0], ev[gi + 1]), "timing")) break;
                generation_ms[gi] = ms;
            }
        }
        auto down = [&](void *dst, const void *src, size_t bytes) {
            if (st == SIGP_OK && dst) check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "download");
        };
        down(fitness_history, hist, hbytes);
        down(best_tokens, btok, nbest * MAXL * 2);
        down(best_lengths, blen, nbest * 4);
    }
    (void)hipStreamSynchronize(g->stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(dev);
    return st;
}

extern "C" sigp_status sigp_gp_set_sampler(sigp_gp *g, const sigp_sampler *smp, const int64_t *run_seed,
                                           int64_t gen_stride, int64_t gen_offset) {
    if (!g || !smp || !run_seed) return gp_fail("NULL argument");
    if (smp->mode != SIGP_SAMPLE_GAUSS && smp->mode != SIGP_SAMPLE_GAMMA) return gp_fail("bad sampler mode");
    if (smp->mode == SIGP_SAMPLE_GAMMA && !(smp->shape >= 1.0))
        return gp_fail("the device Gamma sampler needs shape >= 1");
    for (int r = 0; r < g->runs; ++r)
        if (run_seed[r] < 0) return gp_fail("run seeds must be >= 0");
    GP_TRY(hipStreamSynchronize(g->stream));
    if (!g->run_seed) GP_TRY(hipMalloc((void **)&g->run_seed, (size_t)g->runs * sizeof(long long)));
    GP_TRY(hipMemcpy(g->run_seed, run_seed, (size_t)g->runs * sizeof(long long), hipMemcpyHostToDevice));
    g->smp = *smp;
    g->gen_stride = gen_stride;
    g->gen_offset = gen_offset;
    g->sampling = true;
    return SIGP_OK;
}

extern "C" sigp_status sigp_gp_population(const sigp_gp *g, int16_t *tokens, int32_t *lengths) {
    if (!g || !tokens || !lengths) return gp_fail("NULL argument");
    if (!g->ready) return gp_fail("no population: call sigp_gp_init first");
    GP_TRY(hipStreamSynchronize(g->stream));
    GP_TRY(hipMemcpy(tokens, g->ptok, (size_t)g->P * MAXL * 2, hipMemcpyDeviceToHost));
    GP_TRY(hipMemcpy(lengths, g->plen, (size_t)g->P * 4, hipMemcpyDeviceToHost));
    return SIGP_OK;
}

extern "C" sigp_status sigp_gp_packed(const sigp_gp *g, int32_t *ops, int32_t *args, int32_t *lengths) {
    if (!g || !ops || !args || !lengths) return gp_fail("NULL argument");
    if (!g->ready) return gp_fail("no population: call sigp_gp_init first");
    GP_TRY(hipStreamSynchronize(g->stream));
    GP_TRY(hipMemcpy(ops, g->pset.raw_ops, (size_t)g->P * MAXL * 4, hipMemcpyDeviceToHost));
    GP_TRY(hipMemcpy(args, g->pset.raw_args, (size_t)g->P * MAXL * 8, hipMemcpyDeviceToHost));
    GP_TRY(hipMemcpy(lengths, g->plen, (size_t)g->P * 4, hipMemcpyDeviceToHost));
    return SIGP_OK;
}

extern "C" sigp_status sigp_gp_rng_state(const sigp_gp *g, int32_t run, int32_t sub, uint32_t *state) {
    if (!g || !state) return gp_fail("NULL argument");
    if (run < 0 || run >= g->runs || sub < 0 || sub >= g->prm.subs) return gp_fail("run or sub-step out of range");
    GP_TRY(hipStreamSynchronize(g->stream));
    GP_TRY(hipMemcpy(state, g->mt + ((size_t)run * g->prm.subs + sub) * MT_WORDS, MT_WORDS * 4,
                     hipMemcpyDeviceToHost));
    return SIGP_OK;
}
