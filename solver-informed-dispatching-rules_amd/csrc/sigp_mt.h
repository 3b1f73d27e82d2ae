// CPython's Mersenne Twister (engine/rng.hpp:13-93) for a wavefront that
// runs one sequential generator in lockstep: all 64 lanes execute the same
// scalar program (identical values, no divergence) and lane 0 stores
// results.  Shared by the GP breeding kernels and the scenario sampler.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace sigp_mt {

constexpr int MT_WORDS = 625;   // 624 state words + position

// CPython's Mersenne Twister (engine/rng.hpp:54-93), driven by the whole
// wavefront in lockstep: the 624 state words in LDS, and a window of 64
// consecutive output words tempered in registers, lane l holding word
// base + l, so that a draw is one read of a lane.

// Regenerate the 624 words.  Word i reads words i + 1 (old) and i + 397
// (old for i < 227) or i - 227 (already new): in chunks of 64 ascending
// indices every such new word lies in an earlier chunk, so the chunks run
// lane-parallel with loads and stores separated by barriers.
__device__ __noinline__ void mt_twist(uint32_t *mt) {
    const int lane = (int)threadIdx.x;
    for (int base = 0; base < 623; base += 64) {
        const int i = base + lane;
        uint32_t nv = 0;
        if (i < 623) {
            const uint32_t y = (mt[i] & 0x80000000u) | (mt[i + 1] & 0x7fffffffu);
            nv = mt[i < 227 ? i + 397 : i - 227] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        __syncthreads();
        if (i < 623) mt[i] = nv;
        __syncthreads();
    }
    if (lane == 0) {
        const uint32_t y = (mt[623] & 0x80000000u) | (mt[0] & 0x7fffffffu);
        mt[623] = mt[396] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    __syncthreads();
}

__device__ __forceinline__ uint32_t temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

struct Rng {
    uint32_t *mt;    // LDS state
    int idx;         // next state word; 624 = all used, regenerate first
    int base;        // words base .. base + 63 are tempered in `win`
    uint32_t win;

    __device__ __forceinline__ uint32_t next() {
        if (idx >= 624 || idx - base >= 64) {
            if (idx >= 624) {
                mt_twist(mt);
                idx = 0;
            }
            base = idx;
            const int w = base + (int)threadIdx.x;
            win = w < 624 ? temper(mt[w]) : 0u;
        }
        const int k = idx - base;
        ++idx;
        return (uint32_t)__builtin_amdgcn_readlane((int)win, k);
    }
    // the 53-bit integer m of random() = m / 2^53
    __device__ __forceinline__ uint64_t bits53() {
        const uint64_t a = next() >> 5, b = next() >> 6;
        return (a << 26) | b;
    }
    __device__ __forceinline__ double random() { return (double)bits53() * (1.0 / 9007199254740992.0); }
    // randbelow: rejection on bit_length(n) top bits; 0 for n == 0
    __device__ __forceinline__ uint32_t below(uint32_t n) {
        if (n == 0) return 0;
        const int k = 32 - __clz(n);
        uint32_t r = next() >> (32 - k);
        while (r >= n) r = next() >> (32 - k);
        return r;
    }
};

__device__ __forceinline__ void load_rng(Rng &g, uint32_t *mt, const uint32_t *state) {
    for (int k = (int)threadIdx.x; k < 624; k += 64) mt[k] = state[k];
    __syncthreads();
    g.mt = mt;
    g.idx = (int)state[624];
    g.base = -1000;   // the window is filled at the first draw
    g.win = 0;
}

__device__ __forceinline__ void store_rng(const Rng &g, uint32_t *state) {
    __syncthreads();
    for (int k = (int)threadIdx.x; k < 624; k += 64) state[k] = g.mt[k];
    if (threadIdx.x == 0) state[624] = (uint32_t)g.idx;
}


// Seeding from a non-negative integer (engine/rng.hpp:21-52: init_by_array
// with the seed's 32-bit words, least significant first), by lane 0 into
// the wavefront's LDS state, then a barrier; the position is 624.
__device__ __forceinline__ void seed_lds(uint32_t *mt, uint64_t seed) {
    if (threadIdx.x == 0) {
        uint32_t key[2];
        int kl = 0;
        if (seed == 0) key[kl++] = 0;
        while (seed) {
            key[kl++] = (uint32_t)(seed & 0xffffffffu);
            seed >>= 32;
        }
        uint32_t prev = 19650218u;
        mt[0] = prev;
        for (int i = 1; i < 624; ++i) {
            prev = 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
            mt[i] = prev;
        }
        int i = 1, j = 0;
        for (int k = 624; k; --k) {
            const uint32_t v = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t)j;
            mt[i] = v;
            ++i;
            ++j;
            if (i >= 624) {
                mt[0] = mt[623];
                i = 1;
            }
            if (j >= kl) j = 0;
        }
        for (int k = 623; k; --k) {
            const uint32_t v = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
            mt[i] = v;
            ++i;
            if (i >= 624) {
                mt[0] = mt[623];
                i = 1;
            }
        }
        mt[0] = 0x80000000u;
    }
    __syncthreads();
}

// CPython's seeding of an integer (engine/rng.hpp:21-52 init_by_array with
// the 32-bit words of the seed, least significant first).
inline void mt_seed(uint64_t seed, uint32_t *st) {
    std::vector<uint32_t> key;
    if (seed == 0) key.push_back(0);
    while (seed) {
        key.push_back((uint32_t)(seed & 0xffffffffu));
        seed >>= 32;
    }
    uint32_t *mt = st;
    mt[0] = 19650218u;
    for (int i = 1; i < 624; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
    size_t i = 1, j = 0;
    for (size_t k = std::max<size_t>(624, key.size()); k; --k) {
        mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t)j;
        ++i;
        ++j;
        if (i >= 624) {
            mt[0] = mt[623];
            i = 1;
        }
        if (j >= key.size()) j = 0;
    }
    for (size_t k = 623; k; --k) {
        mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        ++i;
        if (i >= 624) {
            mt[0] = mt[623];
            i = 1;
        }
    }
    mt[0] = 0x80000000u;
    st[624] = 624;
}


}  // namespace sigp_mt
