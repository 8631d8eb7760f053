// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), written from the paper's
// description. Used for block sampling, test vectors and held-out rows so that
// every draw is a pure function of (seed, block, round, purpose, index) and is
// reproduced exactly on the host.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define FARCUR_HD __host__ __device__ __forceinline__
#else
#define FARCUR_HD static inline
#endif

namespace farcur {

struct U4 {
    uint32_t x, y, z, w;
};

FARCUR_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

FARCUR_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const uint32_t hi0 = mulhi32(M0, c.x), lo0 = M0 * c.x;
        const uint32_t hi1 = mulhi32(M1, c.z), lo1 = M1 * c.z;
        U4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Stream purposes (third counter word, low byte).
enum : uint32_t { kStreamCols = 1, kStreamRows = 2, kStreamVtest = 3, kStreamHeldout = 4 };

// Word t of stream (block, round, purpose): counter (t/4, block, round<<8|purpose, tag).
FARCUR_HD uint32_t stream_word(uint64_t seed, uint32_t block, uint32_t round, uint32_t purpose, uint32_t t) {
    U4 c;
    c.x = t >> 2;
    c.y = block;
    c.z = (round << 8) | purpose;
    c.w = 0x46435552u;  // 'FCUR'
    const U4 o = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint32_t l = t & 3u;
    return l == 0 ? o.x : (l == 1 ? o.y : (l == 2 ? o.z : o.w));
}

// 53-bit uniform double in [0, 1) from two words.
FARCUR_HD double u01_from_words(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Test-vector entry j of block `block`: U[0,1) + i U[0,1), the distribution of
// rand(N) + im*rand(N) in the reference (compres_single_block_spheres_for_sweep2.jl:888).
FARCUR_HD void vtest_entry(uint64_t seed, uint32_t block, uint32_t j, double* re, double* im) {
    U4 c;
    c.x = j;
    c.y = block;
    c.z = kStreamVtest;
    c.w = 0x46435552u;
    const U4 o = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    *re = u01_from_words(o.x, o.y);
    *im = u01_from_words(o.z, o.w);
}

// Selection sampling (Knuth, TAOCP vol. 2, 3.4.2, Algorithm S): picks exactly
// n of the indices 0..N-1, uniformly, in increasing order. The acceptance test
// floor(u * (N - t) / 2^32) < n - m is integer-exact on host and device.
FARCUR_HD bool select_step(uint32_t u, uint32_t remaining_candidates, uint32_t remaining_needed) {
    return (uint32_t)(((uint64_t)u * remaining_candidates) >> 32) < remaining_needed;
}

}  // namespace farcur
