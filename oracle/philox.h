// ORACLE — test infrastructure only (see ref.h).
//
// Counter-based sampler that replaces the reference's per-pixel cuRAND XORWOW state
// (sampling/Sampler.cuh:11-82). Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), keyed by (seed, pixel) with counter
// (sample, dimension / 4, 0, 0). Each draw is addressed, not consumed, so the draw for
// (pixel, sample, dimension) is the same on every code path.
//
// Dimension layout (follows the reference's draw order, PathTracingIntegrator.cuh:41-99
// and IntegratorHelper.cuh:142-146):
//   0-1 film sample, 2-3 lens sample,
//   bounce b, base 4 + 8b: +0 light pick, +1..2 light sample, +3..4 MIS-BSDF sample,
//   +5..6 continuation BSDF sample, +7 Russian roulette.
#pragma once
#include <cstdint>

namespace orc {

struct U4 { uint32_t v[4]; };

inline void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

inline U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += W0; k1 += W1; }
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(M0, ctr.v[0], &hi0, &lo0);
        mulhilo32(M1, ctr.v[2], &hi1, &lo1);
        U4 o;
        o.v[0] = hi1 ^ ctr.v[1] ^ k0;
        o.v[1] = lo1;
        o.v[2] = hi0 ^ ctr.v[3] ^ k1;
        o.v[3] = lo0;
        ctr = o;
    }
    return ctr;
}

// Maps 32 random bits to a float in (0, 1], the range curand_uniform documents.
inline float u01(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

struct Sampler {
    uint32_t seed, pixel, sample;
    uint32_t cachedBlock = 0xffffffffu;
    U4 cache;
    float get(int dim) {
        uint32_t blk = (uint32_t)dim >> 2;
        if (blk != cachedBlock) {
            U4 c; c.v[0] = sample; c.v[1] = blk; c.v[2] = 0; c.v[3] = 0;
            cache = philox4x32_10(c, seed, pixel);
            cachedBlock = blk;
        }
        return u01(cache.v[dim & 3]);
    }
};

} // namespace orc
