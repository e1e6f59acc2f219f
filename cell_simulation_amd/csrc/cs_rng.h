// cs_rng.h -- counter RNG of the reference shaders, for host and device code.
//
// res/shader/random.wgsl:14-20 (PCG hash; the LCG line 17 is dead code in the reference, so the output
// hashes the input state), :32-45 (xorshift32), :27-29 (u32 -> f32 scaled by 2^-32, range [0,1]).
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define CS_HD __host__ __device__ __forceinline__
#else
#define CS_HD inline
#endif

namespace cs {

CS_HD uint32_t pcg_hash(uint32_t state)
{
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

CS_HD uint32_t xorshift32(uint32_t s)
{
    s ^= s << 13u;
    s ^= s >> 17u;
    s ^= s << 5u;
    return s;
}

// f32(u) * 2^-32; the conversion rounds to nearest even, so 1.0 is reached for u >= 4294967168.
CS_HD float uint_to_unit(uint32_t u)
{
    return static_cast<float>(u) * 2.3283064365386963e-10f;
}

CS_HD float uniform_float(uint32_t state)
{
    return uint_to_unit(xorshift32(state));
}

// ---- well-mixed (CME) phase: the build's own generator and propensity (the reference has no CME phase) --

// 53 random bits for draw k of a step: two hashed counters per draw, keyed by the step's itime and the
// simulation's CME seed.
CS_HD uint64_t cme_bits(uint32_t itime, uint32_t seed, uint32_t k)
{
    const uint32_t a = pcg_hash(itime + pcg_hash(seed + 2u * k));
    const uint32_t b = pcg_hash(itime + pcg_hash(seed + 2u * k + 1u));
    return ((uint64_t(a) << 32) | b) >> 11;
}

// Voxel draw for placing lattice product `product` of CME event `event`, attempt `attempt` (< 128).
CS_HD uint32_t cme_place_draw(uint32_t itime, uint32_t seed, uint32_t event, uint32_t product, uint32_t attempt)
{
    return pcg_hash(itime + pcg_hash((seed ^ 0x9E3779B9u) + ((event * 4u + product) << 7) + attempt));
}

// Combinatorial factor of a CME reaction: the product over its distinct reactant species of
// binomial(count, multiplicity), in double (exact while counts stay below 2^17 for trimolecular rows).
CS_HD double cme_h(const int32_t *row, const int32_t *counts)
{
    double h = 1.0;
    for (int j = 0; j < 3; ++j) {
        const int32_t id = row[j];
        if (id < 0)
            continue;
        bool seen = false;
        int m = 0;
        for (int q = 0; q < 3; ++q) {
            if (row[q] == id) {
                ++m;
                if (q < j)
                    seen = true;
            }
        }
        if (seen)
            continue;
        const double n = double(counts[id]);
        if (m == 1)
            h = h * n;
        else if (m == 2)
            h = h * (n * (n - 1.0) * 0.5);
        else
            h = h * (n * (n - 1.0) * (n - 2.0) / 6.0);
        if (!(h > 0.0))
            return 0.0;
    }
    return h;
}

}  // namespace cs
