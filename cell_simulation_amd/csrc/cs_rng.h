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

}  // namespace cs
