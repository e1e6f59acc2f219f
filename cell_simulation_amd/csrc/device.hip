// device.hip -- the RDME step on gfx950: one fused gather kernel per step.
//
// The reference runs three passes per step (res/shader/rdme.wgsl, cme.wgsl, final_stage.wgsl) plus two
// buffer copies (src/simulation.rs:99-150); its diffusion pass scatters particles with spin locks and
// CAS first-fit slot claims. Here every voxel is computed by one lane that GATHERS: it recomputes the
// hop decision of each particle in its six neighbours (a pure function of the source lattice, the
// region labels and itime) and lays out the arrivals in the order the canonical serial execution
// (voxels ascending, slots ascending) would produce them. The reaction and the compaction follow in
// registers, and the result goes to the other half of a ping-pong pair, so no copies and no atomics
// on the lattice are needed.
//
// Canonical slot picture of voxel d after diffusion, when no voxel can fill up (see `assemble`):
//   slots 0..occ-1 : own particles that stay, departed slots refilled by arrivals from the higher-index
//                    neighbours (+z, +y, +x, each in source-slot order) while such arrivals remain;
//   then           : arrivals from the lower-index neighbours (-x, -y, -z, each in source-slot order);
//   then           : the remaining higher-index arrivals.
// A voxel whose own particles plus all arrivals exceed S ("tight") is counted in CS_CNT_TIGHT; there
// the capacity hazards of rdme.wgsl:57-87 (rejected claims, the slot-S alias) change the outcome.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cell_sim.h"
#include "cs_rng.h"
#include "device.hpp"
#include "ensemble.hpp"

#include <atomic>
#include <mutex>

// Bins of cs_sim_read_region_counts_noalloc's histogram: one fixed array per GPU, so that entry point allocates
// nothing. Host code passes the address hipGetSymbolAddress gives for the current device (noalloc_bins), never
// this name.
constexpr uint32_t kNoAllocBins = 16384;
static __device__ __align__(256) int32_t scratch[kNoAllocBins];

namespace cs {

namespace {

constexpr int kMaxRW = 4;  // reaction bitmask words: up to 255 reactions
constexpr int kBlock = 256;
constexpr uint32_t kFixCtl = 16;      // control words: [0] voxels noted for the replay, [2] replay scratch used, [8..15] queues
constexpr uint32_t kFixQueue = 8;

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw Error(CS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Occupancy summary of one lattice buffer, kept for the wave kernel: byte ((y / kSumTY) * tiles_z + z / kSumTZ)
// * planes + x is non-zero whenever that wave tile's voxels of plane x may hold particles (a non-zero byte may be
// stale; a zero one never is). Every writer of particles into a buffer sets the byte; the wave kernel also
// clears the bytes of the planes it leaves empty. Plain byte stores: within a step a byte has one writer (the
// work item owning the plane; the passes before and after the step kernel run as kernels of their own), so no
// two writers share a byte, while bits packed in a word would need atomics coherent across the XCDs' L2s.
constexpr uint32_t kSumTY = 4, kSumTZ = 64;  // the wave tile (kWY x kWZ)
struct SumRef {
    uint8_t *bytes;  // null: no summary kept for this buffer
    uint32_t planes, tiles_z, R1, R2;  // planes = R0
};

__device__ __forceinline__ void sum_mark(const SumRef &s, uint32_t v)
{
    if (s.bytes == nullptr)
        return;
    const uint32_t z = v % s.R2, t = v / s.R2, y = t % s.R1, x = t / s.R1;
    s.bytes[((y / kSumTY) * s.tiles_z + z / kSumTZ) * s.planes + x] = 1u;
}

struct StepArgs {
    const uint8_t *src;
    uint8_t *dst;
    const uint8_t *reg;
    const float *p2;
    const float *rates;
    const uint8_t *ridx;
    const int8_t *stoich;
    const uint64_t *rmask;
    const float *bound;
    unsigned long long *counters;
    uint32_t *fix_ctl;    // [0] voxels noted this step for the replay, [kFixQueue..] the wave kernel's queues
    uint32_t *fix_list;   // voxels noted (max_tight), as indices among all voxels of the device state
    uint32_t max_tight;
    uint32_t vox_base;    // index of this lattice's voxel 0 among all voxels on the device (ensembles)
    uint32_t R0, R1, R2, R12;
    uint32_t N;
    uint32_t ns, nr, nreg, nrw;
    uint32_t itime;
    uint32_t mod_zero;
    float tau;
    const uint32_t *live;      // bit per voxel (index vox_base + v): a skewed voxel, whose reaction the replay runs
                               // (null in the wave kernels while the skewed-voxel table is empty: no bit is set)
    const uint32_t *skew_n;    // entries of the skewed-voxel table (no entry is added during a step kernel)
    SumRef sum_dst;            // summary of dst, marked by the passes after the step kernel
};

__device__ __forceinline__ bool voxel_bit(const uint32_t *bits, uint32_t v)
{
    return (bits[v >> 5] >> (v & 31u)) & 1u;
}

// A skewed voxel, whose reaction the replay kernel runs; false for every voxel when a.live is null.
__device__ __forceinline__ bool live_voxel(const StepArgs &a, uint32_t v)
{
    return a.live != nullptr && voxel_bit(a.live, v);
}

__device__ __forceinline__ void wave_sync()
{
    // LDS accesses of one wave complete in order; this keeps the compiler from moving them across.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- slot words: S one-byte slots packed little-endian into S/4 u32 ----------------------------
template <int S>
struct Slots {
    uint32_t w[S / 4];
};

template <int S>
__device__ __forceinline__ Slots<S> load_slots(const uint8_t *base, uint32_t v)
{
    Slots<S> s;
    if constexpr (S == 16) {
        const uint4 q = *reinterpret_cast<const uint4 *>(base + size_t(v) * 16u);
        s.w[0] = q.x; s.w[1] = q.y; s.w[2] = q.z; s.w[3] = q.w;
    } else if constexpr (S == 8) {
        const uint2 q = *reinterpret_cast<const uint2 *>(base + size_t(v) * 8u);
        s.w[0] = q.x; s.w[1] = q.y;
    } else {
        s.w[0] = *reinterpret_cast<const uint32_t *>(base + size_t(v) * 4u);
    }
    return s;
}

template <int S>
__device__ __forceinline__ void store_slots(uint8_t *base, uint32_t v, const Slots<S> &s)
{
    if constexpr (S == 16) {
        *reinterpret_cast<uint4 *>(base + size_t(v) * 16u) = make_uint4(s.w[0], s.w[1], s.w[2], s.w[3]);
    } else if constexpr (S == 8) {
        *reinterpret_cast<uint2 *>(base + size_t(v) * 8u) = make_uint2(s.w[0], s.w[1]);
    } else {
        *reinterpret_cast<uint32_t *>(base + size_t(v) * 4u) = s.w[0];
    }
}

template <int S>
__device__ __forceinline__ Slots<S> zero_slots()
{
    Slots<S> s;
#pragma unroll
    for (int q = 0; q < S / 4; ++q)
        s.w[q] = 0u;
    return s;
}

// high bit of each byte set iff that byte is non-zero
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t w)
{
    return (((w & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | w) & 0x80808080u;
}

template <int S>
__device__ __forceinline__ uint32_t count_nonzero(const Slots<S> &s)
{
    uint32_t n = 0;
#pragma unroll
    for (int q = 0; q < S / 4; ++q)
        n += __builtin_popcount(nonzero_bytes(s.w[q]));
    return n;
}

// number of slots holding species sp (sp >= 1); only the first `words` slot words are examined (a
// wave-uniform bound: the slots are compacted, so words past every lane's occupancy hold no particle)
// count_nonzero of compacted slots: at S = 16 the upper eight hold particles only when the lower eight are full,
// so they are counted only then (no lane of a pass holds a full head, as a rule: eight fewer vector operations)
template <int S>
__device__ __forceinline__ uint32_t count_compacted(const Slots<S> &s)
{
    if constexpr (S == 16) {
        uint32_t n = uint32_t(__builtin_popcount(nonzero_bytes(s.w[0])) + __builtin_popcount(nonzero_bytes(s.w[1])));
        if (n == 8u)
            n += uint32_t(__builtin_popcount(nonzero_bytes(s.w[2])) + __builtin_popcount(nonzero_bytes(s.w[3])));
        return n;
    } else {
        return count_nonzero(s);
    }
}

template <int S>
__device__ __forceinline__ uint32_t count_species(const Slots<S> &s, uint32_t sp, uint32_t words = S / 4)
{
    const uint32_t pat = sp * 0x01010101u;
    uint32_t n = 0;
#pragma unroll
    for (int q = 0; q < S / 4; ++q)
        if (uint32_t(q) < words)
            n += __builtin_popcount(nonzero_bytes(s.w[q] ^ pat) ^ 0x80808080u);
    return n;
}

template <int S>
__device__ __forceinline__ uint32_t get_byte(const Slots<S> &s, int i)
{
    return (s.w[i >> 2] >> ((i & 3) * 8)) & 0xFFu;
}

// Byte at a run-time position j < S. Written as 64-bit shifts over whole values, never as an indexed
// element: an indexed 4-word array is kept in scratch memory by the compiler.
template <int S>
__device__ __forceinline__ uint32_t get_byte_dyn(const Slots<S> &s, uint32_t j)
{
    if constexpr (S == 4) {
        return (s.w[0] >> ((j & 3u) * 8u)) & 0xFFu;
    } else {
        uint64_t v = uint64_t(s.w[0]) | (uint64_t(s.w[1]) << 32);
        if constexpr (S == 16) {
            const uint64_t hi = uint64_t(s.w[2]) | (uint64_t(s.w[3]) << 32);
            v = (j & 8u) ? hi : v;
        }
        return uint32_t(v >> ((j & 7u) * 8u)) & 0xFFu;
    }
}

// OR a byte into a zero slot at run-time position j; positions >= S are dropped
template <int S>
__device__ __forceinline__ void put_byte_dyn(Slots<S> &s, uint32_t j, uint32_t val)
{
    if constexpr (S == 4) {
        s.w[0] |= (j < 4u) ? (val << ((j & 3u) * 8u)) : 0u;
    } else {
        const uint64_t bit = uint64_t(val) << ((j & 7u) * 8u);
        const uint64_t lo = (j < 8u) ? bit : 0ull;
        s.w[0] |= uint32_t(lo);
        s.w[1] |= uint32_t(lo >> 32);
        if constexpr (S == 16) {
            const uint64_t hi = (j >= 8u && j < 16u) ? bit : 0ull;
            s.w[2] |= uint32_t(hi);
            s.w[3] |= uint32_t(hi >> 32);
        }
    }
}

// ---- hop decision (rdme.wgsl:132-136, 138-191, 220-226) -----------------------------------------
struct Hood {
    uint32_t src_reg;
    uint32_t rg[6];     // region of the neighbour in direction -x,+x,-y,+y,-z,+z (if inside)
    uint32_t inside;    // bit d set if that neighbour exists
    uint32_t seed0;     // itime + (x % y) + z + v*S; slot s adds s
};

__device__ __forceinline__ Hood make_hood(const StepArgs &a, uint32_t x, uint32_t y, uint32_t z, uint32_t v,
                                          uint32_t S)
{
    Hood h;
    h.src_reg = a.reg[v];
    h.inside = 0u;
#pragma unroll
    for (int d = 0; d < 6; ++d)
        h.rg[d] = 0u;
    if (x > 0u) { h.inside |= 1u; h.rg[0] = a.reg[v - a.R12]; }
    if (x + 1u < a.R0) { h.inside |= 2u; h.rg[1] = a.reg[v + a.R12]; }
    if (y > 0u) { h.inside |= 4u; h.rg[2] = a.reg[v - a.R2]; }
    if (y + 1u < a.R1) { h.inside |= 8u; h.rg[3] = a.reg[v + a.R2]; }
    if (z > 0u) { h.inside |= 16u; h.rg[4] = a.reg[v - 1u]; }
    if (z + 1u < a.R2) { h.inside |= 32u; h.rg[5] = a.reg[v + 1u]; }
    const uint32_t xmy = (y != 0u) ? (x % y) : a.mod_zero;  // X % 0: hazard H4
    h.seed0 = a.itime + xmy + z + v * S;
    return h;
}

// The model tables as the kernels read them: in LDS for the tiled kernel, in global memory otherwise.
struct Tab {
    const float *p2;        // [nreg][nreg][ns+1] hop probability per direction
    const float *rates;     // [nr+1]
    const uint32_t *ridx;   // [nr+1] three reactant species, one byte each
    const int8_t *stoich;   // [nr+1][ns+1]
    const uint64_t *rmask;  // [ns+1][nrw] reactions having the species as a reactant; row 0: zero-order
    const float *bound;     // [S+1] upper bound of the total propensity of a voxel holding c particles
    const float *pair = nullptr;  // [ns+1][ns+1] upper bound for a voxel holding species a and b (0: no particle),
                                  // or null; see pair_bounds
};

// The pair table follows the S + 1 propensity bounds in the table blob, 8-byte aligned; bound_words is the length
// of both in 32-bit words (n1 = ns + 1, or 0 without a pair table).
__host__ __device__ constexpr uint32_t pair_offset(uint32_t S) { return (S + 2u) & ~1u; }
__host__ __device__ constexpr uint32_t bound_words(uint32_t S, uint32_t n1)
{
    return n1 != 0u ? pair_offset(S) + n1 * n1 : S + 1u;
}

__device__ __forceinline__ Tab global_tab(const StepArgs &a)
{
    return Tab{a.p2, a.rates, reinterpret_cast<const uint32_t *>(a.ridx), a.stoich, a.rmask, a.bound};
}

// Direction 0..5 of an existing neighbour the particle hops to, or 6 if it stays (stay draw or
// reflecting boundary). The cumulative sum is the reference's p[0..5] in f32, in the same order; all six
// table entries are fetched before the comparisons so the loads overlap.
__device__ __forceinline__ uint32_t hop(const StepArgs &a, const Tab &tb, const Hood &h, uint32_t sp, uint32_t slot)
{
    const float r = uniform_float(pcg_hash(h.seed0 + slot));
    const uint32_t n1 = a.ns + 1u;
    // table indices are below 2^24: 24-bit multiplies run at full rate, 32-bit ones at a quarter
    const float *row = tb.p2 + __umul24(n1 * a.nreg, h.src_reg) + sp;
    float val[6];
#pragma unroll
    for (int d = 0; d < 6; ++d)
        val[d] = row[__umul24(n1, h.rg[d])];
    float acc[6];
    float s = 0.0f;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (h.inside & (1u << d))
            s = s + val[d];
        acc[d] = s;
    }
    // the first d with r <= acc[d] decides: the hop if that neighbour exists, else the particle stays; selects
    // from the last direction down, so no branch
    uint32_t dir = 6u;
#pragma unroll
    for (int d = 5; d >= 0; --d)
        dir = !(r > acc[d]) ? (((h.inside >> d) & 1u) ? uint32_t(d) : 6u) : dir;
    return dir;
}

// Append the particles of neighbour u that hop in direction `want` (toward the voxel being gathered),
// in source-slot order. Returns the number found (may exceed the list capacity S).
template <int S>
__device__ __forceinline__ uint32_t gather_from(const StepArgs &a, const Tab &tb, uint32_t ux, uint32_t uy,
                                                uint32_t uz, uint32_t u, uint32_t want, Slots<S> &list, uint32_t n)
{
    const Slots<S> su = load_slots<S>(a.src, u);
    const uint32_t occ = count_nonzero(su);
    if (occ == 0u)
        return n;
    const Hood h = make_hood(a, ux, uy, uz, u, S);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if (uint32_t(s) < occ) {
            const uint32_t sp = get_byte(su, s);
            if (hop(a, tb, h, sp, s) == want) {
                put_byte_dyn(list, n, sp);
                ++n;
            }
        }
    }
    return n;
}

__device__ __forceinline__ void count_event(unsigned long long *counters, int idx, bool pred)
{
    const unsigned long long m = __ballot(pred);
    if (m != 0ull && __lane_id() == (uint32_t)__ffsll((long long)m) - 1u)
        atomicAdd(counters + idx, (unsigned long long)__popcll(m));
}

// ---- reaction (cme.wgsl:25-120) on the compacted post-diffusion slots ------------------------------
template <int S>
__device__ __forceinline__ uint32_t species_count(const Slots<S> &s, uint32_t sp, uint32_t words)
{
    return sp == 0u ? 1u : count_species(s, sp, words);  // C[v][0] = 1 (cme.wgsl:31)
}

template <int S>
__device__ __forceinline__ float propensity(const Tab &tb, const Slots<S> &s, uint32_t r, uint32_t words)
{
    const uint32_t ix = tb.ridx[r];
    const uint32_t prod = species_count(s, ix & 0xFFu, words) * species_count(s, (ix >> 8) & 0xFFu, words) *
                          species_count(s, (ix >> 16) & 0xFFu, words);
    return tb.rates[r] * float(int32_t(prod));
}

// Slot words that hold the particles of every lane of the wave with `cnt` particles (wave-uniform).
template <int S>
__device__ __forceinline__ uint32_t wave_words(uint32_t cnt)
{
    if constexpr (S == 4)
        return 1u;
    else if constexpr (S == 8)
        return __ballot(cnt > 4u) ? 2u : 1u;
    else
        return __ballot(cnt > 8u) ? (__ballot(cnt > 12u) ? 4u : 3u) : (__ballot(cnt > 4u) ? 2u : 1u);
}

// Applies the reaction step to `out` (cnt particles, compacted). Reactions with a reactant absent from
// the voxel have propensity +0 and leave every partial sum unchanged, so only candidate reactions
// (zero-order ones and those with a reactant present) are visited; the f32 sums match the reference's
// sequential sums bit for bit.
// The reject test that opens react(): false when the voxel's draw lies above the upper bound of its total
// propensity, i.e. nothing can fire there this step.
// With the pair table, a voxel of at most four particles (species s0..s3 in slots 0..3, zeros past cnt) is bounded
// by the sum of the pair entries over its pairs, which is at least its total propensity when no reaction has more
// than two reactants (pair_bounds); more particles take bound[cnt].
template <int S>
__device__ __forceinline__ bool react_may_fire(const StepArgs &a, const Tab &tb, const Slots<S> &out, uint32_t cnt,
                                               uint32_t x, uint32_t y, uint32_t z)
{
    // x * y with x, y < 2^24: the low 32 bits of the 24-bit product are those of the 32-bit one
    const float r1 = uniform_float(pcg_hash(a.itime + __umul24(x, y) + z));
    float bnd;
    if (tb.pair != nullptr && cnt <= 4u) {
        const uint32_t n1 = a.ns + 1u, w0 = out.w[0];
        const uint32_t s0 = w0 & 0xFFu, s1 = (w0 >> 8) & 0xFFu;
        bnd = tb.pair[__umul24(s0, n1) + s1];
        if (cnt >= 3u) {
            const uint32_t s2 = (w0 >> 16) & 0xFFu;
            bnd = bnd + tb.pair[__umul24(s0, n1) + s2] + tb.pair[__umul24(s1, n1) + s2];
            if (cnt == 4u) {
                const uint32_t s3 = w0 >> 24;
                bnd = bnd + tb.pair[__umul24(s0, n1) + s3] + tb.pair[__umul24(s1, n1) + s3] +
                      tb.pair[__umul24(s2, n1) + s3];
            }
        }
    } else {
        bnd = tb.bound[cnt];
    }
    return !(r1 > bnd * a.tau * 1.001f + 1e-6f);
}

template <int S>
__device__ __forceinline__ void react(const StepArgs &a, const Tab &tb, Slots<S> &out, uint32_t cnt, uint32_t x,
                                      uint32_t y, uint32_t z, bool &fired, bool &rolled, bool &near, bool &bad)
{
    fired = rolled = near = bad = false;
    const uint32_t P = pcg_hash(a.itime + x * y + z);
    const float r1 = uniform_float(P);
    // The total propensity never exceeds bound[cnt], and f32 products and sums are monotone, so a draw
    // above the bound's reject line is above the exact one below as well: nothing fires, nothing is near.
    if (r1 > tb.bound[cnt] * a.tau * 1.001f + 1e-6f)
        return;
    if (live_voxel(a, a.vox_base + z + y * a.R2 + x * a.R12))
        return;  // a skewed voxel: its reaction runs in the replay kernel after the step
    const uint32_t words = wave_words<S>(cnt);  // slot words holding particles in any lane still here

    uint64_t cand[kMaxRW];
#pragma unroll
    for (int w = 0; w < kMaxRW; ++w)
        cand[w] = (uint32_t(w) < a.nrw) ? tb.rmask[w] : 0ull;
    for (uint32_t j = 0; j < cnt; ++j) {
        const uint32_t sp = get_byte_dyn(out, j);
#pragma unroll
        for (int w = 0; w < kMaxRW; ++w)
            if (uint32_t(w) < a.nrw)
                cand[w] |= tb.rmask[sp * a.nrw + w];
    }

    float tot = 0.0f;
#pragma unroll
    for (int w = 0; w < kMaxRW; ++w) {
        uint64_t m = cand[w];
        while (m) {
            const uint32_t r = uint32_t(w) * 64u + uint32_t(__ffsll((long long)m) - 1);
            m &= m - 1ull;
            tot = tot + propensity(tb, out, r, words);
        }
    }

    // r1 <= 1 - exp(-tot*tau) (cme.wgsl:66). For tot >= 0 that threshold never exceeds
    // x + 2^-23 (x = tot*tau rounded), so larger r1 cannot fire nor sit within 2^-24 of it.
    const float xt = tot * a.tau;
    if (tot >= 0.0f && r1 > xt * 1.001f + 1e-6f)
        return;
    const float e = float(exp(double(-xt)));
    const float thr = 1.0f - e;
    near = xt != 0.0f && fabsf(r1 - thr) <= 5.9604644775390625e-08f;  // exp(-0) = 1 exactly
    if (!(r1 <= thr))
        return;
    fired = true;

    // Reaction choice (cme.wgsl:50-52, 68-73): first i with r2 <= cum[i]/tot, i < nr.
    const float r2 = uniform_float(P + 1u);
    uint32_t i = 0;
    float cum = 0.0f;
    bool chosen = !(r2 > cum / tot) || a.nr == 0u;
    for (int w = 0; w < kMaxRW && !chosen; ++w) {
        uint64_t m = cand[w];
        while (m && !chosen) {
            const uint32_t r = uint32_t(w) * 64u + uint32_t(__ffsll((long long)m) - 1);
            m &= m - 1ull;
            cum = cum + propensity(tb, out, r, words);
            i = r;
            if (r >= a.nr || !(r2 > cum / tot))
                chosen = true;
        }
    }
    if (!chosen)
        i = a.nr;

    // Stoichiometry and slot rewrite in species order (cme.wgsl:76-119). The rewrite is undone when it
    // reaches S slots (cme.wgsl:95-108), leaving the pre-reaction slots.
    const int8_t *row = tb.stoich + i * (a.ns + 1u);
    Slots<S> res = zero_slots<S>();
    uint32_t j = 0;
    for (uint32_t sp = 1; sp <= a.ns; ++sp) {
        const int32_t c = int32_t(count_species(out, sp, words)) + int32_t(row[sp]);
        if (c < 0)
            bad = true;
        if (c <= 0)
            continue;
        if (j + uint32_t(c) >= uint32_t(S)) {
            rolled = true;
            return;
        }
        for (int32_t k = 0; k < c; ++k)
            put_byte_dyn(res, j++, sp);
    }
    out = res;
}

// ---- byte vectors: a voxel's slots as two 64-bit halves (slots 0..7, 8..15) ---------------------------
struct BV {
    uint64_t lo, hi;
};

template <int S>
__device__ __forceinline__ BV to_bv(const Slots<S> &s)
{
    if constexpr (S == 16)
        return BV{uint64_t(s.w[0]) | (uint64_t(s.w[1]) << 32), uint64_t(s.w[2]) | (uint64_t(s.w[3]) << 32)};
    else if constexpr (S == 8)
        return BV{uint64_t(s.w[0]) | (uint64_t(s.w[1]) << 32), 0ull};
    else
        return BV{uint64_t(s.w[0]), 0ull};
}

template <int S>
__device__ __forceinline__ Slots<S> from_bv(const BV &v)
{
    Slots<S> s;
    s.w[0] = uint32_t(v.lo);
    if constexpr (S >= 8)
        s.w[1] = uint32_t(v.lo >> 32);
    if constexpr (S == 16) {
        s.w[2] = uint32_t(v.hi);
        s.w[3] = uint32_t(v.hi >> 32);
    }
    return s;
}

// Moves every byte k positions up (k <= 16); bytes moved past position 15 are dropped.
__device__ __forceinline__ BV shl_bytes(const BV &v, uint32_t k)
{
    const uint32_t b = 8u * (k & 7u);
    const uint64_t lo = v.lo << b;
    const uint64_t hi = (v.hi << b) | (b ? (v.lo >> (64u - b)) : 0ull);
    return BV{k >= 8u ? 0ull : lo, k >= 16u ? 0ull : (k >= 8u ? lo : hi)};
}

// Moves every byte k positions down (k <= 16); zeros come in at the top.
__device__ __forceinline__ BV shr_bytes(const BV &v, uint32_t k)
{
    const uint32_t b = 8u * (k & 7u);
    const uint64_t hi = v.hi >> b;
    const uint64_t lo = (v.lo >> b) | (b ? (v.hi << (64u - b)) : 0ull);
    return BV{k >= 16u ? 0ull : (k >= 8u ? hi : lo), k >= 8u ? 0ull : hi};
}

// Mask of the bytes below position d (d <= 16).
__device__ __forceinline__ BV bytes_below(uint32_t d)
{
    const uint64_t m = (d & 7u) ? ((1ull << (8u * (d & 7u))) - 1ull) : 0ull;
    return BV{d >= 8u ? ~0ull : m, d >= 16u ? ~0ull : (d >= 8u ? m : 0ull)};
}

__device__ __forceinline__ uint32_t bv_byte(const BV &v, uint32_t j)
{
    // half chosen by masking, not by a select: with j uniform the compiler turns a select of the two
    // halves into an indexed load from scratch memory
    const uint64_t w = v.lo ^ ((v.hi ^ v.lo) & (0ull - uint64_t((j >> 3) & 1u)));
    return uint32_t(w >> (8u * (j & 7u))) & 0xFFu;
}

// Byte at position d replaced by val.
__device__ __forceinline__ BV bv_set(const BV &v, uint32_t d, uint32_t val)
{
    const uint64_t m = 0xFFull << (8u * (d & 7u)), x = uint64_t(val) << (8u * (d & 7u));
    return (d & 8u) ? BV{v.lo, (v.hi & ~m) | x} : BV{(v.lo & ~m) | x, v.hi};
}

// Byte at position d removed; the bytes above it move down one place.
__device__ __forceinline__ BV bv_delete(const BV &v, uint32_t d)
{
    const BV m = bytes_below(d), sh = shr_bytes(v, 1u);
    return BV{(v.lo & m.lo) | (sh.lo & ~m.lo), (v.hi & m.hi) | (sh.hi & ~m.hi)};
}

// Bit 4s set for s < n (n <= 16): the nibble mask of the first n slots.
__device__ __forceinline__ uint64_t first_nibbles(uint32_t n)
{
    return (n >= 16u ? ~0ull : ((1ull << (4u * n)) - 1ull)) & 0x1111111111111111ull;
}

// Canonical slot picture of the voxel after diffusion (see the header comment). `stay_mask` has bit 4s
// set when own slot s stays. Departed slots are refilled, in slot order, by the higher-index arrivals
// while they last; departed slots left over close up; the lower-index arrivals and the remaining
// higher-index arrivals follow. Done with byte shifts: the loops run once per departed particle only.
template <int S>
__device__ __forceinline__ Slots<S> assemble(const Slots<S> &own, uint32_t occ, uint64_t stay_mask,
                                             const Slots<S> &lo, uint32_t nlo, const Slots<S> &hi, uint32_t nhi,
                                             uint32_t &cnt)
{
    BV o = to_bv<S>(own), h = to_bv<S>(hi);
    uint64_t rem = first_nibbles(occ) & ~stay_mask;  // departed own slots
    uint32_t used = 0;
    while (rem != 0ull && used < nhi) {
        const uint32_t sl = uint32_t(__ffsll((long long)rem) - 1) >> 2;
        rem &= rem - 1ull;
        o = bv_set(o, sl, bv_byte(h, used));
        ++used;
    }
    const uint32_t j = occ - uint32_t(__popcll(rem));
    while (rem != 0ull) {  // highest first, so the lower positions stay valid
        const uint32_t bit = 63u - uint32_t(__clzll((long long)rem));
        rem &= ~(1ull << bit);
        o = bv_delete(o, bit >> 2);
    }
    const BV l = shl_bytes(to_bv<S>(lo), j), hr = shl_bytes(shr_bytes(h, used), j + nlo);
    o.lo |= l.lo | hr.lo;
    o.hi |= l.hi | hr.hi;
    const uint32_t n = j + nlo + (nhi - used);
    cnt = n < uint32_t(S) ? n : uint32_t(S);  // a tight voxel may collect more than S; extras were dropped
    return from_bv<S>(o);
}

__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
{
    // Consecutive dispatch IDs go round-robin over the 8 XCDs; give each XCD a contiguous range of
    // lattice blocks so neighbouring planes share its L2.
    const uint32_t q = nb / 8u, rem = nb % 8u, xcd = b % 8u, idx = b / 8u;
    return xcd * q + min(xcd, rem) + idx;
}

struct VoxelFlags {
    bool tight = false, fired = false, rolled = false, near = false, bad = false;
};

// The gathered step of voxel v, everything read from global memory (the one-lane-per-voxel form).
template <int S>
__device__ __forceinline__ Slots<S> gather_voxel(const StepArgs &a, const Tab &tb, uint32_t v, VoxelFlags &f)
{
    const uint32_t z = v % a.R2;
    const uint32_t y = (v / a.R2) % a.R1;
    const uint32_t x = v / a.R12;
    const Slots<S> own = load_slots<S>(a.src, v);
    const uint32_t occ = count_nonzero(own);
    uint64_t stay = ~0ull;
    if (occ) {
        const Hood h = make_hood(a, x, y, z, v, S);
        stay = 0ull;
#pragma unroll
        for (int s = 0; s < S; ++s)
            if (uint32_t(s) < occ && hop(a, tb, h, get_byte(own, s), s) == 6u)
                stay |= 1ull << (4 * s);
    }
    Slots<S> lo = zero_slots<S>(), hi = zero_slots<S>();
    uint32_t nlo = 0, nhi = 0;
    if (x > 0u) nlo = gather_from<S>(a, tb, x - 1u, y, z, v - a.R12, 1u, lo, nlo);      // from -x, hopped +x
    if (y > 0u) nlo = gather_from<S>(a, tb, x, y - 1u, z, v - a.R2, 3u, lo, nlo);       // from -y, hopped +y
    if (z > 0u) nlo = gather_from<S>(a, tb, x, y, z - 1u, v - 1u, 5u, lo, nlo);         // from -z, hopped +z
    if (z + 1u < a.R2) nhi = gather_from<S>(a, tb, x, y, z + 1u, v + 1u, 4u, hi, nhi);  // from +z, hopped -z
    if (y + 1u < a.R1) nhi = gather_from<S>(a, tb, x, y + 1u, z, v + a.R2, 2u, hi, nhi);  // from +y, hopped -y
    if (x + 1u < a.R0) nhi = gather_from<S>(a, tb, x + 1u, y, z, v + a.R12, 0u, hi, nhi);  // from +x, hopped -x
    f.tight = occ + nlo + nhi > uint32_t(S);
    uint32_t cnt = 0;
    Slots<S> out = assemble<S>(own, occ, stay, lo, nlo, hi, nhi, cnt);
    react<S>(a, tb, out, cnt, x, y, z, f.fired, f.rolled, f.near, f.bad);
    return out;
}

// ---- exact step where the gather cannot be exact ------------------------------------------------------------
// The reference keeps, beside the lattice, an occupancy word O[v] and per-species counts C[v][sp] for every
// voxel (rdme.wgsl:61-71, cme.wgsl:85-114). Here both are read off the byte slots (O = occupied slots, C[sp] =
// slots holding sp) except at "skewed" voxels, where the reference's words leave that picture and the state
// keeps the difference:
//  * the slot-S alias (rdme.wgsl:30-40, 61-62): a claim into a full voxel d lands in slot 0 of voxel d+1; O[d]
//    and C[d][sp] count it while the slots of d+1 hold it;
//  * a sweep walks O[w] slots from w*S (rdme.wgsl:212-213): past slot S it moves the next voxel's particles and
//    counts them off w; below the occupied slots it leaves the last particles where they are;
//  * a reaction can take a count below zero (k*n^2 of a homodimer row fires on one particle, cme.wgsl:42-44,
//    85-91), and the negative count, which no slot holds, enters later propensities;
//  * a reservoir site holds one unit of C[v][R] in no slot (lattice.rs:292-300), and the reaction skips the
//    species K = reservoirs[v*S] (cme.wgsl:80), whose slots it drops while C[v][K] stays.
// A skewed voxel has an entry (occupancy offset, up to kSkewSp count offsets, K) in an open-addressing table
// keyed by its voxel index; entries are never removed, a voxel whose offsets return to zero keeps a clean one.
// A bitmap marks the voxels with a live entry: the step kernels leave their reaction to the replay kernel.
//
// The step kernels gather every voxel as if nothing were tight or skewed, and list the voxels where that may
// be wrong: "tight" ones (own particles + intended arrivals > S, so a claim can find them full), and those whose
// reaction takes a count below zero. One block after the step (tight_replay_kernel, or ens_epilogue for an
// ensemble) adds every voxel whose occupancy word is off its slot count, sorts the list and cuts it into
// clusters at gaps of kClusterGap planes along x. Each cluster is replayed by one lane, clusters side by side:
// F = the listed voxels, their six neighbours and the voxel after each in memory (slot S aliases its slot 0)
// is the set whose outcome may differ from the gather; the sources W = F and its neighbours act on F. The lane
// runs the sweeps of W in canonical order (voxels ascending, slots ascending; the scatter form of
// rdme.wgsl:45-88) on private copies of F's slots, occupancy words and count offsets, then the reaction
// (cme.wgsl:25-120) and the compaction of each voxel of F, and overwrites F's output; every voxel outside F was
// exact already. The replay checks the assumptions that make this hold, and a voxel outside F that turns out
// to matter (a source whose claim into F fails, a slot-S alias into the voxel after F, a walk into a block
// outside F) is added to F and the cluster replayed again. Clusters kClusterGap planes apart cannot reach each
// other: F extends a cluster by at most kEscX planes, its sources one more.
constexpr int kSkewSp = 4;                  // species with a count offset one voxel can hold
constexpr uint32_t kSkewFree = 0xFFFFFFFFu;  // an unused key
constexpr uint32_t kSkewStatic = 0x100u;     // a reservoir voxel: live for good
constexpr uint32_t kSkewDone = 0x200u;       // replayed this step: its reaction has run
constexpr uint32_t kReplayThreads = 256;
constexpr uint32_t kSkewMargin = kReplayThreads;  // inserts stop this far below the entry capacity
constexpr uint32_t kSortLds = 8192;          // noted voxels sorted in LDS (more: in global memory)
constexpr int kEscMax = 16;                  // voxels a cluster may add to F
constexpr int kEscX = 3;                     // planes F may reach beyond its cluster's noted voxels
constexpr uint32_t kClusterGap = 2u * kEscX + 2u;

struct SkewEntry {
    uint32_t v;               // voxel index among all voxels of the device state
    int32_t off;              // occupancy word - occupied slots (u32 arithmetic in the reference)
    uint32_t sp;              // four species bytes (0: unused) whose count differs from the slots
    uint32_t kf;              // bits 0-7: K = reservoirs[v*S]; kSkewStatic, kSkewDone
    int32_t dc[kSkewSp];      // count - slots holding the species of byte i
};

struct SkewTab {
    uint32_t *keys;   // [mask+1] voxel index or kSkewFree (linear probing)
    uint32_t *idx;    // [mask+1] entry of that key
    SkewEntry *ent;   // [ecap] entries in insertion order
    uint32_t *ctl;    // [0] entries, [1] inserts refused (table full)
    uint32_t *live;   // bit per voxel: a live entry, whose reaction the replay kernel runs
    uint32_t mask, ecap;
};

__host__ __device__ inline uint32_t skew_hash(uint32_t v)
{
    v ^= v >> 16;
    v *= 0x85EBCA6Bu;
    v ^= v >> 13;
    v *= 0xC2B2AE35u;
    v ^= v >> 16;
    return v;
}

__device__ inline int skew_find(const SkewTab &k, uint32_t v)
{
    uint32_t h = skew_hash(v) & k.mask;
    for (uint32_t p = 0; p <= k.mask; ++p) {
        const uint32_t key = k.keys[h];
        if (key == v)
            return int(k.idx[h]);
        if (key == kSkewFree)
            return -1;
        h = (h + 1u) & k.mask;
    }
    return -1;
}

// The entry of voxel v, created clean when missing; -1 when the table is full. Keys are never removed, so a
// probe for a key present stops at it; concurrent inserts are of different voxels.
__device__ inline int skew_insert(const SkewTab &k, uint32_t v)
{
    uint32_t h = skew_hash(v) & k.mask;
    for (uint32_t p = 0; p <= k.mask; ++p) {
        uint32_t key = k.keys[h];
        if (key == v)
            return int(k.idx[h]);
        if (key == kSkewFree) {
            if (atomicAdd(k.ctl, 0u) + kSkewMargin >= k.ecap) {
                atomicAdd(k.ctl + 1, 1u);
                return -1;
            }
            key = atomicCAS(k.keys + h, kSkewFree, v);
            if (key == kSkewFree) {
                const uint32_t i = atomicAdd(k.ctl, 1u);
                SkewEntry e;
                e.v = v;
                e.off = 0;
                e.sp = 0u;
                e.kf = 0u;
                for (int q = 0; q < kSkewSp; ++q)
                    e.dc[q] = 0;
                k.ent[i] = e;
                k.idx[h] = i;
                return int(i);
            }
            if (key == v)
                return int(k.idx[h]);
        }
        h = (h + 1u) & k.mask;
    }
    return -1;
}

__device__ __forceinline__ bool skew_live(const SkewEntry &e)
{
    return e.off != 0 || e.sp != 0u || (e.kf & (0xFFu | kSkewStatic)) != 0u;
}

__device__ __forceinline__ void set_voxel_bit(uint32_t *bits, uint32_t v, bool on)
{
    const uint32_t b = 1u << (v & 31u);
    if (on)
        atomicOr(bits + (v >> 5), b);
    else
        atomicAnd(bits + (v >> 5), ~b);
}

// Count offsets: species bytes `sp` with values dc[0..kSkewSp).
__device__ __forceinline__ int32_t dc_get(uint32_t sp, const int32_t *dc, uint32_t s)
{
    int32_t d = 0;
    for (int q = 0; q < kSkewSp; ++q)
        if (((sp >> (8 * q)) & 0xFFu) == s && s != 0u)
            d = dc[q];
    return d;
}

// Adds d to the offset of species s; false when all kSkewSp places hold other species.
__device__ inline bool dc_add(uint32_t &sp, int32_t *dc, uint32_t s, int32_t d)
{
    if (d == 0 || s == 0u)
        return true;
    for (int q = 0; q < kSkewSp; ++q)
        if (((sp >> (8 * q)) & 0xFFu) == s) {
            dc[q] += d;
            if (dc[q] == 0)
                sp &= ~(0xFFu << (8 * q));
            return true;
        }
    for (int q = 0; q < kSkewSp; ++q)
        if (((sp >> (8 * q)) & 0xFFu) == 0u) {
            sp |= s << (8 * q);
            dc[q] = d;
            return true;
        }
    return false;
}

template <int S>
__device__ __forceinline__ Slots<S> set_byte(const Slots<S> &s, uint32_t j, uint32_t val)
{
    return from_bv<S>(bv_set(to_bv<S>(s), j, val));
}

// Stable partition of the non-zero slots to the front (final_stage.wgsl:25-33).
template <int S>
__device__ __forceinline__ Slots<S> compact(const Slots<S> &s)
{
    Slots<S> out = zero_slots<S>();
    uint32_t j = 0;
    for (uint32_t i = 0; i < uint32_t(S); ++i) {
        const uint32_t b = get_byte_dyn(s, i);
        if (b) {
            put_byte_dyn(out, j, b);
            ++j;
        }
    }
    return out;
}

// One voxel with the reference's counters: its block of slots, occupancy word and count offsets.
template <int S>
struct XVox {
    Slots<S> sl;
    uint32_t occ;
    uint32_t sp;
    int32_t dc[kSkewSp];
};

// cme.wgsl:25-120 then final_stage.wgsl:25-33 on one voxel, with C[s] = slots holding s + offset, C[0] = 1, the
// occupancy word, and K (reservoirs[v*S]) skipped by the rewrite. The voxel leaves compacted, its offsets
// updated: after a rewrite K keeps its whole count in no slot and every other species keeps only a count below
// zero; a rollback subtracts the stoichiometry up to the species being written, K included (cme.wgsl:102-106).
// `bad` is set when the offsets need more than kSkewSp species.
template <int S>
__device__ void exact_react(const StepArgs &a, XVox<S> &vx, uint32_t K, uint32_t x, uint32_t y, uint32_t z, bool &fired,
                            bool &rolled, bool &near, bool &bad)
{
    fired = rolled = near = false;
    auto C = [&](uint32_t s) -> int32_t {
        return s == 0u ? 1 : int32_t(count_species(vx.sl, s)) + dc_get(vx.sp, vx.dc, s);
    };
    auto prop = [&](uint32_t r) -> float {
        const uint32_t ix = reinterpret_cast<const uint32_t *>(a.ridx)[r];
        const uint32_t p = uint32_t(C(ix & 0xFFu)) * uint32_t(C((ix >> 8) & 0xFFu)) * uint32_t(C((ix >> 16) & 0xFFu));
        return a.rates[r] * float(int32_t(p));  // i32 product (cme.wgsl:42-44), wrapping
    };
    float tot = 0.0f;
    for (uint32_t r = 1; r <= a.nr; ++r)
        tot = tot + prop(r);
    const uint32_t P = pcg_hash(a.itime + x * y + z);
    const float r1 = uniform_float(P);
    const float xt = tot * a.tau;
    const float thr = 1.0f - float(exp(double(-xt)));
    near = xt != 0.0f && fabsf(r1 - thr) <= 5.9604644775390625e-08f;  // exp(-0) = 1 exactly
    if (r1 <= thr) {
        fired = true;
        // first i with r2 <= cum[i] / tot, i < nr (cme.wgsl:50-52, 68-73): the same f32 sums in the same order
        const float r2 = uniform_float(P + 1u);
        uint32_t i = 0;
        float cum = 0.0f;
        while (i < a.nr && r2 > cum / tot) {
            ++i;
            cum = cum + prop(i);
        }
        const int8_t *row = a.stoich + size_t(i) * (a.ns + 1u);
        Slots<S> out = zero_slots<S>();
        uint32_t j = 0, sp_stop = 0;
        int32_t occ_new = int32_t(vx.occ);
        for (uint32_t s = 1; s <= a.ns; ++s) {
            if (s == K)
                continue;  // reservoirs[idx_lattice] == species (cme.wgsl:80-83)
            const int32_t c = C(s) + int32_t(row[s]);
            occ_new += int32_t(row[s]);
            if (c > 0) {
                if (j + uint32_t(c) >= uint32_t(S)) {  // the write that reaches S slots rolls back (cme.wgsl:95)
                    sp_stop = s;
                    break;
                }
                for (int32_t t = 0; t < c; ++t)
                    put_byte_dyn(out, j++, s);
            }
        }
        if (sp_stop != 0u) {
            rolled = true;
            if (K != 0u && K <= sp_stop && !dc_add(vx.sp, vx.dc, K, -int32_t(row[K])))
                bad = true;
        } else {
            uint32_t nsp = 0u;
            int32_t ndc[kSkewSp] = {0, 0, 0, 0};
            for (uint32_t s = 1; s <= a.ns; ++s) {
                const int32_t c = C(s);
                const int32_t d = s == K ? c : min(c + int32_t(row[s]), 0);
                if (!dc_add(nsp, ndc, s, d))
                    bad = true;
            }
            vx.sl = out;
            vx.occ = uint32_t(occ_new);
            vx.sp = nsp;
            for (int q = 0; q < kSkewSp; ++q)
                vx.dc[q] = ndc[q];
        }
    }
    vx.sl = compact<S>(vx.sl);
}

__device__ __forceinline__ int find_sorted(const uint32_t *arr, uint32_t n, uint32_t v)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2u;
        if (arr[mid] < v)
            lo = mid + 1u;
        else
            hi = mid;
    }
    return (lo < n && arr[lo] == v) ? int(lo) : -1;
}

__device__ __forceinline__ uint32_t neighbor_or_none(const StepArgs &a, uint32_t v, int k)
{
    const uint32_t z = v % a.R2, y = (v / a.R2) % a.R1, x = v / a.R12;
    switch (k) {
    case 0: return x > 0u ? v - a.R12 : 0xFFFFFFFFu;
    case 1: return x + 1u < a.R0 ? v + a.R12 : 0xFFFFFFFFu;
    case 2: return y > 0u ? v - a.R2 : 0xFFFFFFFFu;
    case 3: return y + 1u < a.R1 ? v + a.R2 : 0xFFFFFFFFu;
    case 4: return z > 0u ? v - 1u : 0xFFFFFFFFu;
    case 5: return z + 1u < a.R2 ? v + 1u : 0xFFFFFFFFu;
    case 6: return v;
    default: return v + 1u < a.N ? v + 1u : 0xFFFFFFFFu;  // the voxel that slot S of v aliases
    }
}

// Ascending, duplicate-free union of the monotone maps k = 0 .. nstreams-1 of neighbor_or_none over the sorted
// list `in`, merged with the sorted list `extra`. Writes it to `out` when given; returns its length.
__device__ uint32_t merge_hood(const StepArgs &a, const uint32_t *in, uint32_t n, int nstreams, const uint32_t *extra,
                               uint32_t n_extra, uint32_t *out)
{
    uint32_t pos[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t pe = 0, m = 0, last = 0xFFFFFFFFu;
    for (;;) {
        uint32_t best = pe < n_extra ? extra[pe] : 0xFFFFFFFFu;
        for (int k = 0; k < nstreams; ++k) {
            while (pos[k] < n && neighbor_or_none(a, in[pos[k]], k) == 0xFFFFFFFFu)
                ++pos[k];
            if (pos[k] < n)
                best = min(best, neighbor_or_none(a, in[pos[k]], k));
        }
        if (best == 0xFFFFFFFFu)
            return m;
        for (int k = 0; k < nstreams; ++k)
            if (pos[k] < n && neighbor_or_none(a, in[pos[k]], k) == best)
                ++pos[k];
        if (pe < n_extra && extra[pe] == best)
            ++pe;
        if (best != last) {
            if (out)
                out[m] = best;
            ++m;
            last = best;
        }
    }
}

__device__ __forceinline__ void counter_add(unsigned long long *c, int idx, long long d)
{
    if (d)
        atomicAdd(c + idx, (unsigned long long)d);
}

// Scratch of the replay kernel (one set per device state).
struct ReplayBuf {
    uint32_t *ctl;      // [0] voxels noted this step, [2] F pool used, [kFixQueue..] the wave kernel's queues
    uint32_t *T;        // [tcap] noted voxels: by the step kernels (note_tight), then the skewed ones; sorted here
    uint32_t tcap;      // power of two
    uint4 *xclus;       // [ccap] runs of noted voxels along x: first, count, cell, lowest plane
    uint4 *clus;        // [ccap] the clusters replayed: first noted voxel, count, cell, lowest plane
    uint32_t ccap;
    uint32_t *mark_lo;  // bit per voxel, and
    uint32_t *mark_hi;  // bit per 1024 voxels: a list longer than lds_sort_max is sorted through them (left clear)
    uint32_t nvox;
    uint32_t lds_sort_max;  // longest list sorted in LDS (kSortLds; lowered by a test hook)
    uint32_t wave_max;      // clusters of up to this many noted voxels are replayed by one wave (kWaveReplayT),
    uint32_t block_max;     // up to this many by the whole block (kBlockReplayT), larger ones by one lane
    uint32_t *F;        // [fcap] the voxels each cluster replays, ascending
    uint8_t *FL;        // [fcap * S] their slots during the replay
    uint32_t *FO;       // [fcap] their occupancy words during the replay
    uint32_t *FW;       // [fcap] their occupancy words when the step began: the sweep's walk (rdme.wgsl:212-213)
    uint32_t *FS;       // [fcap] species bytes of their count offsets
    int32_t *FD;        // [fcap * kSkewSp] the count offsets
    uint32_t fcap;
};

struct ReplayTally {
    long long fired = 0, rolled = 0, near = 0, bad = 0, failed = 0, alias = 0;
};

// Replays cluster c (see the section comment). Returns false when the cluster could not be replayed exactly
// (scratch or table exhausted, or F reaching further than kEscX planes): its gathered output then stands.
template <int S>
__device__ bool replay_cluster(const StepArgs &a, const SkewTab &k, const ReplayBuf &b, const uint32_t *T, uint32_t nt,
                               uint32_t xlo_t, ReplayTally &tally)
{
    const Tab tb = global_tab(a);
    const uint32_t base = a.vox_base;
    const int xlo = int(xlo_t) - kEscX, xhi = int(T[nt - 1u] / a.R12) + kEscX;
    uint32_t ext[kEscMax];
    uint32_t n_ext = 0;
    const uint32_t nF0 = merge_hood(a, T, nt, 8, ext, 0u, nullptr);
    const uint32_t fb = atomicAdd(b.ctl + 2, nF0 + uint32_t(kEscMax));
    if (fb + nF0 + uint32_t(kEscMax) > b.fcap)
        return false;
    uint32_t *F = b.F + fb;
    uint8_t *FL = b.FL + size_t(fb) * S;
    uint32_t *FO = b.FO + fb, *FW = b.FW + fb, *FS = b.FS + fb;
    int32_t *FD = b.FD + size_t(fb) * kSkewSp;
    uint32_t nF = 0;
    long long failed = 0, alias = 0;
    bool done = false, bad = false;
    for (int round = 0; round <= kEscMax && !done && !bad; ++round) {
        nF = merge_hood(a, T, nt, 8, ext, n_ext, F);
        if (int(F[0] / a.R12) < xlo || int(F[nF - 1u] / a.R12) > xhi)
            return false;
        for (uint32_t i = 0; i < nF; ++i) {
            const Slots<S> s = load_slots<S>(a.src, F[i]);
            for (int q = 0; q < S; ++q)
                FL[size_t(i) * S + q] = uint8_t(get_byte(s, q));
            const int ei = skew_find(k, base + F[i]);
            const SkewEntry e = ei >= 0 ? k.ent[ei] : SkewEntry{0u, 0, 0u, 0u, {0, 0, 0, 0}};
            FO[i] = FW[i] = count_nonzero(s) + uint32_t(e.off);
            FS[i] = e.sp;
            for (int q = 0; q < kSkewSp; ++q)
                FD[size_t(i) * kSkewSp + q] = e.dc[q];
        }
        failed = alias = 0;
        uint32_t esc = 0xFFFFFFFFu;
        // sources W = F and its six neighbours, ascending (the sweep order of rdme.wgsl:194-229)
        uint32_t pos[7] = {0, 0, 0, 0, 0, 0, 0};
        for (;;) {
            uint32_t w = 0xFFFFFFFFu;
            for (int q = 0; q < 7; ++q) {
                while (pos[q] < nF && neighbor_or_none(a, F[pos[q]], q) == 0xFFFFFFFFu)
                    ++pos[q];
                if (pos[q] < nF)
                    w = min(w, neighbor_or_none(a, F[pos[q]], q));
            }
            if (w == 0xFFFFFFFFu)
                break;
            for (int q = 0; q < 7; ++q)
                if (pos[q] < nF && neighbor_or_none(a, F[pos[q]], q) == w)
                    ++pos[q];
            const int wF = find_sorted(F, nF, w);
            const Slots<S> sw = load_slots<S>(a.src, w);
            // walk: occupancy word at the step's start, with the reference's u32 end index (an end at or below
            // the start walks nothing)
            const uint32_t occw = wF >= 0 ? FW[wF] : count_nonzero(sw);
            const uint32_t start = w * uint32_t(S), end = start + occw;
            uint32_t walk = end > start ? end - start : 0u;
            if (walk == 0u)
                continue;
            if (walk > 2u * uint32_t(S)) {  // a word wrapped below zero at the first voxels: unrepresentable
                bad = true;
                walk = 2u * uint32_t(S);
            }
            const uint32_t z = w % a.R2, y = (w / a.R2) % a.R1, x = w / a.R12;
            const Hood h = make_hood(a, x, y, z, w, S);
            const Slots<S> sn = (walk > uint32_t(S) && w + 1u < a.N) ? load_slots<S>(a.src, w + 1u) : zero_slots<S>();
            for (uint32_t s = 0; s < walk && esc == 0xFFFFFFFFu; ++s) {
                const uint32_t blk = w + s / uint32_t(S), j = s % uint32_t(S);
                const uint32_t sp = blk >= a.N ? 0u : (blk == w ? get_byte_dyn(sw, j) : get_byte_dyn(sn, j));
                if (sp == 0u)
                    continue;
                const uint32_t d = hop(a, tb, h, sp, s);
                if (d >= 6u)
                    continue;
                const uint32_t t = neighbor_or_none(a, w, int(d));
                const int tF = find_sorted(F, nF, t);
                const int bF = blk == w ? wF : find_sorted(F, nF, blk);
                if (tF < 0 && wF < 0)
                    continue;  // neither end in F: the gather has it
                if (tF >= 0) {
                    const uint32_t old = FO[tF];
                    FO[tF] = old + 1u;
                    bool ok = false;
                    if (old <= uint32_t(S)) {  // rdme.wgsl:62: admits a claim into a voxel holding S
                        for (int q = 0; q < S && !ok; ++q)
                            if (FL[size_t(tF) * S + q] == 0u) {
                                FL[size_t(tF) * S + q] = uint8_t(sp);
                                ok = true;
                            }
                        if (!ok && t + 1u < a.N) {  // slot S: slot 0 of the next voxel (past the lattice: taken)
                            const int t1 = find_sorted(F, nF, t + 1u);
                            if (t1 < 0) {
                                esc = t + 1u;
                                break;
                            }
                            if (FL[size_t(t1) * S] == 0u) {
                                FL[size_t(t1) * S] = uint8_t(sp);
                                ok = true;
                                ++alias;
                                // C[t][sp] counts it, the slots of t+1 hold it
                                bad |= !dc_add(FS[tF], FD + size_t(tF) * kSkewSp, sp, 1);
                                bad |= !dc_add(FS[t1], FD + size_t(t1) * kSkewSp, sp, -1);
                            }
                        }
                    }
                    if (!ok) {
                        FO[tF] = old;
                        ++failed;
                        if (wF < 0) {  // a source outside F keeps a particle the gather let go
                            esc = w;
                            break;
                        }
                        continue;
                    }
                }
                if (wF < 0)
                    continue;
                // the source: O[w] -= 1, lattice[i_part] = 0, C[w][sp] -= 1 (rdme.wgsl:69-71)
                FO[wF] -= 1u;
                if (bF < 0) {  // a walk past slot S into a block outside F
                    esc = blk;
                    break;
                }
                const uint32_t cur = FL[size_t(bF) * S + j];
                FL[size_t(bF) * S + j] = 0u;
                if (!(bF == wF && cur == sp)) {
                    bad |= !dc_add(FS[bF], FD + size_t(bF) * kSkewSp, cur, 1);
                    bad |= !dc_add(FS[wF], FD + size_t(wF) * kSkewSp, sp, -1);
                }
            }
            if (esc != 0xFFFFFFFFu)
                break;
        }
        if (esc == 0xFFFFFFFFu) {
            done = true;
        } else {
            if (n_ext >= uint32_t(kEscMax))
                return false;
            uint32_t q = n_ext++;  // keep ext sorted
            while (q > 0u && ext[q - 1u] > esc) {
                ext[q] = ext[q - 1u];
                --q;
            }
            ext[q] = esc;
        }
    }
    if (!done)
        return false;
    // the counts the step kernel took for F's gathered outputs
    for (uint32_t i = 0; i < nF; ++i) {
        VoxelFlags g;
        (void)gather_voxel<S>(a, tb, F[i], g);
        tally.fired -= g.fired;
        tally.rolled -= g.rolled;
        tally.near -= g.near;
    }
    // reaction and compaction of F; outputs, offsets and live bits
    for (uint32_t i = 0; i < nF; ++i) {
        const uint32_t v = F[i];
        int ei = skew_find(k, base + v);
        const uint32_t kf = ei >= 0 ? k.ent[ei].kf : 0u;
        XVox<S> vx;
        vx.sl = zero_slots<S>();
        for (int q = 0; q < S; ++q)
            put_byte_dyn(vx.sl, uint32_t(q), FL[size_t(i) * S + q]);
        vx.occ = FO[i];
        vx.sp = FS[i];
        for (int q = 0; q < kSkewSp; ++q)
            vx.dc[q] = FD[size_t(i) * kSkewSp + q];
        bool fired, rolled, near, xbad = false;
        exact_react<S>(a, vx, kf & 0xFFu, v / a.R12, (v / a.R2) % a.R1, v % a.R2, fired, rolled, near, xbad);
        tally.fired += fired;
        tally.rolled += rolled;
        tally.near += near;
        bad |= xbad;
        store_slots<S>(a.dst, v, vx.sl);
        const int32_t off = int32_t(vx.occ - count_nonzero(vx.sl));
        if (count_nonzero(vx.sl) != 0u)
            sum_mark(a.sum_dst, v);
        if (ei < 0 && (off != 0 || vx.sp != 0u))
            ei = skew_insert(k, base + v);
        if (ei >= 0) {
            SkewEntry &e = k.ent[ei];
            e.off = off;
            e.sp = vx.sp;
            for (int q = 0; q < kSkewSp; ++q)
                e.dc[q] = vx.dc[q];
            e.kf = kf | kSkewDone;
            set_voxel_bit(k.live, base + v, skew_live(e));
        } else if (off != 0 || vx.sp != 0u) {
            bad = true;  // table full
        }
    }
    tally.failed += failed;
    tally.alias += alias;
    tally.bad += bad;
    return true;
}

// The same replay with the cluster's arrays in the LDS of one wave, for clusters of at most kWaveReplayT noted
// voxels (F then holds at most 8 * kWaveReplayT + kEscMax voxels). The lanes load F's state, list the sources W
// and compute their hops side by side, keeping each move with an end in F as an event; the events are sorted
// into the canonical order (source ascending, walk position ascending) and one lane applies them; the lanes then
// run the reactions and write F's outputs. The one-lane form replays a cluster whose events overflow the list.
constexpr uint32_t kWaveReplayT = 14;
constexpr uint32_t kWaveReplayF = 8u * kWaveReplayT + uint32_t(kEscMax);
constexpr uint32_t kWaveReplayW = 7u * kWaveReplayF;
constexpr uint32_t kWaveReplayEv = 1024;
constexpr uint32_t kNoF = 0xFFu;          // an event end outside F
constexpr uint32_t kPastLattice = 0xFEu;  // slot S of the lattice's last voxel: no voxel after it
static_assert(kWaveReplayF < kPastLattice && kWaveReplayW < 1024u, "event fields");

template <int S>
struct WaveReplayLds {
    uint64_t ev[kWaveReplayEv];           // events: (W index << 8 | walk position << 3 | direction) << 32 |
                                          // species | F index of target << 8, of target + 1 << 16, of block << 24
    uint32_t F[kWaveReplayF];
    uint32_t FL[kWaveReplayF * (S / 4)];  // F's slots during the replay
    uint32_t FO[kWaveReplayF];            // F's occupancy words during the replay
    uint32_t FW[kWaveReplayF];            // F's occupancy words when the step began
    uint32_t FS[kWaveReplayF];            // species bytes of F's count offsets
    int32_t FD[kWaveReplayF * kSkewSp];   // the count offsets
    uint32_t W[kWaveReplayW];             // the sources, ascending
    uint8_t WF[kWaveReplayW];             // F index of each source (kNoF: none)
    uint32_t T[kWaveReplayT];             // the noted voxels
    uint32_t ext[kEscMax];                // voxels added to F, ascending
    uint32_t n_ev, n_ext, esc;
};

// Ascending sort of the wave's arr[0..n) by bitonic merging (arr padded to n2, a power of two, with all ones).
template <class K>
__device__ void wave_bitonic(K *arr, uint32_t n, uint32_t n2)
{
    const uint32_t l = __lane_id();
    for (uint32_t i = n + l; i < n2; i += 64u)
        arr[i] = K(~K(0));
    wave_sync();
    for (uint32_t kk = 2; kk <= n2; kk <<= 1)
        for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
            for (uint32_t i = l; i < n2; i += 64u) {
                const uint32_t m = i ^ jj;
                if (m > i) {
                    const K p = arr[i], q = arr[m];
                    if ((i & kk) == 0u ? p > q : p < q) {
                        arr[i] = q;
                        arr[m] = p;
                    }
                }
            }
            wave_sync();
        }
}

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t n)
{
    uint32_t n2 = 1u;
    while (n2 < n)
        n2 <<= 1;
    return n2;
}

// The voxels cand[0..n) (0xFFFFFFFF: none) sorted, without repeats, into out; returns their number (one wave;
// cand is overwritten).
__device__ uint32_t wave_sort_unique(uint32_t *cand, uint32_t n, uint32_t *out)
{
    const uint32_t l = __lane_id();
    wave_bitonic<uint32_t>(cand, n, pow2_at_least(n));
    uint32_t m = 0;
    for (uint32_t b0 = 0; b0 < n; b0 += 64u) {
        const uint32_t i = b0 + l;
        const uint32_t v = i < n ? cand[i] : 0xFFFFFFFFu;
        const bool keep = v != 0xFFFFFFFFu && (i == 0u || cand[i - 1u] != v);
        const uint64_t mk = __ballot(keep);
        if (keep)
            out[m + __builtin_amdgcn_mbcnt_hi(uint32_t(mk >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mk), 0u))] = v;
        m += uint32_t(__popcll(mk));
    }
    wave_sync();
    return m;
}

// neighbor_or_none(a, v, k) for k = 0 .. nk-1 (the six neighbours, v, v + 1), the coordinates formed once.
__device__ __forceinline__ void hood_list(const StepArgs &a, uint32_t v, uint32_t nk, uint32_t *out)
{
    const uint32_t z = v % a.R2, y = (v / a.R2) % a.R1, x = v / a.R12;
    out[0] = x > 0u ? v - a.R12 : 0xFFFFFFFFu;
    out[1] = x + 1u < a.R0 ? v + a.R12 : 0xFFFFFFFFu;
    out[2] = y > 0u ? v - a.R2 : 0xFFFFFFFFu;
    out[3] = y + 1u < a.R1 ? v + a.R2 : 0xFFFFFFFFu;
    out[4] = z > 0u ? v - 1u : 0xFFFFFFFFu;
    out[5] = z + 1u < a.R2 ? v + 1u : 0xFFFFFFFFu;
    out[6] = v;
    if (nk > 7u)
        out[7] = v + 1u < a.N ? v + 1u : 0xFFFFFFFFu;
}

// Called by every lane of a wave with the same cluster. Returns 1 when replayed, 0 when the cluster cannot be
// replayed exactly (as replay_cluster's false: its gathered output stands), -1 when the events overflow the list
// (nothing written: the caller replays it in the one-lane form).
template <int S>
__device__ int replay_cluster_wave(const StepArgs &a, const SkewTab &k, WaveReplayLds<S> &L, const uint32_t *T,
                                   uint32_t nt, uint32_t xlo_t, ReplayTally &tally)
{
    constexpr uint32_t NW = S / 4;
    const uint32_t l = __lane_id();
    const Tab tb = global_tab(a);
    const uint32_t base = a.vox_base;
    const int xlo = int(xlo_t) - kEscX, xhi = int(T[nt - 1u] / a.R12) + kEscX;
    for (uint32_t i = l; i < nt; i += 64u)
        L.T[i] = T[i];
    if (l == 0u)
        L.n_ext = 0u;
    wave_sync();
    long long failed = 0, alias = 0;  // lane 0
    bool bad = false;
    uint32_t nF = 0;
    bool done = false;
    for (int round = 0; round <= kEscMax && !done; ++round) {
        // F: the noted voxels' neighbours, themselves and the voxels after them, and the voxels added
        uint32_t *cand = reinterpret_cast<uint32_t *>(L.ev);
        const uint32_t n_ext = L.n_ext;
        for (uint32_t i = l; i < nt; i += 64u)
            hood_list(a, L.T[i], 8u, cand + 8u * i);
        for (uint32_t i = l; i < n_ext; i += 64u)
            cand[8u * nt + i] = L.ext[i];
        wave_sync();
        nF = wave_sort_unique(cand, 8u * nt + n_ext, L.F);
        if (l == 0u)
            L.n_ev = 0u;
        if (int(L.F[0] / a.R12) < xlo || int(L.F[nF - 1u] / a.R12) > xhi)
            return 0;
        for (uint32_t i = l; i < nF; i += 64u) {
            const uint32_t v = L.F[i];
            const Slots<S> sv = load_slots<S>(a.src, v);
            for (uint32_t q = 0; q < NW; ++q)
                L.FL[i * NW + q] = sv.w[q];
            const int ei = skew_find(k, base + v);
            const SkewEntry e = ei >= 0 ? k.ent[ei] : SkewEntry{0u, 0, 0u, 0u, {0, 0, 0, 0}};
            L.FO[i] = L.FW[i] = count_nonzero(sv) + uint32_t(e.off);
            L.FS[i] = e.sp;
            for (int q = 0; q < kSkewSp; ++q)
                L.FD[i * kSkewSp + q] = e.dc[q];
        }
        for (uint32_t i = l; i < nF; i += 64u)
            hood_list(a, L.F[i], 7u, cand + 7u * i);
        wave_sync();
        const uint32_t nW = wave_sort_unique(cand, 7u * nF, L.W);
        // the moves of the sources W = F and its six neighbours that have an end in F
        for (uint32_t wi = l; wi < nW; wi += 64u) {
            const uint32_t w = L.W[wi];
            const int wF = find_sorted(L.F, nF, w);
            L.WF[wi] = uint8_t(wF >= 0 ? uint32_t(wF) : kNoF);
            const Slots<S> sw = load_slots<S>(a.src, w);
            // the walk: occupancy word at the step's start, with the reference's u32 end index
            const uint32_t occw = wF >= 0 ? L.FW[wF] : count_nonzero(sw);
            const uint32_t start = w * uint32_t(S), end = start + occw;
            uint32_t walk = end > start ? end - start : 0u;
            if (walk == 0u)
                continue;
            if (walk > 2u * uint32_t(S)) {  // a word wrapped below zero at the first voxels: unrepresentable
                bad = true;
                walk = 2u * uint32_t(S);
            }
            const uint32_t z = w % a.R2, y = (w / a.R2) % a.R1, x = w / a.R12;
            const Hood h = make_hood(a, x, y, z, w, S);
            const Slots<S> sn = (walk > uint32_t(S) && w + 1u < a.N) ? load_slots<S>(a.src, w + 1u) : zero_slots<S>();
            for (uint32_t s = 0; s < walk; ++s) {
                const uint32_t blk = w + s / uint32_t(S), j = s % uint32_t(S);
                const uint32_t sp = blk >= a.N ? 0u : (blk == w ? get_byte_dyn(sw, j) : get_byte_dyn(sn, j));
                if (sp == 0u)
                    continue;
                const uint32_t d = hop(a, tb, h, sp, s);
                if (d >= 6u)
                    continue;
                const uint32_t t = neighbor_or_none(a, w, int(d));
                const int tF = find_sorted(L.F, nF, t);
                if (tF < 0 && wF < 0)
                    continue;  // neither end in F: the gather has it
                const int bF = blk == w ? wF : find_sorted(L.F, nF, blk);
                uint32_t t1F = kNoF;
                if (tF >= 0) {
                    if (t + 1u >= a.N) {
                        t1F = kPastLattice;
                    } else {
                        const int f1 = find_sorted(L.F, nF, t + 1u);
                        t1F = f1 >= 0 ? uint32_t(f1) : kNoF;
                    }
                }
                const uint32_t idx = atomicAdd(&L.n_ev, 1u);
                if (idx < kWaveReplayEv)
                    L.ev[idx] = (uint64_t((wi << 8) | (s << 3) | d) << 32) |
                                (sp | (tF >= 0 ? uint32_t(tF) : kNoF) << 8 | t1F << 16 |
                                 (bF >= 0 ? uint32_t(bF) : kNoF) << 24);
            }
        }
        wave_sync();
        const uint32_t n_ev = L.n_ev;
        if (n_ev > kWaveReplayEv)
            return -1;
        if (n_ev > 1u) {
            uint32_t n2 = 1u;
            while (n2 < n_ev)
                n2 <<= 1;
            wave_bitonic<uint64_t>(L.ev, n_ev, n2);
        }
        // the claims and departures in the canonical order (the scatter form of rdme.wgsl:45-88), on one lane
        if (l == 0u) {
            failed = alias = 0;
            uint32_t esc = 0xFFFFFFFFu;
            uint8_t *FLb = reinterpret_cast<uint8_t *>(L.FL);
            for (uint32_t i = 0; i < n_ev; ++i) {
                const uint64_t e = L.ev[i];
                const uint32_t key = uint32_t(e >> 32), lo = uint32_t(e);
                const uint32_t wi = key >> 8, s = (key >> 3) & 31u, d = key & 7u;
                const uint32_t sp = lo & 0xFFu, tF = (lo >> 8) & 0xFFu, t1F = (lo >> 16) & 0xFFu, bF = lo >> 24;
                const uint32_t wF = L.WF[wi];
                if (tF != kNoF) {
                    const uint32_t old = L.FO[tF];
                    L.FO[tF] = old + 1u;
                    bool ok = false;
                    if (old <= uint32_t(S)) {  // rdme.wgsl:62: admits a claim into a voxel holding S
                        for (uint32_t q = 0; q < NW && !ok; ++q) {
                            const uint32_t zb = ~nonzero_bytes(L.FL[tF * NW + q]) & 0x80808080u;
                            if (zb != 0u) {  // the first empty slot
                                FLb[tF * uint32_t(S) + 4u * q + (uint32_t(__ffs(zb) - 1) >> 3)] = uint8_t(sp);
                                ok = true;
                            }
                        }
                        if (!ok && t1F != kPastLattice) {  // slot S: slot 0 of the next voxel
                            if (t1F == kNoF) {
                                esc = neighbor_or_none(a, L.W[wi], int(d)) + 1u;
                                break;
                            }
                            if (FLb[t1F * uint32_t(S)] == 0u) {
                                FLb[t1F * uint32_t(S)] = uint8_t(sp);
                                ok = true;
                                ++alias;
                                // C[t][sp] counts it, the slots of t+1 hold it
                                bad |= !dc_add(L.FS[tF], L.FD + tF * kSkewSp, sp, 1);
                                bad |= !dc_add(L.FS[t1F], L.FD + t1F * kSkewSp, sp, -1);
                            }
                        }
                    }
                    if (!ok) {
                        L.FO[tF] = old;
                        ++failed;
                        if (wF == kNoF) {  // a source outside F keeps a particle the gather let go
                            esc = L.W[wi];
                            break;
                        }
                        continue;
                    }
                }
                if (wF == kNoF)
                    continue;
                // the source: O[w] -= 1, lattice[i_part] = 0, C[w][sp] -= 1 (rdme.wgsl:69-71)
                L.FO[wF] -= 1u;
                if (bF == kNoF) {  // a walk past slot S into a block outside F
                    esc = L.W[wi] + s / uint32_t(S);
                    break;
                }
                const uint32_t at = bF * uint32_t(S) + s % uint32_t(S);
                const uint32_t cur = FLb[at];
                FLb[at] = 0u;
                if (!(bF == wF && cur == sp)) {
                    bad |= !dc_add(L.FS[bF], L.FD + bF * kSkewSp, cur, 1);
                    bad |= !dc_add(L.FS[wF], L.FD + wF * kSkewSp, sp, -1);
                }
            }
            if (esc != 0xFFFFFFFFu && L.n_ext < uint32_t(kEscMax)) {
                uint32_t q = L.n_ext++;  // keep ext sorted
                while (q > 0u && L.ext[q - 1u] > esc) {
                    L.ext[q] = L.ext[q - 1u];
                    --q;
                }
                L.ext[q] = esc;
            } else if (esc != 0xFFFFFFFFu) {
                L.n_ext = uint32_t(kEscMax) + 1u;  // more voxels than F may add
            }
            L.esc = esc;
        }
        wave_sync();
        if (L.esc == 0xFFFFFFFFu)
            done = true;
        else if (L.n_ext > uint32_t(kEscMax) || __ballot(bad) != 0ull)
            return 0;
    }
    if (!done)
        return 0;
    // the counts the step kernel took for F's gathered outputs; reaction and compaction of F; outputs, offsets and
    // live bits (one voxel per lane)
    for (uint32_t i = l; i < nF; i += 64u) {
        const uint32_t v = L.F[i];
        VoxelFlags g;
        (void)gather_voxel<S>(a, tb, v, g);
        tally.fired -= g.fired;
        tally.rolled -= g.rolled;
        tally.near -= g.near;
        int ei = skew_find(k, base + v);
        const uint32_t kf = ei >= 0 ? k.ent[ei].kf : 0u;
        XVox<S> vx;
        for (uint32_t q = 0; q < NW; ++q)
            vx.sl.w[q] = L.FL[i * NW + q];
        vx.occ = L.FO[i];
        vx.sp = L.FS[i];
        for (int q = 0; q < kSkewSp; ++q)
            vx.dc[q] = L.FD[i * kSkewSp + q];
        bool fired, rolled, near, xbad = false;
        exact_react<S>(a, vx, kf & 0xFFu, v / a.R12, (v / a.R2) % a.R1, v % a.R2, fired, rolled, near, xbad);
        tally.fired += fired;
        tally.rolled += rolled;
        tally.near += near;
        bad |= xbad;
        store_slots<S>(a.dst, v, vx.sl);
        const int32_t off = int32_t(vx.occ - count_nonzero(vx.sl));
        if (count_nonzero(vx.sl) != 0u)
            sum_mark(a.sum_dst, v);
        if (ei < 0 && (off != 0 || vx.sp != 0u))
            ei = skew_insert(k, base + v);
        if (ei >= 0) {
            SkewEntry &e = k.ent[ei];
            e.off = off;
            e.sp = vx.sp;
            for (int q = 0; q < kSkewSp; ++q)
                e.dc[q] = vx.dc[q];
            e.kf = kf | kSkewDone;
            set_voxel_bit(k.live, base + v, skew_live(e));
        } else if (off != 0 || vx.sp != 0u) {
            bad = true;  // table full
        }
    }
    const bool any_bad = __ballot(bad) != 0ull;
    if (l == 0u) {
        tally.failed += failed;
        tally.alias += alias;
        tally.bad += any_bad;
    }
    return 1;
}

// Ascending sort of arr[0..n) in place by bitonic merging, the whole block taking part (arr padded to n2, a
// power of two, with 0xFFFFFFFF).
template <class K>
__device__ void block_bitonic(K *arr, uint32_t n, uint32_t n2)
{
    for (uint32_t i = n + threadIdx.x; i < n2; i += blockDim.x)
        arr[i] = K(~K(0));
    __syncthreads();
    for (uint32_t kk = 2; kk <= n2; kk <<= 1)
        for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
            for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x) {
                const uint32_t l = i ^ jj;
                if (l > i) {
                    const K p = arr[i], q = arr[l];
                    const bool up = (i & kk) == 0u;
                    if (up ? p > q : p < q) {
                        arr[i] = q;
                        arr[l] = p;
                    }
                }
            }
            __syncthreads();
        }
}

// The same replay by the whole block, for clusters of more than kWaveReplayT and at most kBlockReplayT noted
// voxels whose F holds at most kBlockReplayF voxels (W then at most seven times that) and whose sources make at
// most kBlockReplayEv events, the arrays in the block's LDS: the threads load F's state, sort F and W, compute the
// hops of W and keep the moves with an end in F as events, sort the events into the canonical order, and thread 0
// applies them; the threads then run the reactions and write F's outputs. F indices take 16 bits here: an event
// carries the source's W index, the walk position, the direction, the species and the target's F index, and the
// applying thread finds the voxel after the target and a walk's block past slot S in F by binary search (both
// rare). Returns as replay_cluster_wave (-1 also when F or the events pass their capacity).
constexpr uint32_t kBlockReplayT = 512;
constexpr uint32_t kBlockReplayF = 1024;
constexpr uint32_t kBlockReplayW = 7u * kBlockReplayF;
constexpr uint32_t kBlockReplayEv = 4096;
constexpr uint32_t kBNoF = 0xFFFFu;
constexpr uint32_t kBEvNoF = 0xFFFu;  // the same in a 12-bit event field
static_assert(kBlockReplayW < (1u << 24) && kBlockReplayF < kBEvNoF, "block replay event fields");
static_assert(kBlockReplayEv * sizeof(uint64_t) >= 8192u * sizeof(uint32_t) && 8192u >= kBlockReplayW &&
                  8192u >= 8u * kBlockReplayT + kEscMax,
              "the candidate lists (padded to powers of two) fit the event list");

template <int S>
struct BlockReplayLds {
    uint64_t ev[kBlockReplayEv];            // events: (W index << 8 | walk position << 3 | direction) << 32 |
                                            // species | F index of the target << 8 | of the source << 20
                                            // (kBEvNoF: none); first the scratch of block_hood_set
    uint32_t F[kBlockReplayF];
    uint32_t FL[kBlockReplayF * (S / 4)];   // F's slots during the replay
    uint32_t FO[kBlockReplayF];             // F's occupancy words during the replay
    uint32_t FW[kBlockReplayF];             // F's occupancy words when the step began
    uint32_t FS[kBlockReplayF];             // species bytes of F's count offsets
    int32_t FD[kBlockReplayF * kSkewSp];    // the count offsets
    uint32_t W[kBlockReplayW];              // the sources, ascending
    uint32_t T[kBlockReplayT];              // the noted voxels
    uint32_t ext[kEscMax];                  // voxels added to F, ascending
    uint32_t scan[kReplayThreads / 64u];    // per-wave sums of block_offset
    uint32_t box[6];                        // bounding box of block_hood_set: x, y, z minima and maxima
    uint32_t n_ev, n_ext, esc, n_out;
};

// Exclusive prefix sum of c over the block (threads in order); total gets the block's sum. Every thread calls it.
__device__ uint32_t block_offset(uint32_t c, uint32_t *wave_sums, uint32_t &total)
{
    const uint32_t l = __lane_id(), wv = threadIdx.x / 64u, nw = blockDim.x / 64u;
    uint32_t x = c;
    for (uint32_t d = 1; d < 64u; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (l >= d)
            x += y;
    }
    if (l == 63u)
        wave_sums[wv] = x;
    __syncthreads();
    uint32_t off = x - c;
    total = 0u;
    for (uint32_t w = 0; w < nw; ++w) {
        const uint32_t t = wave_sums[w];
        off += w < wv ? t : 0u;
        total += t;
    }
    __syncthreads();
    return off;
}

// Sorts cand[0, n) (room for pow2_at_least(n) entries), drops repeats and 0xFFFFFFFF, writes the rest to out in
// ascending order (the first cap of them); returns their number. Every thread calls it.
__device__ uint32_t block_sort_unique(uint32_t *cand, uint32_t n, uint32_t *out, uint32_t cap, uint32_t *wave_sums)
{
    block_bitonic<uint32_t>(cand, n, pow2_at_least(n));
    const uint32_t nth = blockDim.x, per = (n + nth - 1u) / nth;
    const uint32_t b0 = min(threadIdx.x * per, n), e0 = min(b0 + per, n);
    uint32_t c = 0;
    for (uint32_t i = b0; i < e0; ++i)
        c += cand[i] != 0xFFFFFFFFu && (i == 0u || cand[i - 1u] != cand[i]);
    uint32_t total;
    uint32_t o = block_offset(c, wave_sums, total);
    for (uint32_t i = b0; i < e0; ++i)
        if (cand[i] != 0xFFFFFFFFu && (i == 0u || cand[i - 1u] != cand[i])) {
            if (o < cap)
                out[o] = cand[i];
            ++o;
        }
    __syncthreads();
    return total;
}

// The voxels hood_list(a, src[i], nk) for i < n, and extra[0, n_extra), sorted and without repeats (0xFFFFFFFF
// entries dropped), written to out (the first cap of them); returns their number. They are marked in a bitmap over
// their bounding box, whose row-major order is the lattice's, in `bits` (nwords words of scratch), and read back in
// order; a box with more voxels than the bitmap has bits takes the sort instead (the candidates in `bits`, which
// then holds n * nk + n_extra of them). A sort of the candidates costs a bitonic network over the next power of
// two (91 stages with a block barrier each for 4096 candidates); the bitmap, a few passes. Every thread calls it.
__device__ uint32_t block_hood_set(const StepArgs &a, const uint32_t *src, uint32_t n, uint32_t nk,
                                   const uint32_t *extra, uint32_t n_extra, uint32_t *out, uint32_t cap,
                                   uint32_t *bits, uint32_t nwords, uint32_t *box, uint32_t *wave_sums)
{
    const uint32_t tid = threadIdx.x, nth = blockDim.x;
    if (tid < 6u)
        box[tid] = (tid & 1u) ? 0u : 0xFFFFFFFFu;
    __syncthreads();
    uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
    auto take = [&](uint32_t v) {
        if (v == 0xFFFFFFFFu)
            return;
        const uint32_t c[3] = {v / a.R12, (v / a.R2) % a.R1, v % a.R2};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = min(lo[d], c[d]);
            hi[d] = max(hi[d], c[d]);
        }
    };
    uint32_t h[8];
    for (uint32_t i = tid; i < n; i += nth) {
        hood_list(a, src[i], nk, h);
        for (uint32_t q = 0; q < nk; ++q)
            take(h[q]);
    }
    for (uint32_t i = tid; i < n_extra; i += nth)
        take(extra[i]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        for (int o = 32; o >= 1; o >>= 1) {
            lo[d] = min(lo[d], uint32_t(__shfl_xor(int(lo[d]), o)));
            hi[d] = max(hi[d], uint32_t(__shfl_xor(int(hi[d]), o)));
        }
    }
    if (__lane_id() == 0u) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            atomicMin(&box[2 * d], lo[d]);
            atomicMax(&box[2 * d + 1], hi[d]);
        }
    }
    __syncthreads();
    const uint32_t x0 = box[0], y0 = box[2], z0 = box[4];
    if (x0 > box[1]) {  // no voxel at all
        __syncthreads();
        return 0u;
    }
    const uint32_t BY = box[3] - y0 + 1u, BZ = box[5] - z0 + 1u;
    const uint64_t vol = uint64_t(box[1] - x0 + 1u) * BY * BZ;
    __syncthreads();  // every thread has read the box
    if (vol > 32ull * nwords) {
        for (uint32_t i = tid; i < n; i += nth)
            hood_list(a, src[i], nk, bits + nk * i);
        for (uint32_t i = tid; i < n_extra; i += nth)
            bits[nk * n + i] = extra[i];
        __syncthreads();
        return block_sort_unique(bits, nk * n + n_extra, out, cap, wave_sums);
    }
    const uint32_t nw = uint32_t((vol + 31u) / 32u);
    for (uint32_t i = tid; i < nw; i += nth)
        bits[i] = 0u;
    __syncthreads();
    auto mark = [&](uint32_t v) {
        if (v == 0xFFFFFFFFu)
            return;
        const uint32_t idx = ((v / a.R12 - x0) * BY + ((v / a.R2) % a.R1 - y0)) * BZ + (v % a.R2 - z0);
        atomicOr(&bits[idx >> 5], 1u << (idx & 31u));
    };
    for (uint32_t i = tid; i < n; i += nth) {
        hood_list(a, src[i], nk, h);
        for (uint32_t q = 0; q < nk; ++q)
            mark(h[q]);
    }
    for (uint32_t i = tid; i < n_extra; i += nth)
        mark(extra[i]);
    __syncthreads();
    // each thread reads back a run of words, placed by one prefix sum of their bit counts
    const uint32_t per = (nw + nth - 1u) / nth, w0 = min(tid * per, nw), w1 = min(w0 + per, nw);
    uint32_t c = 0u;
    for (uint32_t w = w0; w < w1; ++w)
        c += uint32_t(__popc(bits[w]));
    uint32_t total;
    uint32_t o = block_offset(c, wave_sums, total);
    for (uint32_t w = w0; w < w1; ++w) {
        uint32_t m = bits[w];
        while (m != 0u) {
            const uint32_t idx = w * 32u + uint32_t(__ffs(m) - 1);
            m &= m - 1u;
            const uint32_t bz = idx % BZ, t = idx / BZ, by = t % BY, bx = t / BY;
            if (o < cap)
                out[o] = ((x0 + bx) * a.R1 + (y0 + by)) * a.R2 + (z0 + bz);
            ++o;
        }
    }
    __syncthreads();
    return total;
}

template <int S>
__device__ int replay_cluster_block(const StepArgs &a, const SkewTab &k, BlockReplayLds<S> &L, const uint32_t *T,
                                    uint32_t nt, uint32_t xlo_t, ReplayTally &tally)
{
    constexpr uint32_t NW = S / 4;
    const uint32_t tid = threadIdx.x, nth = blockDim.x;
    const Tab tb = global_tab(a);
    const uint32_t base = a.vox_base;
    const int xlo = int(xlo_t) - kEscX, xhi = int(T[nt - 1u] / a.R12) + kEscX;
    for (uint32_t i = tid; i < nt; i += nth)
        L.T[i] = T[i];
    if (tid == 0u)
        L.n_ext = 0u;
    __syncthreads();
    long long failed = 0, alias = 0;  // thread 0
    bool bad = false;
    uint32_t nF = 0;
    bool done = false;
    for (int round = 0; round <= kEscMax && !done; ++round) {
        // F: the noted voxels' neighbours, themselves and the voxels after them, and the voxels added
        uint32_t *cand = reinterpret_cast<uint32_t *>(L.ev);
        const uint32_t n_ext = L.n_ext;
        nF = block_hood_set(a, L.T, nt, 8u, L.ext, n_ext, L.F, kBlockReplayF, cand, 2u * kBlockReplayEv, L.box,
                            L.scan);
        if (nF > kBlockReplayF)
            return -1;  // uniform
        if (tid == 0u)
            L.n_ev = 0u;
        if (int(L.F[0] / a.R12) < xlo || int(L.F[nF - 1u] / a.R12) > xhi)
            return 0;  // uniform: every thread reads the same F
        for (uint32_t i = tid; i < nF; i += nth) {
            const uint32_t v = L.F[i];
            const Slots<S> sv = load_slots<S>(a.src, v);
            for (uint32_t q = 0; q < NW; ++q)
                L.FL[i * NW + q] = sv.w[q];
            const int ei = skew_find(k, base + v);
            const SkewEntry e = ei >= 0 ? k.ent[ei] : SkewEntry{0u, 0, 0u, 0u, {0, 0, 0, 0}};
            L.FO[i] = L.FW[i] = count_nonzero(sv) + uint32_t(e.off);
            L.FS[i] = e.sp;
            for (int q = 0; q < kSkewSp; ++q)
                L.FD[i * kSkewSp + q] = e.dc[q];
        }
        __syncthreads();
        const uint32_t nW =
            block_hood_set(a, L.F, nF, 7u, nullptr, 0u, L.W, kBlockReplayW, cand, 2u * kBlockReplayEv, L.box, L.scan);
        // the moves of the sources W = F and its six neighbours that have an end in F
        for (uint32_t wi = tid; wi < nW; wi += nth) {
            const uint32_t w = L.W[wi];
            const int wF = find_sorted(L.F, nF, w);
            const Slots<S> sw = load_slots<S>(a.src, w);
            // the walk: occupancy word at the step's start, with the reference's u32 end index
            const uint32_t occw = wF >= 0 ? L.FW[wF] : count_nonzero(sw);
            const uint32_t start = w * uint32_t(S), end = start + occw;
            uint32_t walk = end > start ? end - start : 0u;
            if (walk == 0u)
                continue;
            if (walk > 2u * uint32_t(S)) {  // a word wrapped below zero at the first voxels: unrepresentable
                bad = true;
                walk = 2u * uint32_t(S);
            }
            const uint32_t z = w % a.R2, y = (w / a.R2) % a.R1, x = w / a.R12;
            const Hood h = make_hood(a, x, y, z, w, S);
            const Slots<S> sn = (walk > uint32_t(S) && w + 1u < a.N) ? load_slots<S>(a.src, w + 1u) : zero_slots<S>();
            for (uint32_t s = 0; s < walk; ++s) {
                const uint32_t blk = w + s / uint32_t(S), j = s % uint32_t(S);
                const uint32_t sp = blk >= a.N ? 0u : (blk == w ? get_byte_dyn(sw, j) : get_byte_dyn(sn, j));
                if (sp == 0u)
                    continue;
                const uint32_t d = hop(a, tb, h, sp, s);
                if (d >= 6u)
                    continue;
                const int tF = find_sorted(L.F, nF, neighbor_or_none(a, w, int(d)));
                if (tF < 0 && wF < 0)
                    continue;  // neither end in F: the gather has it
                const uint32_t idx = atomicAdd(&L.n_ev, 1u);
                if (idx < kBlockReplayEv)  // species | F index of the target << 8 | of the source << 20
                    L.ev[idx] = (uint64_t((wi << 8) | (s << 3) | d) << 32) |
                                (sp | (tF >= 0 ? uint32_t(tF) : kBEvNoF) << 8 |
                                 (wF >= 0 ? uint32_t(wF) : kBEvNoF) << 20);
            }
        }
        __syncthreads();
        const uint32_t n_ev = L.n_ev;
        if (n_ev > kBlockReplayEv)
            return -1;  // uniform
        if (n_ev > 1u)
            block_bitonic<uint64_t>(L.ev, n_ev, pow2_at_least(n_ev));
        __syncthreads();
        // the claims and departures in the canonical order (the scatter form of rdme.wgsl:45-88), on one thread
        if (tid == 0u) {
            failed = alias = 0;
            uint32_t esc = 0xFFFFFFFFu;
            uint8_t *FLb = reinterpret_cast<uint8_t *>(L.FL);
            // each event carries both F indices, and the next event is read while this one runs: the state reads
            // of an event are then its only round trips to LDS
            uint64_t e_next = n_ev != 0u ? L.ev[0] : 0ull;
            for (uint32_t i = 0; i < n_ev; ++i) {
                const uint64_t e = e_next;
                if (i + 1u < n_ev)
                    e_next = L.ev[i + 1u];
                const uint32_t key = uint32_t(e >> 32), lo = uint32_t(e);
                const uint32_t wi = key >> 8, s = (key >> 3) & 31u, d = key & 7u;
                const uint32_t sp = lo & 0xFFu, tF0 = (lo >> 8) & 0xFFFu, wF0 = lo >> 20;
                const uint32_t tF = tF0 == kBEvNoF ? kBNoF : tF0, wF = wF0 == kBEvNoF ? kBNoF : wF0;
                if (tF != kBNoF) {
                    const uint32_t old = L.FO[tF];
                    L.FO[tF] = old + 1u;
                    bool ok = false;
                    if (old <= uint32_t(S)) {  // rdme.wgsl:62: admits a claim into a voxel holding S
                        for (uint32_t q = 0; q < NW && !ok; ++q) {
                            const uint32_t zb = ~nonzero_bytes(L.FL[tF * NW + q]) & 0x80808080u;
                            if (zb != 0u) {  // the first empty slot
                                FLb[tF * uint32_t(S) + 4u * q + (uint32_t(__ffs(zb) - 1) >> 3)] = uint8_t(sp);
                                ok = true;
                            }
                        }
                        const uint32_t t = L.F[tF];
                        if (!ok && t + 1u < a.N) {  // slot S: slot 0 of the next voxel
                            const int f1 = find_sorted(L.F, nF, t + 1u);
                            if (f1 < 0) {
                                esc = t + 1u;
                                break;
                            }
                            const uint32_t t1F = uint32_t(f1);
                            if (FLb[t1F * uint32_t(S)] == 0u) {
                                FLb[t1F * uint32_t(S)] = uint8_t(sp);
                                ok = true;
                                ++alias;
                                // C[t][sp] counts it, the slots of t+1 hold it
                                bad |= !dc_add(L.FS[tF], L.FD + tF * kSkewSp, sp, 1);
                                bad |= !dc_add(L.FS[t1F], L.FD + t1F * kSkewSp, sp, -1);
                            }
                        }
                    }
                    if (!ok) {
                        L.FO[tF] = old;
                        ++failed;
                        if (wF == kBNoF) {  // a source outside F keeps a particle the gather let go
                            esc = L.W[wi];
                            break;
                        }
                        continue;
                    }
                }
                if (wF == kBNoF)
                    continue;
                // the source: O[w] -= 1, lattice[i_part] = 0, C[w][sp] -= 1 (rdme.wgsl:69-71)
                L.FO[wF] -= 1u;
                uint32_t bF = wF;
                if (s >= uint32_t(S)) {  // a walk past slot S: the particle sits in the next voxel's block
                    const int f = find_sorted(L.F, nF, L.W[wi] + 1u);
                    if (f < 0) {  // a block outside F
                        esc = L.W[wi] + 1u;
                        break;
                    }
                    bF = uint32_t(f);
                }
                const uint32_t at = bF * uint32_t(S) + s % uint32_t(S);
                const uint32_t cur = FLb[at];
                FLb[at] = 0u;
                if (!(bF == wF && cur == sp)) {
                    bad |= !dc_add(L.FS[bF], L.FD + bF * kSkewSp, cur, 1);
                    bad |= !dc_add(L.FS[wF], L.FD + wF * kSkewSp, sp, -1);
                }
            }
            if (esc != 0xFFFFFFFFu && L.n_ext < uint32_t(kEscMax)) {
                uint32_t q = L.n_ext++;  // keep ext sorted
                while (q > 0u && L.ext[q - 1u] > esc) {
                    L.ext[q] = L.ext[q - 1u];
                    --q;
                }
                L.ext[q] = esc;
            } else if (esc != 0xFFFFFFFFu) {
                L.n_ext = uint32_t(kEscMax) + 1u;  // more voxels than F may add
            }
            L.esc = esc;
        }
        __syncthreads();
        const bool any_bad = __syncthreads_or(bad ? 1 : 0) != 0;
        if (L.esc == 0xFFFFFFFFu)
            done = true;
        else if (L.n_ext > uint32_t(kEscMax) || any_bad)
            return 0;
    }
    if (!done)
        return 0;
    // the counts the step kernel took for F's gathered outputs; reaction and compaction of F; outputs, offsets and
    // live bits (one voxel per thread)
    for (uint32_t i = tid; i < nF; i += nth) {
        const uint32_t v = L.F[i];
        VoxelFlags g;
        (void)gather_voxel<S>(a, tb, v, g);
        tally.fired -= g.fired;
        tally.rolled -= g.rolled;
        tally.near -= g.near;
        int ei = skew_find(k, base + v);
        const uint32_t kf = ei >= 0 ? k.ent[ei].kf : 0u;
        XVox<S> vx;
        for (uint32_t q = 0; q < NW; ++q)
            vx.sl.w[q] = L.FL[i * NW + q];
        vx.occ = L.FO[i];
        vx.sp = L.FS[i];
        for (int q = 0; q < kSkewSp; ++q)
            vx.dc[q] = L.FD[i * kSkewSp + q];
        bool fired, rolled, near, xbad = false;
        exact_react<S>(a, vx, kf & 0xFFu, v / a.R12, (v / a.R2) % a.R1, v % a.R2, fired, rolled, near, xbad);
        tally.fired += fired;
        tally.rolled += rolled;
        tally.near += near;
        bad |= xbad;
        store_slots<S>(a.dst, v, vx.sl);
        const int32_t off = int32_t(vx.occ - count_nonzero(vx.sl));
        if (count_nonzero(vx.sl) != 0u)
            sum_mark(a.sum_dst, v);
        if (ei < 0 && (off != 0 || vx.sp != 0u))
            ei = skew_insert(k, base + v);
        if (ei >= 0) {
            SkewEntry &e = k.ent[ei];
            e.off = off;
            e.sp = vx.sp;
            for (int q = 0; q < kSkewSp; ++q)
                e.dc[q] = vx.dc[q];
            e.kf = kf | kSkewDone;
            set_voxel_bit(k.live, base + v, skew_live(e));
        } else if (off != 0 || vx.sp != 0u) {
            bad = true;  // table full
        }
    }
    const bool any_bad = __syncthreads_or(bad ? 1 : 0) != 0;
    if (tid == 0u) {
        tally.failed += failed;
        tally.alias += alias;
        tally.bad += any_bad;
    }
    return 1;
}

// The reaction of a live entry that no cluster replayed this step: the step kernel stored the voxel's
// post-diffusion slots and left its reaction here.
template <int S>
__device__ void skew_react(const StepArgs &a, const SkewTab &k, uint32_t ei, ReplayTally &tally)
{
    SkewEntry e = k.ent[ei];
    if (e.kf & kSkewDone) {
        k.ent[ei].kf = e.kf & ~kSkewDone;
        return;
    }
    if (!skew_live(e))
        return;
    const uint32_t v = e.v - a.vox_base;
    XVox<S> vx;
    vx.sl = load_slots<S>(a.dst, v);
    vx.occ = count_nonzero(vx.sl) + uint32_t(e.off);
    vx.sp = e.sp;
    for (int q = 0; q < kSkewSp; ++q)
        vx.dc[q] = e.dc[q];
    bool fired, rolled, near, bad = false;
    exact_react<S>(a, vx, e.kf & 0xFFu, v / a.R12, (v / a.R2) % a.R1, v % a.R2, fired, rolled, near, bad);
    tally.fired += fired;
    tally.rolled += rolled;
    tally.near += near;
    tally.bad += bad;
    if (fired && !rolled) {
        store_slots<S>(a.dst, v, vx.sl);
        if (count_nonzero(vx.sl) != 0u)
            sum_mark(a.sum_dst, v);
    }
    e.off = int32_t(vx.occ - count_nonzero(vx.sl));
    e.sp = vx.sp;
    for (int q = 0; q < kSkewSp; ++q)
        e.dc[q] = vx.dc[q];
    k.ent[ei] = e;
    if (!skew_live(e))
        set_voxel_bit(k.live, e.v, false);
}


// A skewed voxel whose replay can be left out: no particle, a walk that stays in its own (empty) slots, and no
// particle of a neighbour hopping into it (so no claim meets its occupancy word). Its gathered output (empty) is
// exact, its neighbours receive nothing from it, and its reaction runs with its offsets in skew_react.
template <int S>
__device__ bool skew_inert(const StepArgs &a, const SkewEntry &e)
{
    const uint32_t v = e.v - a.vox_base;
    if (count_nonzero(load_slots<S>(a.src, v)) != 0u)
        return false;
    const uint32_t start = v * uint32_t(S), end = start + uint32_t(e.off);
    if ((end > start ? end - start : 0u) > uint32_t(S))
        return false;
    const Tab tb = global_tab(a);
    for (int k = 0; k < 6; ++k) {
        const uint32_t u = neighbor_or_none(a, v, k);
        if (u == 0xFFFFFFFFu)
            continue;
        const Slots<S> su = load_slots<S>(a.src, u);
        const uint32_t occ = count_nonzero(su);
        if (occ == 0u)
            continue;
        const uint32_t z = u % a.R2, y = (u / a.R2) % a.R1, x = u / a.R12;
        const Hood h = make_hood(a, x, y, z, u, S);
        const uint32_t back = uint32_t(k ^ 1);  // the direction from u to v
        for (uint32_t j = 0; j < occ; ++j)
            if (hop(a, tb, h, get_byte_dyn(su, j), j) == back)
                return false;
    }
    return true;
}

// Ascending, duplicate-free list of the noted voxels from the bitmaps: T[0..n) marked, then read back in order.
// Returns the list's length.
__device__ uint32_t bitmap_sort(const ReplayBuf &b, uint32_t n)
{
    __shared__ uint32_t s_part[kReplayThreads + 1];
    const uint32_t tid = threadIdx.x, nth = blockDim.x;
    const uint32_t n_hi = (b.nvox + 1023u) / 1024u, hw = (n_hi + 31u) / 32u;
    for (uint32_t i = tid; i < n; i += nth) {
        const uint32_t v = b.T[i];
        atomicOr(b.mark_lo + (v >> 5), 1u << (v & 31u));
        atomicOr(b.mark_hi + (v >> 15), 1u << ((v >> 10) & 31u));
    }
    __syncthreads();
    // voxels under each word of mark_hi, then its first position: contiguous words per thread, block prefix
    const uint32_t per = (hw + nth - 1u) / nth, w0 = min(tid * per, hw), w1 = min(w0 + per, hw);
    uint32_t mine = 0;
    for (uint32_t w = w0; w < w1; ++w) {
        uint32_t c = 0, hb = b.mark_hi[w];
        while (hb) {
            const uint32_t bit = uint32_t(__ffs(hb) - 1);
            hb &= hb - 1u;
            const uint32_t lo0 = (w * 32u + bit) * 32u;
            for (uint32_t q = 0; q < 32u; ++q)
                c += uint32_t(__popc(b.mark_lo[lo0 + q]));
        }
        mine += c;
    }
    s_part[tid] = mine;
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        for (uint32_t t = 0; t < nth; ++t) {
            const uint32_t c = s_part[t];
            s_part[t] = acc;
            acc += c;
        }
        s_part[nth] = acc;
    }
    __syncthreads();
    uint32_t pos = s_part[tid];
    for (uint32_t w = w0; w < w1; ++w) {
        uint32_t hb = b.mark_hi[w];
        while (hb) {
            const uint32_t bit = uint32_t(__ffs(hb) - 1);
            hb &= hb - 1u;
            const uint32_t lo0 = (w * 32u + bit) * 32u;
            for (uint32_t q = 0; q < 32u; ++q) {
                uint32_t lb = b.mark_lo[lo0 + q];
                if (!lb)
                    continue;
                b.mark_lo[lo0 + q] = 0u;
                while (lb) {
                    const uint32_t t = uint32_t(__ffs(lb) - 1);
                    lb &= lb - 1u;
                    b.T[pos++] = (lo0 + q) * 32u + t;
                }
            }
        }
        b.mark_hi[w] = 0u;
    }
    __syncthreads();
    return s_part[nth];
}

// Splits the run of noted voxels T[r.x .. r.x + r.y) of one cell (ascending, all within kClusterGap planes of the
// next) where no two voxels kClusterGap apart along y, then along z, belong to one piece; appends each piece as a
// cluster, its voxels ascending. Runs longer than 64 stay whole (a crowded run is one cluster anyway).
__device__ void split_run(const StepArgs &a, const ReplayBuf &b, uint4 r, uint32_t *n_clus)
{
    uint32_t *T = b.T + r.x;
    const uint32_t m = r.y;
    auto emit = [&](uint32_t first, uint32_t count, uint32_t xlo) {
        const uint32_t c = atomicAdd(n_clus, 1u);
        if (c < b.ccap)
            b.clus[c] = make_uint4(r.x + first, count, r.z, xlo);
    };
    if (m > 64u) {
        emit(0u, m, r.w);
        return;
    }
    uint32_t key[64], idx[64];
    // order by (y, v)
    for (uint32_t i = 0; i < m; ++i) {
        const uint32_t y = (T[i] / a.R2) % a.R1;
        uint32_t q = i;
        while (q > 0u && key[q - 1u] > y) {
            key[q] = key[q - 1u];
            idx[q] = idx[q - 1u];
            --q;
        }
        key[q] = y;
        idx[q] = T[i];
    }
    uint32_t out[64];
    uint32_t n_out = 0;
    uint32_t g0 = 0;
    while (g0 < m) {
        uint32_t g1 = g0 + 1u;
        while (g1 < m && key[g1] - key[g1 - 1u] < kClusterGap)
            ++g1;
        // the y-group idx[g0..g1): order by (z, v), split along z
        uint32_t zk[64], zv[64];
        const uint32_t gm = g1 - g0;
        for (uint32_t i = 0; i < gm; ++i) {
            const uint32_t v = idx[g0 + i], z = v % a.R2;
            uint32_t q = i;
            while (q > 0u && (zk[q - 1u] > z || (zk[q - 1u] == z && zv[q - 1u] > v))) {
                zk[q] = zk[q - 1u];
                zv[q] = zv[q - 1u];
                --q;
            }
            zk[q] = z;
            zv[q] = v;
        }
        uint32_t h0 = 0;
        while (h0 < gm) {
            uint32_t h1 = h0 + 1u;
            while (h1 < gm && zk[h1] - zk[h1 - 1u] < kClusterGap)
                ++h1;
            // a piece: its voxels ascending
            const uint32_t first = n_out;
            uint32_t xlo = 0xFFFFFFFFu;
            for (uint32_t i = h0; i < h1; ++i) {
                const uint32_t v = zv[i];
                uint32_t q = n_out++;
                while (q > first && out[q - 1u] > v) {
                    out[q] = out[q - 1u];
                    --q;
                }
                out[q] = v;
                xlo = min(xlo, v / a.R12);
            }
            emit(first, h1 - h0, xlo);
            h0 = h1;
        }
        g0 = g1;
    }
    for (uint32_t i = 0; i < m; ++i)
        T[i] = out[i];
}

// The replay kernel's work for one step (see the section comment). `cells` maps a voxel index to the cell that
// holds it: find(v) -> cell, args(cell) -> its StepArgs (vox_base, extents, tables, dst).
template <int S, class Cells>
__device__ void replay_step(const Cells &cells, const SkewTab &k, const ReplayBuf &b, unsigned long long *counters)
{
    __shared__ uint32_t s_n, s_nx, s_ncl;
    // the sort's list, and afterwards one replay area per wave
    constexpr size_t kWaves = kReplayThreads / 64u;
    constexpr size_t kRaw = std::max(std::max(size_t(kSortLds) * sizeof(uint32_t), kWaves * sizeof(WaveReplayLds<S>)),
                                     sizeof(BlockReplayLds<S>));
    __shared__ __align__(16) uint8_t s_raw[kRaw];
    uint32_t *s_sort = reinterpret_cast<uint32_t *>(s_raw);
    __shared__ unsigned long long s_tally[6];  // fired, rolled back, near, exceptions, failed claims, aliases
    const uint32_t tid = threadIdx.x, nth = blockDim.x;
    const uint32_t ne = k.ctl[0];
    if (tid == 0) {
        s_n = b.ctl[0];
        s_ncl = 0u;
    }
    if (tid < 6u)
        s_tally[tid] = 0ull;
    __syncthreads();
    ReplayTally tally;
    // every voxel whose occupancy word is off its slot count (its walk and the claims into it differ from the
    // gather), unless nothing reaches it this step
    for (uint32_t i = tid; i < ne; i += nth) {
        const SkewEntry e = k.ent[i];
        if (e.off != 0 && !skew_inert<S>(cells.args(cells.find(e.v)), e)) {
            const uint32_t j = atomicAdd(&s_n, 1u);
            if (j < b.tcap)
                b.T[j] = e.v;
        }
    }
    __syncthreads();
    uint32_t n = s_n;
    if (n > b.tcap) {
        if (tid == 0)
            tally.bad += n;  // more voxels than the list holds: none replayed
    } else if (n > 0u) {
        uint32_t n2 = 1u;
        while (n2 < n)
            n2 <<= 1;
        uint32_t *arr = b.T;
        if (n2 <= b.lds_sort_max) {
            for (uint32_t i = tid; i < n; i += nth)
                s_sort[i] = b.T[i];
            arr = s_sort;
            block_bitonic(arr, n, n2);
        } else {
            n = bitmap_sort(b, n);  // ascending and duplicate-free
        }
        if (tid == 0) {
            // runs of one cell whose consecutive noted voxels are fewer than kClusterGap planes apart
            uint32_t m = 0, nx = 0, prev = 0xFFFFFFFFu, cell = 0, cbase = 0, cend = 0, last_x = 0;
            uint32_t R12 = 1;
            for (uint32_t i = 0; i < n; ++i) {
                const uint32_t v = arr[i];
                if (v == prev)
                    continue;
                prev = v;
                bool fresh = nx == 0u;
                if (fresh || v >= cend) {
                    cell = cells.find(v);
                    const StepArgs ca = cells.args(cell);
                    cbase = ca.vox_base;
                    cend = cbase + ca.N;
                    R12 = ca.R12;
                    fresh = true;
                }
                const uint32_t local = v - cbase, x = local / R12;
                if (!fresh && x - last_x >= kClusterGap)
                    fresh = true;
                if (fresh) {
                    if (nx == b.ccap)
                        break;  // more runs than clusters fit: cannot happen (runs are kClusterGap planes apart)
                    b.xclus[nx++] = make_uint4(m, 0u, cell, x);
                }
                b.T[m++] = local;
                b.xclus[nx - 1u].y += 1u;
                last_x = x;
            }
            s_nx = nx;
            b.ctl[2] = 0u;
        }
        __syncthreads();
        for (uint32_t c = tid; c < s_nx; c += nth) {
            const uint4 r = b.xclus[c];
            split_run(cells.args(r.z), b, r, &s_ncl);
        }
        __syncthreads();
        const uint32_t ncl = min(s_ncl, b.ccap);
        if (tid == 0 && s_ncl > b.ccap)
            tally.bad += s_ncl - b.ccap;
        // clusters of more than block_max noted voxels: one lane each, the arrays in global memory
        for (uint32_t c = tid; c < ncl; c += nth) {
            const uint4 cl = b.clus[c];
            if (cl.y > b.block_max && !replay_cluster<S>(cells.args(cl.z), k, b, b.T + cl.x, cl.y, cl.w, tally))
                tally.bad += cl.y;
        }
        // more than wave_max: the whole block, one cluster after another, the arrays in its LDS
        BlockReplayLds<S> &bl = *reinterpret_cast<BlockReplayLds<S> *>(s_raw);
        for (uint32_t c = 0; c < ncl; ++c) {
            const uint4 cl = b.clus[c];
            if (cl.y <= b.wave_max || cl.y > b.block_max)
                continue;
            const StepArgs ca = cells.args(cl.z);
            const int r = replay_cluster_block<S>(ca, k, bl, b.T + cl.x, cl.y, cl.w, tally);
            if (r < 0 && tid == 0u) {
                if (!replay_cluster<S>(ca, k, b, b.T + cl.x, cl.y, cl.w, tally))
                    tally.bad += cl.y;
            } else if (r == 0 && tid == 0u) {
                tally.bad += cl.y;
            }
            __syncthreads();
        }
        // the others: one wave each, the arrays in its LDS
        WaveReplayLds<S> &wl = reinterpret_cast<WaveReplayLds<S> *>(s_raw)[tid / 64u];
        for (uint32_t c = tid / 64u; c < ncl; c += nth / 64u) {
            const uint4 cl = b.clus[c];
            if (cl.y > b.wave_max)
                continue;
            const StepArgs ca = cells.args(cl.z);
            const int r = replay_cluster_wave<S>(ca, k, wl, b.T + cl.x, cl.y, cl.w, tally);
            if (r < 0 && __lane_id() == 0u) {
                if (!replay_cluster<S>(ca, k, b, b.T + cl.x, cl.y, cl.w, tally))
                    tally.bad += cl.y;
            } else if (r == 0 && __lane_id() == 0u) {
                tally.bad += cl.y;
            }
        }
        __syncthreads();
    }
    // the reactions of the live entries no cluster replayed (and the replayed ones re-armed)
    const uint32_t ne2 = k.ctl[0];
    for (uint32_t i = tid; i < ne2; i += nth)
        skew_react<S>(cells.args(cells.find(k.ent[i].v)), k, i, tally);
    const long long mine[6] = {tally.fired, tally.rolled, tally.near, tally.bad, tally.failed, tally.alias};
    for (int q = 0; q < 6; ++q)
        if (mine[q])
            atomicAdd(&s_tally[q], (unsigned long long)mine[q]);
    __syncthreads();
    if (tid == 0) {
        const int idx[6] = {CS_CNT_FIRED, CS_CNT_ROLLBACK, CS_CNT_NEAR_THRESHOLD, CS_CNT_EXCEPTION, CS_CNT_FAILED_CLAIM,
                            CS_CNT_ALIAS};
        for (int q = 0; q < 6; ++q)
            counter_add(counters, idx[q], (long long)s_tally[q]);
        b.ctl[0] = 0u;
    }
}

__device__ __forceinline__ void note_tight(const StepArgs &a, uint32_t v)
{
    const uint32_t i = atomicAdd(a.fix_ctl, 1u);
    if (i < a.max_tight)
        a.fix_list[i] = a.vox_base + v;
}

struct PlainCells {
    StepArgs a;
    __device__ uint32_t find(uint32_t) const { return 0u; }
    __device__ StepArgs args(uint32_t) const { return a; }
};

// One block launched after every step kernel on the same stream: replays what the step noted (if anything) and
// the skewed voxels, runs the reactions of the live entries, and re-arms the wave kernel's work queues. With
// nothing noted and no entries it returns at once, the normal case.
template <int S>
__global__ __launch_bounds__(kReplayThreads) void tight_replay_kernel(StepArgs a, SkewTab k, ReplayBuf b)
{
    if (threadIdx.x < 8u)
        b.ctl[kFixQueue + threadIdx.x] = 0u;  // the wave kernel's work queues, for the next step
    if (b.ctl[0] == 0u && k.ctl[0] == 0u)
        return;
    replay_step<S>(PlainCells{a}, k, b, a.counters);
}

// ---- view (final_stage.wgsl:35-67): one value per voxel for a renderer -----------------------------
// Voxel order (v = z + y*R2 + x*R2*R1). Hidden (0) when res[axis] - coord[axis] < slice. View 0: the species
// found most often among slots 0..occupancy-1 in lattice-index order (slot S and on are the next voxel's,
// past the lattice 0; ties keep the first in slot order), over 255; view 1: region label / 255; view 2:
// reservoir species / 255. The occupancy word is the slot count plus the skewed voxel's offset; a word above 2S
// (only one wrapped below zero, whose loop in the reference walks the whole buffer) gives 0.
template <int S>
__global__ __launch_bounds__(kBlock) void view_kernel(StepArgs a, SkewTab k, const uint8_t *resv, uint32_t view,
                                                      uint32_t axis, uint32_t slice, float *out)
{
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= a.N)
        return;
    const uint32_t z = v % a.R2, y = (v / a.R2) % a.R1, x = v / a.R12;
    const uint32_t c = axis == 0u ? x : (axis == 1u ? y : z);
    const uint32_t R = axis == 0u ? a.R0 : (axis == 1u ? a.R1 : a.R2);
    float val = 0.0f;
    if (!(R - c < slice)) {
        if (view == 1u) {
            val = float(a.reg[v]) / 255.0f;
        } else if (view == 2u) {
            val = float(resv != nullptr ? resv[v] : 0u) / 255.0f;
        } else {
            const Slots<S> s = load_slots<S>(a.src, v);
            int32_t off = 0;
            if (voxel_bit(a.live, a.vox_base + v)) {
                const int ei = skew_find(k, a.vox_base + v);
                off = ei >= 0 ? k.ent[ei].off : 0;
            }
            const uint32_t occ = count_nonzero(s) + uint32_t(off);
            uint32_t best = 0u, maxc = 0u;
            if (occ <= 2u * uint32_t(S)) {
                const Slots<S> nx = (occ > uint32_t(S) && v + 1u < a.N) ? load_slots<S>(a.src, v + 1u) : zero_slots<S>();
                for (uint32_t i = 0; i < occ; ++i) {
                    const uint32_t si = i < uint32_t(S) ? get_byte_dyn(s, i) : get_byte_dyn(nx, i - uint32_t(S));
                    uint32_t cnt = 0u;
                    for (uint32_t j = 0; j < occ; ++j)
                        cnt += (j < uint32_t(S) ? get_byte_dyn(s, j) : get_byte_dyn(nx, j - uint32_t(S))) == si;
                    if (cnt > maxc) {
                        maxc = cnt;
                        best = si;
                    }
                }
            }
            val = float(best) / 255.0f;
        }
    }
    out[v] = val;
}

template <int S>
__global__ __launch_bounds__(kBlock) void rdme_step_kernel(StepArgs a)
{
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t v = blk * kBlock + threadIdx.x;
    VoxelFlags f;
    if (v < a.N) {
        const Slots<S> out = gather_voxel<S>(a, global_tab(a), v, f);
        store_slots<S>(a.dst, v, out);
        if (f.tight || f.bad)  // a tight voxel, or a reaction that took a count below zero: replayed
            note_tight(a, v);
    }
    count_event(a.counters, CS_CNT_TIGHT, f.tight);
    count_event(a.counters, CS_CNT_NEAR_THRESHOLD, f.near);
    count_event(a.counters, CS_CNT_FIRED, f.fired);
    count_event(a.counters, CS_CNT_ROLLBACK, f.rolled);
}

// ---- tiled plane-march kernel ------------------------------------------------------------------------
// A block owns a kTY x kTZ tile of (y, z) and marches along x through `xc` planes. Each plane is staged in
// LDS with a one-voxel halo in y and z ("records": slots and per-slot hop codes), so every particle's hop
// is computed once per block instead of once per neighbour, and the gather reads neighbour records from
// LDS. Region labels are staged with a two-voxel halo (a halo record's hop needs its own neighbours'
// labels). The slots and labels of the next plane are loaded into registers one iteration ahead, so their
// HBM latency overlaps the current plane's work. The model tables live in LDS for the whole block.
//
// Particles are sparse, so the per-particle work runs from lists: iteration p stages plane p and lists
// its records that hold particles; the hops of those records mark, on every voxel a particle enters, the
// direction it came by ("arrive" bytes) and a "needs gather" bit. The gather runs two planes behind
// (plane p-2), when every mark it depends on is final, so the hops of plane p and the gather of plane p-2
// run in the same phase on different waves: hop items are taken from the first lanes up, gather items
// from the last lanes down. Lists are per-wave segments filled with ballots (no LDS atomics); two
// barriers per plane.
constexpr int kTZ = 64;                    // one wave per z-row
constexpr int kTY = 4;                     // four rows: 256 lanes
constexpr int kWaves = kBlock / 64;        // == kTY: wave w owns interior row w
constexpr int kRZ = kTZ + 2;               // record row with z halo
constexpr int kRec = (kTY + 2) * kRZ;      // records per plane (396)
constexpr int kHalo = kRec - kTY * kTZ;    // halo records per plane (140)
constexpr int kGZ = kTZ + 4;               // label row with a two-voxel z halo (68 bytes, 17 words)
constexpr int kGrid = (kTY + 4) * kGZ;     // labels per plane (544)
constexpr int kLabWords = kGrid / 4;       // 136: one word of four labels per lane
constexpr int kRing = 4;                   // planes of records: the gather of p-2 reads p-3..p-1 while p is staged
constexpr uint32_t kTabLdsMax = 24 * 1024;  // tables larger than this run on the simple kernel
constexpr int kTightLds = 64;               // tight voxels a block can note in LDS per step
static_assert(kWaves == kTY && kGZ % 4 == 0 && kLabWords <= kBlock, "tile geometry");

struct TileArgs {
    uint32_t tiles_z, tiles_y, chunks, xc;
    uint32_t has_zero_order;               // a reaction without reactants exists (fires in empty voxels)
    uint32_t tab_words;                    // size of the table blob in 32-bit words
    uint32_t off_p2, off_rates, off_ridx, off_stoich, off_rmask, off_bound;  // byte offsets in the blob
    const uint32_t *tab;                   // the blob in global memory
};

// A voxel's slots as one vector, so LDS moves them with one 4/8/16-byte access per lane.
template <int S> struct SlotVec;
template <> struct SlotVec<4> { using T = uint32_t; };
template <> struct SlotVec<8> { using T = uint2; };
template <> struct SlotVec<16> { using T = uint4; };

template <int S>
__device__ __forceinline__ typename SlotVec<S>::T to_vec(const Slots<S> &s)
{
    if constexpr (S == 16)
        return make_uint4(s.w[0], s.w[1], s.w[2], s.w[3]);
    else if constexpr (S == 8)
        return make_uint2(s.w[0], s.w[1]);
    else
        return s.w[0];
}

template <int S>
__device__ __forceinline__ Slots<S> from_vec(const typename SlotVec<S>::T &v)
{
    Slots<S> s;
    if constexpr (S == 16) {
        s.w[0] = v.x; s.w[1] = v.y; s.w[2] = v.z; s.w[3] = v.w;
    } else if constexpr (S == 8) {
        s.w[0] = v.x; s.w[1] = v.y;
    } else {
        s.w[0] = v;
    }
    return s;
}

constexpr int kHopSeg = 192;  // hop items a wave can list per plane

// The part of a record the wave kernel stages: all of it up to 8 slots; the first 8 of 16. A voxel holding more
// (its slot 7 is taken) is rare and reads its whole record from global memory where needed.
template <int S> struct HeadVec { using T = typename SlotVec<S>::T; };
template <> struct HeadVec<16> { using T = uint2; };
template <int S> constexpr int kHeadSlots = S > 8 ? 8 : S;

template <int S>
__device__ __forceinline__ Slots<S> from_head(const typename HeadVec<S>::T &h)
{
    if constexpr (S == 16) {
        Slots<S> s;
        s.w[0] = h.x; s.w[1] = h.y; s.w[2] = 0u; s.w[3] = 0u;
        return s;
    } else {
        return from_vec<S>(h);
    }
}

template <int S>
__device__ __forceinline__ typename HeadVec<S>::T to_head(const Slots<S> &s)
{
    if constexpr (S == 16)
        return make_uint2(s.w[0], s.w[1]);
    else
        return to_vec<S>(s);
}

// The head's last slot is taken: the voxel may hold more particles than the head shows.
template <int S>
__device__ __forceinline__ bool head_full(const typename HeadVec<S>::T &h)
{
    if constexpr (S == 16)
        return (h.y >> 24) != 0u;
    else
        return false;
}

template <int S>
struct TileLds {
    typename SlotVec<S>::T slots[kRing][kRec];
    uint64_t codes[kRing][kRec];            // nibble per slot (see record hops); valid for occupied records
    uint64_t need[kRing][kTY];              // bit z of row y: the voxel holds particles or receives one
    uint32_t arrive[kRing][kTY * kTZ / 4];  // byte per interior voxel: bit d = a particle arrives by a d-hop
    uint32_t reg[3][kLabWords];             // region labels, four per word
    uint32_t hop_list[kWaves][kHopSeg];     // hop items listed by wave w: record | first slot << 9 | slots << 13
    uint16_t slow_list[kWaves][64];         // interior voxels of wave w that need the full gather
    uint32_t hop_cnt[kWaves];
    uint32_t slow_cnt[kWaves];
    uint32_t xmy[kTY + 2];                  // x % y of the staged plane for the record rows y0-1 .. y0+kTY
    uint32_t n_tight;
    uint32_t tight[kTightLds];
};

// Writes `item` at the lane's rank among the lanes where `pred` holds into a per-wave list segment and
// returns the segment's length (the same in every lane).
__device__ __forceinline__ uint32_t wave_list(bool pred, uint32_t item, uint16_t *seg)
{
    const uint64_t m = __ballot(pred);
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
    if (pred)
        seg[rank] = uint16_t(item);
    return uint32_t(__popcll(m));
}

// Item i of the concatenation of n segments with the given lengths: (segment, position).
template <int N>
__device__ __forceinline__ void locate(uint32_t i, const uint32_t (&len)[N], uint32_t &seg, uint32_t &pos)
{
    seg = 0u;
    pos = i;
#pragma unroll
    for (int k = 0; k < N - 1; ++k) {
        const bool past = seg == uint32_t(k) && pos >= len[k];
        pos = past ? pos - len[k] : pos;
        seg = past ? seg + 1u : seg;
    }
}

// record index -> tile coordinates (yy in [-1, kTY], zz in [-1, kTZ])
__device__ __forceinline__ int rec_of(int yy, int zz) { return (yy + 1) * kRZ + (zz + 1); }

// The halo record handled by lane t < kHalo.
__device__ __forceinline__ void halo_coords(int t, int &yy, int &zz)
{
    if (t < kRZ) { yy = -1; zz = t - 1; }
    else if (t < 2 * kRZ) { yy = kTY; zz = t - kRZ - 1; }
    else if (t < 2 * kRZ + kTY) { yy = t - 2 * kRZ; zz = -1; }
    else { yy = t - 2 * kRZ - kTY; zz = kTZ; }
}

template <int S>
__device__ __forceinline__ Slots<S> masked(const Slots<S> &s, bool ok)
{
    Slots<S> r;
#pragma unroll
    for (int q = 0; q < S / 4; ++q)
        r.w[q] = ok ? s.w[q] : 0u;
    return r;
}

// Hops of `ns` slots from `s0` of one record at plane X: the hop code of slot s (0..5 = hops in that
// direction, 6 = stays; an empty slot keeps 0xF) is returned as the bits to clear from an all-0xF nibble
// `clear`, and `summ` gets bit d for every hop in direction d. Labels come from the staged planes X-1, X,
// X+1; `xmy` is X % Y (the reference's seed term, rdme.wgsl:220-226).
template <int S>
__device__ __forceinline__ void item_hops(const StepArgs &a, const Tab &tb, const Slots<S> &sl, int X, int Y, int Z,
                                          int yy, int zz, uint32_t xmy, const uint8_t *rprev, const uint8_t *rcur,
                                          const uint8_t *rnext, uint32_t s0, uint32_t ns, uint64_t &clear,
                                          uint32_t &summ)
{
    clear = 0ull;
    summ = 0u;
    const int gi = (yy + 2) * kGZ + (zz + 2);
    Hood h;
    h.src_reg = rcur[gi];
    h.inside = 0u;
    h.rg[0] = rprev[gi];
    h.rg[1] = rnext[gi];
    h.rg[2] = rcur[gi - kGZ];
    h.rg[3] = rcur[gi + kGZ];
    h.rg[4] = rcur[gi - 1];
    h.rg[5] = rcur[gi + 1];
    if (X > 0) h.inside |= 1u;
    if (uint32_t(X) + 1u < a.R0) h.inside |= 2u;
    if (Y > 0) h.inside |= 4u;
    if (uint32_t(Y) + 1u < a.R1) h.inside |= 8u;
    if (Z > 0) h.inside |= 16u;
    if (uint32_t(Z) + 1u < a.R2) h.inside |= 32u;
    const uint32_t ux = uint32_t(X), uy = uint32_t(Y), uz = uint32_t(Z);
    h.seed0 = a.itime + xmy + uz + (uz + uy * a.R2 + ux * a.R12) * uint32_t(S);
    for (uint32_t s = s0; s < s0 + ns; ++s) {
        const uint32_t d = hop(a, tb, h, get_byte_dyn(sl, s), s);
        clear |= uint64_t(0xFu ^ d) << (4u * s);
        if (d < 6u)
            summ |= 1u << d;
    }
}

// Bit 4k set iff nibble k of `codes` equals `want`.
__device__ __forceinline__ uint64_t nibble_matches(uint64_t codes, uint32_t want)
{
    uint64_t x = codes ^ (uint64_t(want) * 0x1111111111111111ull);
    x |= x >> 1;
    x |= x >> 2;
    return ~x & 0x1111111111111111ull;
}

// Appends the slots of a neighbour record selected by `m` (bit 4s for slot s), in slot order.
template <int S>
__device__ __forceinline__ uint32_t take_arrivals(const typename SlotVec<S>::T &src, uint64_t m, Slots<S> &list,
                                                  uint32_t n)
{
    const Slots<S> sl = from_vec<S>(src);
    while (m) {
        const uint32_t s = uint32_t(__ffsll((long long)m) - 1) >> 2;
        m &= m - 1ull;
        put_byte_dyn(list, n, get_byte_dyn(sl, s));
        ++n;
    }
    return n;
}

// Event counts a block accumulates over the tiles it marches.
struct TileCounts {
    uint32_t tight = 0, near = 0, fired = 0, rolled = 0;
};

// One tile (tz, ty) of the lattice described by (a, t), marched through x-chunk `chunk`. Every lane of the
// block calls it; it starts and ends with the block's LDS free for the next tile.
template <int S>
__device__ __forceinline__ void tile_march(TileLds<S> &L, uint32_t *tab_lds, const StepArgs &a, const TileArgs &t,
                                           uint32_t tz, uint32_t ty, uint32_t chunk, TileCounts &tc)
{
    const int y0 = int(ty) * kTY, z0 = int(tz) * kTZ;
    const int xb = int(chunk * t.xc);
    const int xe = min(int(a.R0), xb + int(t.xc));
    const int lane = int(threadIdx.x);
    const int wv = lane / 64, wl = lane % 64;  // wave and lane in wave
    const int iy = wv, iz = wl;                 // interior voxel of this lane
    int hy = 0, hz = 0;
    const bool has_halo = lane < kHalo;
    if (has_halo)
        halo_coords(lane, hy, hz);
    const int irec = rec_of(iy, iz);
    const int hrec = has_halo ? rec_of(hy, hz) : 0;
    const int Y = y0 + iy, Z = z0 + iz;
    const bool in_lattice = uint32_t(Y) < a.R1 && uint32_t(Z) < a.R2;
    // (y, z) parts of this lane's global offsets. A voxel outside the lattice is loaded from offset 0 of
    // its plane (in bounds) and masked, so no load sits under a branch.
    const bool hal_in = has_halo && y0 + hy >= 0 && uint32_t(y0 + hy) < a.R1 && z0 + hz >= 0 &&
                        uint32_t(z0 + hz) < a.R2;
    const uint32_t mine_yz = in_lattice ? uint32_t(Z) + uint32_t(Y) * a.R2 : 0u;
    const uint32_t hal_yz = hal_in ? uint32_t(z0 + hz) + uint32_t(y0 + hy) * a.R2 : 0u;
    // Labels: lane l < kLabWords loads grid row l / 17, columns 4*(l % 17) .. +3 (z0-2 .. z0+65 over a row).
    const bool lab_lane = lane < kLabWords;
    uint32_t lab_off[4];
    uint32_t lab_ok = 0u;
    {
        const int gy = y0 + lane / (kGZ / 4) - 2, gz0 = z0 + 4 * (lane % (kGZ / 4)) - 2;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int gz = gz0 + b;
            const bool ok = lab_lane && gy >= 0 && uint32_t(gy) < a.R1 && gz >= 0 && uint32_t(gz) < a.R2;
            lab_off[b] = ok ? uint32_t(gz) + uint32_t(gy) * a.R2 : 0u;
            lab_ok |= uint32_t(ok) << b;
        }
    }

    for (uint32_t i = uint32_t(lane); i < t.tab_words; i += kBlock)
        tab_lds[i] = t.tab[i];
    const uint8_t *tbase = reinterpret_cast<const uint8_t *>(tab_lds);
    const Tab tb{reinterpret_cast<const float *>(tbase + t.off_p2), reinterpret_cast<const float *>(tbase + t.off_rates),
                 reinterpret_cast<const uint32_t *>(tbase + t.off_ridx),
                 reinterpret_cast<const int8_t *>(tbase + t.off_stoich),
                 reinterpret_cast<const uint64_t *>(tbase + t.off_rmask),
                 reinterpret_cast<const float *>(tbase + t.off_bound)};
    if (lane == 0)
        L.n_tight = 0u;

    auto plane_base = [&](int X) { return (X >= 0 && X < int(a.R0)) ? uint32_t(X) * a.R12 : 0u; };
    // Label loads are unconditional into full registers; whether a value is real is decided when it is
    // stored a plane later. (A load under a branch, merged into a packed byte, makes the compiler wait for
    // every outstanding load, the slot prefetch included, on the spot.)
    auto fetch_labels = [&](int X, uint32_t (&r)[4]) {
        const uint32_t base = plane_base(X);
#pragma unroll
        for (int b = 0; b < 4; ++b)
            r[b] = a.reg[base + lab_off[b]];
    };
    auto store_labels = [&](int X, const uint32_t (&r)[4]) {
        if (lab_lane) {
            const uint32_t ok = (X >= 0 && X < int(a.R0)) ? lab_ok : 0u;
            uint32_t w = 0u;
#pragma unroll
            for (int b = 0; b < 4; ++b)
                w |= ((ok >> b) & 1u) ? (r[b] & 0xFFu) << (8 * b) : 0u;
            L.reg[(X + 3) % 3][lane] = w;
        }
    };
    auto labels_of = [&](int X) { return reinterpret_cast<const uint8_t *>(L.reg[(X + 3) % 3]); };
    auto ring = [](int p) { return (p + 4) & (kRing - 1); };

    {   // labels of planes xb-2 and xb-1 (needed by the hops of plane xb-1)
        uint32_t r[4];
        fetch_labels(xb - 2, r);
        store_labels(xb - 2, r);
        fetch_labels(xb - 1, r);
        store_labels(xb - 1, r);
    }
    // Registers holding the next plane to stage (its slots, and the labels one plane further), in two
    // sets that the planes use in turn: a set is refilled by the plane that consumes the other one, so
    // no register is copied while its load is in flight (a copy would wait for the load on the spot).
    struct Staged {
        Slots<S> mine, hal;
        uint32_t lab[4];
        bool ok;
    };
    Staged st0, st1;
    st0.ok = xb - 1 >= 0;
    st0.mine = load_slots<S>(a.src, plane_base(xb - 1) + mine_yz);
    st0.hal = load_slots<S>(a.src, plane_base(xb - 1) + hal_yz);
    fetch_labels(xb, st0.lab);

    uint32_t n_tight = 0, n_near = 0, n_fired = 0, n_rolled = 0;
    __syncthreads();  // the tables and the first labels are in LDS

    auto plane = [&](int p, Staged &use, Staged &fill) __attribute__((always_inline)) {
        const int rp = ring(p);
        // ---- T0: stage plane p and list its hop items; list the voxels of plane p-2 to gather
        // prefetch plane p+1 and the labels of plane p+2 (out of the lattice or past the chunk: clamped,
        // never used). Outside any branch, so the loaded registers are never merged at a join point.
        fill.ok = p + 1 < int(a.R0);
        fill.mine = load_slots<S>(a.src, plane_base(p + 1) + mine_yz);
        fill.hal = load_slots<S>(a.src, plane_base(p + 1) + hal_yz);
        fetch_labels(p + 2, fill.lab);
        uint32_t n_items = 0u;
        if (p <= xe) {
            const Slots<S> mine = masked(use.mine, use.ok && in_lattice), hal = masked(use.hal, use.ok && hal_in);
            store_labels(p + 1, use.lab);
            L.slots[rp][irec] = to_vec<S>(mine);
            if (has_halo)
                L.slots[rp][hrec] = to_vec<S>(hal);
            const uint32_t om = count_nonzero(mine), oh = count_nonzero(hal);
            // codes of occupied records start all-0xF; the hop items clear the nibbles of their slots
            if (om)
                L.codes[rp][irec] = ~0ull;
            if (oh)
                L.codes[rp][hrec] = ~0ull;
            if (lane < kTY + 2) {
                const int y = y0 + lane - 1;
                L.xmy[lane] = y > 0 ? uint32_t(p) % uint32_t(y) : (y == 0 ? a.mod_zero : 0u);
            }
            // plane p+1 starts unmarked before the hops of plane p mark it
            if (wl < kTZ / 4)
                L.arrive[ring(p + 1)][wv * (kTZ / 4) + wl] = 0u;
            if (wl == 0)
                L.need[ring(p + 1)][wv] = 0ull;
            const uint64_t occm = __ballot(om != 0u);
            if (wl == 0 && occm != 0ull)
                atomicOr(reinterpret_cast<unsigned long long *>(&L.need[rp][wv]), (unsigned long long)occm);
            // Hop items: one per particle when the wave's particles fit its list, else one per record.
            uint32_t tot = 0u;
#pragma unroll
            for (int b = 0; b < 5; ++b)
                tot += uint32_t(__popcll(__ballot((om >> b) & 1u)) + __popcll(__ballot((oh >> b) & 1u))) << b;
            uint32_t *seg = L.hop_list[wv];
            auto rank = [&](uint64_t m) {
                return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
            };
            if (tot <= uint32_t(kHopSeg)) {
                for (uint32_t j = 0;; ++j) {
                    const bool pm = om > j, ph = oh > j;
                    const uint64_t mm = __ballot(pm), mh = __ballot(ph);
                    if ((mm | mh) == 0ull)
                        break;
                    const uint32_t cm = uint32_t(__popcll(mm));
                    if (pm)
                        seg[n_items + rank(mm)] = uint32_t(irec) | (j << 9) | (1u << 13);
                    if (ph)
                        seg[n_items + cm + rank(mh)] = uint32_t(hrec) | (j << 9) | (1u << 13);
                    n_items += cm + uint32_t(__popcll(mh));
                }
            } else {
                const uint64_t mm = __ballot(om != 0u), mh = __ballot(oh != 0u);
                const uint32_t cm = uint32_t(__popcll(mm));
                if (om)
                    seg[rank(mm)] = uint32_t(irec) | (om << 13);
                if (oh)
                    seg[cm + rank(mh)] = uint32_t(hrec) | (oh << 13);
                n_items = cm + uint32_t(__popcll(mh));
            }
        }
        const int x = p - 2;  // plane gathered in this iteration
        const bool gather_here = x >= xb && x < xe && in_lattice;
        bool slow = false;
        if (x >= xb && x < xe)
            slow = gather_here && (((L.need[ring(x)][wv] >> wl) & 1ull) != 0ull || t.has_zero_order != 0u);
        if (gather_here && !slow) {
            // empty and nothing arrives: the output is empty; only the zero-propensity draw at seed 0 "fires"
            store_slots<S>(a.dst, mine_yz + uint32_t(x) * a.R12, zero_slots<S>());
            n_fired += (a.itime + uint32_t(x) * uint32_t(Y) + uint32_t(Z)) == 0u &&
                       !voxel_bit(a.live, a.vox_base + mine_yz + uint32_t(x) * a.R12);
        }
        const uint32_t n_slow = wave_list(slow, uint32_t(lane), L.slow_list[wv]);
        if (wl == 0) {
            L.hop_cnt[wv] = n_items;
            L.slow_cnt[wv] = n_slow;
        }
        __syncthreads();

        // ---- T1: hops of plane p (first lanes up) and the gather of plane p-2 (last lanes down)
        uint32_t hc[kWaves], sc[kWaves];
        uint32_t nh = 0u, nsl = 0u;
#pragma unroll
        for (int k = 0; k < kWaves; ++k) {
            hc[k] = L.hop_cnt[k];
            nh += hc[k];
            sc[k] = L.slow_cnt[k];
            nsl += sc[k];
        }
        {
            const uint8_t *r_prev = labels_of(p - 1), *r_cur = labels_of(p), *r_next = labels_of(p + 1);
            for (uint32_t i = uint32_t(lane); i < nh; i += kBlock) {
                uint32_t seg, pos;
                locate(i, hc, seg, pos);
                const uint32_t item = L.hop_list[seg][pos];
                const int rec = int(item & 511u);
                const int yy = rec / kRZ - 1, zz = rec % kRZ - 1;
                const Slots<S> sl = from_vec<S>(L.slots[rp][rec]);
                uint64_t clear;
                uint32_t sm;
                item_hops<S>(a, tb, sl, p, y0 + yy, z0 + zz, yy, zz, L.xmy[yy + 1], r_prev, r_cur, r_next,
                             (item >> 9) & 15u, item >> 13, clear, sm);
                if (clear)
                    atomicAnd(reinterpret_cast<unsigned long long *>(&L.codes[rp][rec]), (unsigned long long)~clear);
                while (sm) {  // mark the interior voxels that receive a particle, with the hop direction
                    const uint32_t d = uint32_t(__ffs(sm) - 1);
                    sm &= sm - 1u;
                    const int vy = yy + (d == 2u ? -1 : (d == 3u ? 1 : 0));
                    const int vz = zz + (d == 4u ? -1 : (d == 5u ? 1 : 0));
                    const int rg = d == 0u ? ring(p - 1) : (d == 1u ? ring(p + 1) : rp);
                    if (uint32_t(vy) < uint32_t(kTY) && uint32_t(vz) < uint32_t(kTZ)) {
                        atomicOr(reinterpret_cast<unsigned long long *>(&L.need[rg][vy]),
                                 (unsigned long long)(1ull << vz));
                        atomicOr(&L.arrive[rg][(vy * kTZ + vz) >> 2], (1u << d) << (8 * (vz & 3)));
                    }
                }
            }
        }
        {
            const int r0 = ring(x - 1), r1 = ring(x), r2 = ring(x + 1);
            for (uint32_t i = uint32_t(kBlock - 1 - lane); i < nsl; i += kBlock) {
                uint32_t seg, pos;
                locate(i, sc, seg, pos);
                const int vl = int(L.slow_list[seg][pos]);
                const int vy = vl / kTZ, vz = vl % kTZ;
                const int vr = rec_of(vy, vz);
                const uint32_t gy = uint32_t(y0 + vy), gz = uint32_t(z0 + vz);
                // every read is issued up front; records without a matching arrival are masked out
                const Slots<S> own = from_vec<S>(L.slots[r1][vr]);
                const uint64_t own_c = L.codes[r1][vr];
                const uint32_t arr = (L.arrive[r1][vl >> 2] >> (8 * (vl & 3))) & 0x3Fu;
                const typename SlotVec<S>::T sx0 = L.slots[r0][vr], sy0 = L.slots[r1][vr - kRZ],
                                             sz0 = L.slots[r1][vr - 1], sz1 = L.slots[r1][vr + 1],
                                             sy1 = L.slots[r1][vr + kRZ], sx1 = L.slots[r2][vr];
                const uint64_t cx0 = L.codes[r0][vr], cy0 = L.codes[r1][vr - kRZ], cz0 = L.codes[r1][vr - 1],
                               cz1 = L.codes[r1][vr + 1], cy1 = L.codes[r1][vr + kRZ], cx1 = L.codes[r2][vr];
                const uint32_t occ = count_nonzero(own);
                // nibble >= 6: stays (6) or empty (0xF)
                const uint64_t stay = ((own_c >> 3) | ((own_c >> 2) & (own_c >> 1))) & 0x1111111111111111ull;
                auto sel = [&](uint32_t bit, uint64_t c, uint32_t want) {
                    return (arr & bit) ? nibble_matches(c, want) : 0ull;
                };
                Slots<S> lo = zero_slots<S>(), hi = zero_slots<S>();
                uint32_t nlo = 0, nhi = 0;
                nlo = take_arrivals<S>(sx0, sel(2u, cx0, 1u), lo, nlo);   // from -x, hopped +x
                nlo = take_arrivals<S>(sy0, sel(8u, cy0, 3u), lo, nlo);   // from -y, hopped +y
                nlo = take_arrivals<S>(sz0, sel(32u, cz0, 5u), lo, nlo);  // from -z, hopped +z
                nhi = take_arrivals<S>(sz1, sel(16u, cz1, 4u), hi, nhi);  // from +z, hopped -z
                nhi = take_arrivals<S>(sy1, sel(4u, cy1, 2u), hi, nhi);   // from +y, hopped -y
                nhi = take_arrivals<S>(sx1, sel(1u, cx1, 0u), hi, nhi);   // from +x, hopped -x
                const uint32_t V = gz + gy * a.R2 + uint32_t(x) * a.R12;
                const bool tight = occ + nlo + nhi > uint32_t(S);
                n_tight += tight;
                uint32_t cnt = 0;
                Slots<S> out = assemble<S>(own, occ, occ ? stay : 0ull, lo, nlo, hi, nhi, cnt);
                bool fired, rolled, near, bad;
                react<S>(a, tb, out, cnt, uint32_t(x), gy, gz, fired, rolled, near, bad);
                n_fired += fired;
                n_rolled += rolled;
                n_near += near;
                // a tight voxel, or a reaction that took a count below zero: replayed after the step
                if (tight || bad) {
                    const uint32_t k = atomicAdd(&L.n_tight, 1u);
                    if (k < uint32_t(kTightLds))
                        L.tight[k] = V;
                    else
                        note_tight(a, V);  // block list full: straight to the global list
                }
                store_slots<S>(a.dst, V, out);
            }
        }
        __syncthreads();
    };
    // Planes xb-1 .. xe+1 in pairs; an odd count ends with plane xe+2, which stages and gathers nothing
    // (calling it unconditionally keeps the register sets free of merges at a branch).
    for (int p = xb - 1; p <= xe + 1; p += 2) {
        plane(p, st0, st1);
        plane(p + 1, st1, st0);
    }
    // the block's tight list is final: the last barrier of the loop follows every append
    const uint32_t nt = min(L.n_tight, uint32_t(kTightLds));
    for (uint32_t i = uint32_t(lane); i < nt; i += kBlock)
        note_tight(a, L.tight[i]);
    tc.tight += n_tight;
    tc.near += n_near;
    tc.fired += n_fired;
    tc.rolled += n_rolled;
    __syncthreads();  // every lane is done with the LDS before the next tile reuses it
}

__device__ __forceinline__ void flush_counts(unsigned long long *counters, const TileCounts &tc)
{
    const uint32_t vals[4] = {tc.tight, tc.near, tc.fired, tc.rolled};
    const int idx[4] = {CS_CNT_TIGHT, CS_CNT_NEAR_THRESHOLD, CS_CNT_FIRED, CS_CNT_ROLLBACK};
#pragma unroll
    for (int k = 0; k < 4; ++k)  // events are rare: one atomic per lane that saw any
        if (vals[k])
            atomicAdd(counters + idx[k], (unsigned long long)vals[k]);
}

// The plain step: one block per (tile, x-chunk) of one lattice.
template <int S>
__global__ __launch_bounds__(kBlock) void rdme_step_tiled(StepArgs a, TileArgs t)
{
    __shared__ TileLds<S> L;
    extern __shared__ uint32_t tab_lds[];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t tz = blk % t.tiles_z;
    const uint32_t ty = (blk / t.tiles_z) % t.tiles_y;
    const uint32_t chunk = blk / (t.tiles_z * t.tiles_y);
    TileCounts tc;
    tile_march<S>(L, tab_lds, a, t, tz, ty, chunk, tc);
    flush_counts(a.counters, tc);
}

// ---- wave-march kernel --------------------------------------------------------------------------------
// Each wave owns a tile of kWY rows (y) x kWZ columns (z) and marches along x through `xc` planes on its
// own: no block barrier sits in the march, so the waves of a CU hide each other's latency. The slots of
// the tile plus a one-voxel halo ("records") and the region labels with a two-voxel halo are loaded one
// plane ahead into registers. Per plane X:
//   T0  the occupied records of plane X are listed, with their slots, in the wave's LDS (ballot ranks);
//   T1  one lane per listed record computes its particles' hops; a hop that ends on a voxel of the tile
//       (planes X-1, X, X+1) appends an arrival event (destination, direction, source slot, species) to
//       the destination plane's list, and an interior particle that leaves sets its bit in the voxel's
//       leave mask;
//   T2  plane X-1 is final (every hop into it came from planes X-2 .. X): voxels without particles or
//       arrivals are written as zeros by their own lane; the others are listed and gathered one lane
//       each: own stays from the leave mask, arrivals from the events sorted into the canonical order,
//       then the reaction.
// Particles are sparse, so the lists keep the lanes busy where one lane per voxel would leave most idle.
// A plane whose event list overflows is gathered by the one-lane-per-voxel global form instead.
constexpr int kWY = 4;                            // rows of a wave tile
constexpr int kWZ = 64;                           // columns of a wave tile: one lane each
constexpr int kWRZ = kWZ + 2;                     // record row with its z halo
constexpr int kWLY = kWY + 4;                     // label rows y0-2 .. y0+kWY+1
constexpr int kWLZ = kWZ + 4;                     // label columns z0-2 .. z0+kWZ+1 (68 bytes, 17 words)
constexpr int kWLabWords = kWLY * kWLZ / 4;       // 136
constexpr int kWBatches = kWY + 3;                // records per lane: interior rows, rows -1 and kWY, z halo
static_assert(kWBatches < 8 && kWY < 8, "wave_offset takes per-lane counts below 8");
constexpr int kWHop = kWY * kWZ + 2 * kWZ + 2 * kWY;  // records a plane can list (392)
constexpr int kWEv = 96;                          // arrival events per destination plane
constexpr int kWReact = 80;                       // deferred reactions a wave holds (runs of up to 64 once 16 wait)
constexpr int kWReactRun = 16;
constexpr uint32_t kWEvEnd = 0xFFu;               // end of an event chain
constexpr int kWaveBlock = 256;                   // four waves: four tiles stacked in y share the tables
constexpr int kWavesPerBlock = kWaveBlock / 64;
constexpr int kWLabLane = (kWLabWords + 63) / 64;  // label words a lane loads per plane
static_assert(kWZ == 64 && kWLZ % 4 == 0 && kWY * kWZ <= 512 && kWY % 2 == 0, "wave tile geometry");

struct WaveArgs {
    uint32_t tiles_z, wtiles_y, btiles_y, chunks, xc;  // wave tiles in z and y, block rows of tiles, x-chunks
    uint32_t has_zero_order;
    uint32_t ev_cap;        // events a plane may hold (<= kWEv); a plane with more takes the global form
    uint32_t tab_words;
    uint32_t off_p2, off_rates, off_ridx, off_stoich, off_rmask, off_bound;
    uint32_t pair_n1;       // ns + 1 when the pair table follows the bounds (pair_offset), else 0
    const uint32_t *tab;
    const uint32_t *regp;   // padded labels [R0+4][PY][PZ] (0xFF outside the lattice), read as words
    uint32_t PY, PZ4;       // PY = kWY*wtiles_y + 4 rows; PZ4 = (kWZ*tiles_z + 4) / 4 words per row
    uint32_t nitems;        // work items: kWavesPerBlock tiles stacked in y over an x-range
    uint32_t *queue;        // [8] work counters, one per XCD; zero at the start of every step
    const uint2 *items;     // [nitems] tz | by << 16, x_begin | x_end << 16; entry e in queue e % 8 (by 0xFFFF: empty)
    const uint8_t *sum_src;   // occupancy summary of src (SumRef layout, geometry of StepArgs::sum_dst), or null
};
static_assert(kWY == int(kSumTY) && kWZ == int(kSumTZ), "the occupancy summary counts wave tiles");

template <int S>
struct WaveLds {
    typename HeadVec<S>::T hop_slots[kWHop];  // listed records of the plane being hopped (their heads); then the
                                               // own slots of the voxels of plane X-1 to gather
    uint16_t hop_rec[kWHop];                   // their record ids: (row + 1) * kWRZ + (column + 1); bit 15: full head;
                                               // then the gathered voxels' leave masks
    uint32_t ev[3][kWEv];                      // arrival events per destination plane (X % 3): direction |
                                               // species << 3 | next event to the same voxel << 11
    uint32_t head[3][kWY * kWZ / 2];           // latest event into each interior voxel (kWEvEnd: none), 16 bits each
    uint32_t ev_n[4];                          // events appended (may pass the cap: the plane overflowed)
    uint32_t leave[2][kWY * kWZ / 2];          // 16-bit leave mask per interior voxel (X % 2)
    uint32_t lab[3][kWLabWords];               // label planes (X % 3)
    uint16_t slow[kWY * kWZ];                  // voxels of plane X-1 to gather: row * kWZ + column | chain head << 8
    typename SlotVec<S>::T rb_slots[kWReact];  // gathered voxels whose reaction draw passed the bound test,
    uint32_t rb_v[kWReact];                    // their voxel index,
    uint32_t rb_xc[kWReact];                   // x | particle count << 24
    uint32_t rb_yz[kWReact];                   // and y << 16 | z (a circular list)
};

// Waits for this wave's outstanding vector-memory operations. A rare path that loads from global memory ends
// with it, so that no load result is pending where the paths join: the wait the compiler would put there
// instead runs on every path and, as loads complete in order, waits for the next plane's prefetch too.
__device__ __forceinline__ void vm_drain()
{
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt left alone
}

__device__ __forceinline__ int ring3(int x) { return (x + 3) % 3; }

// Exclusive prefix sum over the wave of per-lane counts c < 8 (three bit planes, one ballot each); `total` gets
// the wave's sum.
__device__ __forceinline__ uint32_t wave_offset(uint32_t c, uint32_t &total)
{
    const uint64_t b0 = __ballot(c & 1u), b1 = __ballot(c & 2u), b2 = __ballot(c & 4u);
    auto below = [](uint64_t m) {
        return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
    };
    total = uint32_t(__popcll(b0)) + 2u * uint32_t(__popcll(b1)) + 4u * uint32_t(__popcll(b2));
    return below(b0) + 2u * below(b1) + 4u * below(b2);
}

template <int S>
__device__ __forceinline__ bool vec_nonzero(const typename SlotVec<S>::T &v)
{
    if constexpr (S == 16)
        return (v.x | v.y | v.z | v.w) != 0u;
    else if constexpr (S == 8)
        return (v.x | v.y) != 0u;
    else
        return v != 0u;
}

template <int S>
__device__ __forceinline__ typename SlotVec<S>::T load_vec(const uint8_t *base, uint32_t v)
{
    return *reinterpret_cast<const typename SlotVec<S>::T *>(base + size_t(v) * uint32_t(S));
}

// The whole record of voxel v for the lanes in `m` (a rare case: more particles than the staged head). Read with
// scalar loads one lane at a time: a vector load here would make the wave wait for every vector load issued
// before it, the next plane's prefetch included (vector loads complete in order). The lattice is not written
// during the step, so the scalar cache may hold it.
template <int S>
__device__ __forceinline__ void full_records(const uint8_t *src, uint32_t v, uint64_t m, Slots<S> &out)
{
    typedef __attribute__((address_space(4))) const uint32_t *cptr;
    while (m != 0ull) {
        const uint32_t ln = uint32_t(__ffsll((long long)m) - 1);
        m &= m - 1ull;
        const uint32_t vv = __builtin_amdgcn_readlane(v, ln);
        const cptr p = (cptr)(src + size_t(vv) * uint32_t(S));
        Slots<S> q;
#pragma unroll
        for (int k = 0; k < S / 4; ++k)
            q.w[k] = p[k];
        if (__lane_id() == ln)
            out = q;
    }
}

// A pointer that is the same on every lane, in scalar registers: a buffer resource must be, and the ensemble
// kernel's cell pointers come from a per-lane load of the cell's descriptor (without this the compiler wraps every
// buffer store in a loop over the distinct values).
__device__ __forceinline__ uint8_t *uniform_ptr(uint8_t *p)
{
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return reinterpret_cast<uint8_t *>((uint64_t(hi) << 32) | lo);
}

// Raw buffer stores (stride 0, no scalar offset): a lane whose byte offset is at or past the resource's size stores
// nothing. Word 3 of the resource is the gfx9 default (32-bit data format, no swizzle); kBufferDrop lies past any
// plane of a lattice the wave kernels take (R1, R2 <= 4096, S <= 16: below 2^28 bytes). The plane's offset goes
// into the base: passed as the scalar offset instead, one store of the gene-expression parity case was lost.
constexpr int kBufferRsrcWord3 = 0x00020000;
constexpr uint32_t kBufferDrop = 0x80000000u;
// The row stores are non-temporal (cache policy bit 1, `nt`): no step reads an output line again before the lattice
// has streamed through L2, and lines kept there would evict the halo lines that neighbouring tiles read shortly after
// their owners (256^3: 3% fewer bytes read, 192.5 -> 189.9 us).
constexpr int kStoreNT = 2;
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

template <int S>
__device__ __forceinline__ void buffer_store_vec(const typename SlotVec<S>::T &v, __amdgpu_buffer_rsrc_t rs,
                                                 uint32_t off)
{
    if constexpr (S == 16) {
        const u32x4_t d = {v.x, v.y, v.z, v.w};
        __builtin_amdgcn_raw_buffer_store_b128(d, rs, int(off), 0, kStoreNT);
    } else if constexpr (S == 8) {
        const u32x2_t d = {v.x, v.y};
        __builtin_amdgcn_raw_buffer_store_b64(d, rs, int(off), 0, kStoreNT);
    } else {
        __builtin_amdgcn_raw_buffer_store_b32(v, rs, int(off), 0, kStoreNT);
    }
}

template <int S>
__device__ __forceinline__ typename HeadVec<S>::T load_head(const uint8_t *base, uint32_t v)
{
    return *reinterpret_cast<const typename HeadVec<S>::T *>(base + size_t(v) * uint32_t(S));
}

template <int S>
__device__ __forceinline__ bool head_nonzero(const typename HeadVec<S>::T &h)
{
    if constexpr (S == 16)
        return (h.x | h.y) != 0u;
    else
        return vec_nonzero<S>(h);
}

// Bit 4s set for every bit s of a 16-bit mask.
__device__ __forceinline__ uint64_t spread_nibbles(uint32_t m)
{
    uint64_t x = m & 0xFFFFu;
    x = (x | (x << 24)) & 0x000000FF000000FFull;
    x = (x | (x << 12)) & 0x000F000F000F000Full;
    x = (x | (x << 6)) & 0x0303030303030303ull;
    x = (x | (x << 3)) & 0x1111111111111111ull;
    return x;
}

// `val` inserted at byte position d < 16; the bytes from d up move one place up (the top byte is lost).
__device__ __forceinline__ BV bv_insert(const BV &v, uint32_t d, uint32_t val)
{
    const BV m = bytes_below(d), sh = shl_bytes(v, 1u);
    const BV r{(v.lo & m.lo) | (sh.lo & ~m.lo), (v.hi & m.hi) | (sh.hi & ~m.hi)};
    return bv_set(r, d, val);
}

// The arrivals of voxel V = (x, y, z) recomputed from global memory (the neighbours' slots and labels), in
// the canonical order: from -x, -y, -z into `lo`, from +z, +y, +x into `hi`. Loops are kept rolled: this
// serves planes whose event list overflowed, and must not add registers to the main path.
template <int S>
__device__ __forceinline__ void gather_arrivals_rolled(const StepArgs &a, const Tab &tb, int x, uint32_t y,
                                                       uint32_t z, uint32_t V, Slots<S> &lo, uint32_t &nlo,
                                                       Slots<S> &hi, uint32_t &nhi)
{
#pragma unroll 1
    for (uint32_t k = 0; k < 6u; ++k) {
        // k: source -x, -y, -z (hopped +x, +y, +z), then +z, +y, +x (hopped -z, -y, -x)
        const uint32_t want = (0x024531u >> (4u * k)) & 0xFu;
        const bool exists = k == 0u ? x > 0 : k == 1u ? y > 0u : k == 2u ? z > 0u
                          : k == 3u ? z + 1u < a.R2 : k == 4u ? y + 1u < a.R1 : uint32_t(x) + 1u < a.R0;
        if (!exists)
            continue;
        const uint32_t u = k == 0u ? V - a.R12 : k == 1u ? V - a.R2 : k == 2u ? V - 1u
                         : k == 3u ? V + 1u : k == 4u ? V + a.R2 : V + a.R12;
        const uint32_t ux = k == 0u ? uint32_t(x) - 1u : k == 5u ? uint32_t(x) + 1u : uint32_t(x);
        const uint32_t uy = k == 1u ? y - 1u : k == 4u ? y + 1u : y;
        const uint32_t uz = k == 2u ? z - 1u : k == 3u ? z + 1u : z;
        const Slots<S> su = load_slots<S>(a.src, u);
        const uint32_t occ = count_nonzero(su);
        if (occ == 0u)
            continue;
        const Hood h = make_hood(a, ux, uy, uz, u, S);
#pragma unroll 1
        for (uint32_t s = 0; s < occ; ++s) {
            const uint32_t sp = get_byte_dyn(su, s);
            if (hop(a, tb, h, sp, s) == want) {
                if (k < 3u) {
                    put_byte_dyn(lo, nlo, sp);
                    ++nlo;
                } else {
                    put_byte_dyn(hi, nhi, sp);
                    ++nhi;
                }
            }
        }
    }
}

// The byte-vector helpers on a voxel's first four slots held in one word (positions 0..3).
__device__ __forceinline__ uint32_t b4_below(uint32_t d) { return d >= 4u ? ~0u : ((1u << (8u * d)) - 1u); }
__device__ __forceinline__ uint32_t b4_shl(uint32_t v, uint32_t k) { return k >= 4u ? 0u : v << (8u * k); }
__device__ __forceinline__ uint32_t b4_shr(uint32_t v, uint32_t k) { return k >= 4u ? 0u : v >> (8u * k); }
__device__ __forceinline__ uint32_t b4_byte(uint32_t v, uint32_t j) { return (v >> (8u * (j & 3u))) & 0xFFu; }
__device__ __forceinline__ uint32_t b4_set(uint32_t v, uint32_t d, uint32_t val)
{
    const uint32_t sh = 8u * (d & 3u);
    return (v & ~(0xFFu << sh)) | (val << sh);
}
__device__ __forceinline__ uint32_t b4_delete(uint32_t v, uint32_t d)
{
    const uint32_t m = b4_below(d);
    return (v & m) | ((v >> 8) & ~m);
}
// `val` inserted at position d (< 4); the bytes from d up move one place (the top one is lost).
__device__ __forceinline__ uint32_t b4_insert(uint32_t v, uint32_t d, uint32_t val)
{
    const uint32_t m = b4_below(d);
    return d >= 4u ? v : b4_set((v & m) | ((v << 8) & ~m), d, val);
}

// assemble() for a voxel that ends with at most four particles: the same canonical picture on one word.
__device__ __forceinline__ uint32_t assemble4(uint32_t own, uint32_t occ, uint32_t stay_mask, uint32_t lo,
                                              uint32_t nlo, uint32_t hi, uint32_t nhi, uint32_t &cnt)
{
    uint32_t o = own;
    uint32_t rem = (occ >= 4u ? 0x1111u : ((1u << (4u * occ)) - 1u) & 0x1111u) & ~stay_mask;
    uint32_t used = 0;
    while (rem != 0u && used < nhi) {
        const uint32_t sl = uint32_t(__ffs(rem) - 1) >> 2;
        rem &= rem - 1u;
        o = b4_set(o, sl, b4_byte(hi, used));
        ++used;
    }
    const uint32_t j = occ - uint32_t(__popc(rem));
    while (rem != 0u) {  // highest first, so the lower positions stay valid
        const uint32_t bit = 31u - uint32_t(__clz(rem));
        rem &= ~(1u << bit);
        o = b4_delete(o, bit >> 2);
    }
    o |= b4_shl(lo, j) | b4_shl(b4_shr(hi, used), j + nlo);
    cnt = j + nlo + (nhi - used);
    return o;
}

// Event-chain heads, 16 bits per voxel in 32-bit words: exchange (returns the previous head) and read. The first
// compare-and-swap takes both halves of the word as still empty (kWEvEnd), true for most arrivals, so that no read
// of the word goes before it; when the guess is wrong the swap returns the word and is retried with it.
__device__ __forceinline__ uint32_t head_exchange(uint32_t *heads, uint32_t v, uint32_t idx)
{
    uint32_t *word = heads + (v >> 1);
    const uint32_t sh = (v & 1u) * 16u;
    uint32_t old = kWEvEnd | (kWEvEnd << 16), assumed;
    do {
        assumed = old;
        old = atomicCAS(word, assumed, (assumed & ~(0xFFFFu << sh)) | (idx << sh));
    } while (old != assumed);
    return (assumed >> sh) & 0xFFFFu;
}
__device__ __forceinline__ uint32_t head_read(const uint32_t *heads, uint32_t v)
{
    return (heads[v >> 1] >> ((v & 1u) * 16u)) & 0xFFFFu;
}

template <int S>
struct WaveStaged {
    typename HeadVec<S>::T rec[kWBatches];
    uint32_t lab[kWLabLane];
};

// The occupancy summaries of one march (SumRef): bit i of act is plane xb - 1 + i, set when a record of that plane
// (this tile's voxels or the four face-neighbour tiles') may hold a particle; bit i of old is plane xb + i, set
// when the destination may hold one there. One byte load per lane and tile, then ballots.
struct MarchSummary {
    uint64_t act, old;
};

__device__ __forceinline__ MarchSummary march_summary(const StepArgs &a, const WaveArgs &w, uint32_t tz, uint32_t wy,
                                                      int xb, int xe)
{
    const int l = int(__lane_id());
    const uint32_t P = a.sum_dst.planes;
    const int xa = xb - 1 + l;
    bool any = false;
    if (xa >= 0 && xa <= xe && xa < int(P)) {
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const int ty = int(wy) + (q == 1 ? -1 : (q == 2 ? 1 : 0));
            const int tq = int(tz) + (q == 3 ? -1 : (q == 4 ? 1 : 0));
            if (ty >= 0 && ty < int(w.wtiles_y) && tq >= 0 && tq < int(w.tiles_z))
                any = any || w.sum_src[(uint32_t(ty) * w.tiles_z + uint32_t(tq)) * P + uint32_t(xa)] != 0u;
        }
    }
    const uint8_t *own = a.sum_dst.bytes + (wy * w.tiles_z + tz) * P;
    return MarchSummary{__ballot(any), __ballot(xb + l < xe && own[xb + l] != 0u)};
}

// A march with nothing to load and nothing to clear does nothing, except that an empty voxel's zero-propensity
// draw still fires at seed 0 (itime + x*y + z == 0): counted here (lane 0) for the skipped planes.
__device__ __forceinline__ void count_idle_fires(const StepArgs &a, uint32_t tz, uint32_t wy, int xb, int xe,
                                                 TileCounts &tc)
{
    const int l = int(__lane_id());
    const int y0 = int(wy) * kWY, z0 = int(tz) * kWZ;
    // one lane per plane Q of the range, row by row: the voxel (Q, y0 + r, z0 + zfire) fires if it lies in the
    // lattice (bit zfire of the row's in-lattice lanes)
    uint32_t nf = 0u;
#pragma unroll
    for (int r = 0; r < kWY; ++r) {
        const uint64_t im = __ballot(y0 + r < int(a.R1) && z0 + l < int(a.R2));
        for (int qb = xb; qb < xe; qb += 64) {
            const int Q = qb + l;
            const uint32_t zfire = 0u - a.itime - uint32_t(Q) * uint32_t(y0 + r) - uint32_t(z0);
            bool f = Q < xe && zfire < uint32_t(kWZ) && ((im >> (zfire & 63u)) & 1ull) != 0ull;
            if (f)  // a skewed voxel's reaction runs in the replay kernel
                f = !live_voxel(a, a.vox_base + z0 + zfire + uint32_t(y0 + r) * a.R2 + uint32_t(Q) * a.R12);
            nf += uint32_t(__popcll(__ballot(f)));
        }
    }
    if (l == 0)
        tc.fired += nf;
}

// One wave tile (tz, wy) of the lattice (a, w), marched through x-chunk `chunk`.
template <int S>
__device__ __forceinline__ void wave_march(WaveLds<S> &L, const Tab &tb, const StepArgs &a, const WaveArgs &w,
                                           uint32_t tz, uint32_t wy, int xb, int xe, TileCounts &tc)
{
    using Vec = typename SlotVec<S>::T;
    using HVec = typename HeadVec<S>::T;
    // the tile and range are the same on every lane: scalar registers
    tz = __builtin_amdgcn_readfirstlane(tz);
    wy = __builtin_amdgcn_readfirstlane(wy);
    xb = __builtin_amdgcn_readfirstlane(xb);
    xe = __builtin_amdgcn_readfirstlane(xe);
    const int l = int(__lane_id());
    const int y0 = int(wy) * kWY, z0 = int(tz) * kWZ;
    const int R0 = int(a.R0), R1 = int(a.R1), R2 = int(a.R2);

    // records of this lane: k < kWY interior row k; kWY, kWY+1 halo rows -1 and kWY; the last one a z-halo
    // record for lanes < 2*kWY (rows 0..kWY-1 at column -1, then at column kWZ)
    uint32_t off[kWBatches];
    uint32_t okm = 0u, inl = 0u;  // record inside the lattice; interior voxel inside the lattice
#pragma unroll
    for (int k = 0; k < kWBatches; ++k) {
        int row = k, col = l;
        if (k == kWY) row = -1;
        if (k == kWY + 1) row = kWY;
        if (k == kWY + 2) {
            row = (l < kWY) ? l : l - kWY;
            col = (l < kWY) ? -1 : kWZ;
        }
        const int Y = y0 + row, Z = z0 + col;
        const bool ok = (k < kWY + 2 || l < 2 * kWY) && Y >= 0 && Y < R1 && Z >= 0 && Z < R2;
        off[k] = ok ? uint32_t(Z + Y * R2) : 0u;
        okm |= uint32_t(ok) << k;
        if (k < kWY)
            inl |= uint32_t(ok) << k;
    }
    // label words of this lane: word i of the plane's (kWY+4) x 17-word grid is row i / 17, word i % 17
    uint32_t loff[kWLabLane];
#pragma unroll
    for (int i = 0; i < kWLabLane; ++i) {
        const int wd = min(l + 64 * i, kWLabWords - 1);
        loff[i] = uint32_t((y0 + wd / (kWLZ / 4)) * int(w.PZ4) + z0 / 4 + wd % (kWLZ / 4));
    }
    const uint32_t plane_words = w.PY * w.PZ4;
    auto lab_base = [&](int X) { return uint32_t(min(max(X, -2), R0 + 1) + 2) * plane_words; };
    auto slot_base = [&](int X) { return (X >= 0 && X < R0) ? uint32_t(X) * a.R12 : 0u; };

    // X % Y of the record rows -1 .. kWY on lanes 0 .. kWY+1 (the reference's seed term, rdme.wgsl:220-226);
    // kept for the plane being hopped, advanced by one per plane
    const int xrow = y0 + l - 1;
    const bool xrow_ok = l < kWY + 2 && xrow > 0 && xrow < R1;
    uint32_t xm = 0u;
    if (xrow_ok)
        xm = xb - 1 >= 0 ? uint32_t(xb - 1) % uint32_t(xrow) : uint32_t(xrow - 1);
    const uint32_t xm_zero = a.mod_zero;
    // the label planes xb-2 and xb-1, loaded before the summaries' loads so that one wait covers both
    uint32_t lab0[kWLabLane], lab1[kWLabLane];
#pragma unroll
    for (int i = 0; i < kWLabLane; ++i) {
        lab0[i] = w.regp[lab_base(xb - 2) + loff[i]];
        lab1[i] = w.regp[lab_base(xb - 1) + loff[i]];
    }

    // Occupancy summaries: a plane whose records (this tile's voxels and the four face-neighbour tiles') hold
    // no particle is neither loaded nor listed; a plane gathered empty where the destination was empty
    // already is not stored; a march with neither is skipped. Bit i of actm: plane xb - 1 + i; of oldm: xb + i.
    const bool summ = w.sum_src != nullptr && a.sum_dst.bytes != nullptr && xe - xb <= 60;
    uint64_t actm = ~0ull, oldm = ~0ull;
    uint8_t *own_sum = nullptr;
    if (summ) {
        const MarchSummary ms = march_summary(a, w, tz, wy, xb, xe);
        actm = ms.act;
        oldm = ms.old;
        own_sum = a.sum_dst.bytes + (wy * w.tiles_z + tz) * a.sum_dst.planes;
    }
    auto pact = [&](int X) -> bool {
        const int k = X - (xb - 1);
        return k >= 0 && k < 64 && ((actm >> k) & 1ull) != 0ull;
    };
    auto pold = [&](int Q) -> bool {
        const int k = Q - xb;
        return k >= 0 && k < 64 && ((oldm >> k) & 1ull) != 0ull;
    };
    if (summ && w.has_zero_order == 0u && (actm | oldm) == 0ull) {
        count_idle_fires(a, tz, wy, xb, xe, tc);
        return;
    }

    // labels of planes xb-2 and xb-1; slots of plane xb-1 and labels of plane xb into the first register set
    auto store_lab = [&](int X, const uint32_t (&r)[kWLabLane]) {
        uint32_t *dst = L.lab[ring3(X)];
#pragma unroll
        for (int i = 0; i < kWLabLane; ++i)
            if (i + 1 < kWLabLane || l + 64 * i < kWLabWords)
                dst[l + 64 * i] = r[i];
    };
    WaveStaged<S> st0, st1;
    store_lab(xb - 2, lab0);
    store_lab(xb - 1, lab1);
#pragma unroll
    for (int k = 0; k < kWBatches; ++k)  // the side plane xb - 1: interior rows only (the other loads at one line)
        st0.rec[k] = load_head<S>(a.src, k < kWY ? slot_base(xb - 1) + off[k] : (off[k] & 7u));
#pragma unroll
    for (int i = 0; i < kWLabLane; ++i)
        st0.lab[i] = w.regp[lab_base(xb) + loff[i]];
    {
        // four stores that store nothing: with them, the memory operations in flight when the march enters its
        // plane loop are those of every later plane (a prefetch, then the row stores), and the first plane's wait
        // for its registers is the same as every other's, which leaves the previous plane's stores in flight (with
        // fewer operations on this path the compiler's wait at the top of the loop covered some of those stores)
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc(uniform_ptr(a.dst), 0, int(a.R12 * uint32_t(S)), kBufferRsrcWord3);
#pragma unroll
        for (int r = 0; r < kWY; ++r)  // at different offsets, which the compiler does not merge into one store
            buffer_store_vec<S>(Vec{}, rs, kBufferDrop + uint32_t(r) * 1024u);
    }
    if (l < 3)
        L.ev_n[l] = 0u;
#pragma unroll
    for (int q = 0; q < 3 * kWY / 2; ++q)
        L.head[q / (kWY / 2)][(q % (kWY / 2)) * 64 + l] = kWEvEnd | (kWEvEnd << 16);

    uint32_t n_tight = 0, n_near = 0, n_fired = 0, n_rolled = 0;
    // bit r of this lane: its voxel of interior row r holds particles, in the plane gathered next and in the
    // plane listed now (per-lane bits rather than wave masks: these live through the march, and scalar
    // registers are the scarcer ones here)
    uint32_t occ_prev = 0u, occ_cur = 0u;
    HVec prev[kWY];          // interior slots of that plane: the gather reads no global memory, where a
                             // load would also wait for the prefetch issued before it (loads return in order)
#pragma unroll
    for (int r = 0; r < kWY; ++r)
        prev[r] = HVec{};
    // Reactions whose draw passes the bound test are rare in sparse tiles and common in dense ones; either way
    // only a few lanes of a gather pass have one. They wait in a circular list and run 64 at a time.
    uint32_t rb_head = 0u, rb_n = 0u;
    auto react_flush = [&](uint32_t k) __attribute__((always_inline)) {
        wave_sync();
        const bool act = uint32_t(l) < k;
        uint32_t idx = rb_head + uint32_t(l);
        idx -= idx >= uint32_t(kWReact) ? uint32_t(kWReact) : 0u;
        Slots<S> out = from_vec<S>(act ? L.rb_slots[idx] : Vec{});
        const uint32_t V = L.rb_v[idx], xc = L.rb_xc[idx], yz = L.rb_yz[idx];
        if (act) {
            bool fired, rolled, near, bad;
            react<S>(a, tb, out, xc >> 24, xc & 0xFFFFFFu, yz >> 16, yz & 0xFFFFu, fired, rolled, near, bad);
            n_fired += fired;
            n_rolled += rolled;
            n_near += near;
            if (bad)  // a count taken below zero: the replay kernel redoes this voxel
                note_tight(a, V);
            store_slots<S>(a.dst, V, out);
        }
        rb_head += k;
        rb_head -= rb_head >= uint32_t(kWReact) ? uint32_t(kWReact) : 0u;
        rb_n -= k;
        wave_sync();
    };

    auto plane = [&](int X, WaveStaged<S> &use, WaveStaged<S> &fill) __attribute__((always_inline)) {
        // The per-lane record bits are taken as new on every plane: otherwise the compiler hoists one lane mask per
        // record batch out of the march into scalar registers, which then spill and are read back on every plane;
        asm volatile("" : "+v"(okm), "+v"(inl));
        // the lane id, likewise: values derived from it are recomputed in the plane rather than kept (and spilled)
        int l = int(__lane_id());
        asm volatile("" : "+v"(l));
        // ---- prefetch plane X+1's slots and plane X+2's labels (clamped; unused past the chunk)
        {
            // planes past xe are never listed nor gathered; on the side plane xe only the interior rows are
            // listed (a particle there reaches the tile with a hop along x or not at all), and a plane the
            // summaries mark empty is not listed either. The loads are issued on every plane all the same, the
            // ones not needed at one line (the lattice's first record, the label volume's first word), and their
            // registers are then never read: with the same number of loads on every path, the wait for the
            // previous plane's registers below leaves all of this plane's loads in flight (with some loads on some
            // paths only, the compiler has to assume the fewest and waits for part of this prefetch too). The
            // unneeded loads of one batch read different words of one line, so that no two of them are the same
            // load (which the compiler would issue once and copy, waiting for it).
            const uint32_t sb = slot_base(X + 1), lb = lab_base(X + 2);
            const bool go = X + 1 <= xe && (!summ || pact(X + 1));
            const int nb = !go ? 0 : (X + 1 == xe ? kWY : kWBatches);
#pragma unroll
            for (int k = 0; k < kWBatches; ++k)
                fill.rec[k] = load_head<S>(a.src, k < nb ? sb + off[k] : (off[k] & 7u));
            // labels up to plane xe + 1 (the hops of plane xe look at them)
            const bool lgo = X + 1 <= xe;
#pragma unroll
            for (int i = 0; i < kWLabLane; ++i)
                fill.lab[i] = w.regp[lgo ? lb + loff[i] : (loff[i] & 31u)];
        }
        store_lab(X + 1, use.lab);
        // plane X+1's event list starts empty (it was consumed as plane X-2's); plane X's leave masks too
        if (l == 0)
            L.ev_n[ring3(X + 1)] = 0u;
#pragma unroll
        for (int q = 0; q < kWY / 2; ++q)
            L.head[ring3(X + 1)][q * 64 + l] = kWEvEnd | (kWEvEnd << 16);
#pragma unroll
        for (int q = 0; q < kWY / 2; ++q)
            L.leave[X & 1][l + 64 * q] = 0u;
        wave_sync();

        // ---- T0: list the occupied records of plane X
        uint32_t n = 0u;
        const bool hop_plane = X >= 0 && X < R0 && X <= xe && (!summ || pact(X));
        if (hop_plane) {
            // on the planes beside the x-range only the interior rows matter: a particle there reaches the
            // tile with a hop along x or not at all
            const uint32_t bmask = (X == xb - 1 || X == xe) ? (1u << kWY) - 1u : (1u << kWBatches) - 1u;
            // this lane's occupied records as bits; each lane writes its records to consecutive list entries,
            // placed by one exclusive prefix sum of the per-lane counts (list order is free: see T1)
            uint32_t om = 0u;
#pragma unroll
            for (int k = 0; k < kWBatches; ++k)
                om |= uint32_t(head_nonzero<S>(use.rec[k])) << k;
            om &= okm & bmask;
            uint32_t pos = wave_offset(__popc(om), n);
#pragma unroll
            for (int k = 0; k < kWBatches; ++k) {
                if ((om >> k) & 1u) {
                    uint32_t rec;
                    if (k < kWY) rec = uint32_t((k + 1) * kWRZ + l + 1);
                    else if (k == kWY) rec = uint32_t(l + 1);
                    else if (k == kWY + 1) rec = uint32_t((kWY + 1) * kWRZ + l + 1);
                    else rec = uint32_t((l < kWY ? l + 1 : l - kWY + 1) * kWRZ + (l < kWY ? 0 : kWZ + 1));
                    L.hop_slots[pos] = use.rec[k];
                    L.hop_rec[pos] = uint16_t(rec | (head_full<S>(use.rec[k]) ? 0x8000u : 0u));
                    ++pos;
                }
            }
            occ_cur |= om & ((1u << kWY) - 1u);
        }
        wave_sync();

        // ---- T1: hops of plane X
        if (n != 0u) {
            const uint8_t *lprev = reinterpret_cast<const uint8_t *>(L.lab[ring3(X - 1)]);
            const uint8_t *lcur = reinterpret_cast<const uint8_t *>(L.lab[ring3(X)]);
            const uint8_t *lnext = reinterpret_cast<const uint8_t *>(L.lab[ring3(X + 1)]);
            const uint32_t n1 = a.ns + 1u;
            // event lists of planes X-1, X, X+1: nibbles 0, 1, 2
            const uint32_t rings = uint32_t(ring3(X - 1)) | uint32_t(ring3(X)) << 4 | uint32_t(ring3(X + 1)) << 8;
            for (uint32_t base = 0; base < n; base += 64u) {
                const uint32_t i = base + uint32_t(l);
                const bool act = i < n;
                const uint32_t rcode = act ? L.hop_rec[i] : uint32_t(kWRZ + 1);
                const uint32_t rec = rcode & 0x7FFFu;
                Slots<S> sl = from_head<S>(act ? L.hop_slots[i] : HVec{});
                const int rho = int(rec / kWRZ) - 1, zeta = int(rec % kWRZ) - 1;
                const uint32_t Y = uint32_t(y0 + rho), Z = uint32_t(z0 + zeta);
                if constexpr (S > kHeadSlots<S>) {
                    const uint64_t fm = __ballot((rcode & 0x8000u) != 0u);  // full heads: whole records (rare)
                    if (fm != 0ull) {
                        full_records<S>(a.src, Z + __umul24(Y, a.R2) + uint32_t(X) * a.R12, fm, sl);
                        vm_drain();
                    }
                }
                const int gi = (rho + 2) * kWLZ + (zeta + 2);
                Hood h;
                h.src_reg = lcur[gi];
                const uint32_t rg[6] = {lprev[gi], lnext[gi], lcur[gi - kWLZ], lcur[gi + kWLZ], lcur[gi - 1],
                                        lcur[gi + 1]};
                h.inside = 0u;
#pragma unroll
                for (int d = 0; d < 6; ++d) {
                    const bool in = rg[d] != 0xFFu;
                    h.inside |= uint32_t(in) << d;
                    h.rg[d] = in ? rg[d] : 0u;
                }
                if (h.src_reg == 0xFFu)
                    h.src_reg = 0u;  // an empty out-of-lattice filler lane
                const uint32_t xmr = __shfl(xm, rho + 1);
                const uint32_t xmy = Y == 0u ? xm_zero : xmr;
                h.seed0 = a.itime + xmy + Z + (Z + __umul24(Y, a.R2) + uint32_t(X) * a.R12) * uint32_t(S);
                const uint32_t occ = act ? count_compacted(sl) : 0u;
                const bool interior = rho >= 0 && rho < kWY && zeta >= 0 && zeta < kWZ;
                uint32_t lv = 0u;
                for (uint32_t j = 0;; ++j) {
                    const bool pj = j < occ;
                    if (__ballot(pj) == 0ull)
                        break;
                    const uint32_t sp = get_byte_dyn(sl, j);
                    const uint32_t d = pj ? hop(a, tb, h, sp, j) : 6u;
                    (void)n1;
                    const bool moves = d < 6u;
                    lv |= uint32_t(moves) << j;
                    // the step along x, y, z of direction d, plus one: nibble d of each constant (d = 6: none)
                    const uint32_t dsh = 4u * d;
                    const uint32_t dx1 = (0x1111120u >> dsh) & 0xFu;
                    const int r2 = rho + int((0x1112011u >> dsh) & 0xFu) - 1;
                    const int z2 = zeta + int((0x1201111u >> dsh) & 0xFu) - 1;
                    const bool ev = moves && uint32_t(r2) < uint32_t(kWY) && uint32_t(z2) < uint32_t(kWZ) &&
                                    uint32_t(X + int(dx1) - 1 - xb) < uint32_t(xe - xb);
                    // One event per arriving particle, appended to its destination plane's list with an LDS
                    // atomic (list order is free: the per-voxel chains carry the order, and within one slot index
                    // j no two events reach one voxel from one direction, so the chains see the source-slot order).
                    if (ev) {
                        const uint32_t dest = uint32_t(r2 * kWZ + z2);
                        const int rg3 = int((rings >> (4u * dx1)) & 0xFu);
                        const uint32_t idx = atomicAdd(&L.ev_n[rg3], 1u);
                        if (idx < w.ev_cap) {
                            const uint32_t prev = head_exchange(L.head[rg3], dest, idx);
                            L.ev[rg3][idx] = d | (sp << 3) | (prev << 11);
                        }
                    }
                }
                if (act && interior && lv != 0u)
                    reinterpret_cast<uint16_t *>(L.leave[X & 1])[rho * kWZ + zeta] = uint16_t(lv);
            }
        }
        wave_sync();

        // ---- T2: gather plane Q = X-1
        const int Q = X - 1;
        const bool t2 = Q >= xb && Q < xe;
        uint32_t ns = 0u, needm = 0u, sidx[kWY];  // voxels to gather: count, this lane's rows, list positions
#pragma unroll
        for (int r = 0; r < kWY; ++r)
            sidx[r] = 0u;
        if (t2) {
            const int rq = ring3(Q);
            // this lane's voxels: their chain heads (kWEvEnd: no arrival) and leave masks, read together with the
            // event count
            const uint16_t *hq = reinterpret_cast<const uint16_t *>(L.head[rq]);
            const uint16_t *lq = reinterpret_cast<const uint16_t *>(L.leave[Q & 1]);
            uint32_t hd[kWY], lvr[kWY];
#pragma unroll
            for (int r = 0; r < kWY; ++r) {
                hd[r] = hq[r * kWZ + l];
                lvr[r] = lq[r * kWZ + l];
            }
            // a plane whose event list overflowed gathers every interior voxel's arrivals from global memory
            const bool over = L.ev_n[rq] > w.ev_cap;
            // the voxels to gather: occupied, or a particle arrives (everything when the event list overflowed,
            // whose chains miss events, or when empty voxels can react)
            needm = (over || w.has_zero_order != 0u) ? (1u << kWY) - 1u : occ_prev;
#pragma unroll
            for (int r = 0; r < kWY; ++r)
                needm |= uint32_t(hd[r] != kWEvEnd) << r;
            needm &= inl;
#pragma unroll
            for (int r = 0; r < kWY; ++r) {
                // an empty voxel's zero-propensity draw fires at seed 0 (itime + x*y + z == 0)
                const uint32_t zfire = 0u - a.itime - uint32_t(Q) * uint32_t(y0 + r) - uint32_t(z0);
                if (zfire < uint32_t(kWZ)) {
                    const uint64_t inm = __ballot((inl >> r) & 1u), m = __ballot((needm >> r) & 1u);
                    if (l == 0 &&
                        !live_voxel(a, a.vox_base + z0 + zfire + uint32_t(y0 + r) * a.R2 + uint32_t(Q) * a.R12))
                        n_fired += uint32_t((inm & ~m) >> zfire) & 1u;
                }
            }
            // the voxels to gather, this lane's consecutively (one exclusive prefix sum of the counts)
            uint32_t pos = wave_offset(__popc(needm), ns);
#pragma unroll
            for (int r = 0; r < kWY; ++r) {
                sidx[r] = pos;
                if ((needm >> r) & 1u) {
                    const bool occ = (occ_prev >> r) & 1u;
                    // the entry: voxel | chain head << 8, the leave mask, the own slots (the hop list and its
                    // record ids are done with)
                    L.slow[pos] = uint16_t(uint32_t(r * kWZ + l) | (hd[r] << 8));
                    L.hop_rec[pos] = uint16_t(lvr[r]);
                    L.hop_slots[pos] = occ ? prev[r] : HVec{};
                    ++pos;
                }
            }
            wave_sync();
            for (uint32_t base = 0; base < ns; base += 64u) {
                const uint32_t i = base + uint32_t(l);
                const bool act = i < ns;
                const uint32_t code = act ? L.slow[i] : (kWEvEnd << 8);
                const uint32_t me = code & 0xFFu, e0 = code >> 8;
                const uint32_t lv = act ? L.hop_rec[i] : 0u;
                const int r = int(me) / kWZ, z = int(me) % kWZ;
                const uint32_t Y = uint32_t(y0 + r), Z = uint32_t(z0 + z);
                const uint32_t V = Z + __umul24(Y, a.R2) + uint32_t(Q) * a.R12;
                const HVec hv = act ? L.hop_slots[i] : HVec{};
                Slots<S> own = from_head<S>(hv);
                if constexpr (S > kHeadSlots<S>) {
                    const uint64_t fm = __ballot(act && head_full<S>(hv));  // full heads (rare)
                    if (fm != 0ull) {
                        full_records<S>(a.src, V, fm, own);
                        vm_drain();
                    }
                }
                const uint32_t occ = count_compacted(own);
                // Most voxels end with at most four particles: then the arrivals and the assembled slots fit
                // one 32-bit word, and the wave takes that path when every lane fits it.
                uint32_t cnt = 0u;
                Slots<S> out = zero_slots<S>();
                bool done = false, tight = false;
                if constexpr (S > 4) {
                    if (!over && __ballot(occ > 4u) == 0ull) {
                        uint32_t arr = 0u, tot = 0u;
                        // arrivals per rank, one nibble each: exact while the voxel gets at most 15 per rank, and
                        // a lane with more has more than four arrivals, so it leaves this path anyway
                        uint32_t cn = 0u;
                        uint32_t e = e0;
                        while (__ballot(e != kWEvEnd) != 0ull) {
                            if (e != kWEvEnd) {
                                const uint32_t c = L.ev[rq][e];
                                const uint32_t k = (0x231405u >> (4u * (c & 7u))) & 0xFu;
                                // arrivals of lower ranks so far: the nibbles below k summed (24-bit product)
                                const uint32_t pos = k ? (__umul24(cn, 0x111111u) >> (4u * k - 4u)) & 0xFu : 0u;
                                arr = b4_insert(arr, pos, (c >> 3) & 0xFFu);
                                cn += 1u << (4u * k);
                                ++tot;
                                e = c >> 11;
                            }
                        }
                        if (__ballot(occ + tot > 4u) == 0ull) {
                            const uint32_t nlo4 = (__umul24(cn, 0x111111u) >> 8) & 0xFu;  // ranks 0..2: from below
                            // own slots that stay: bit 4s for each slot s < occ whose particle did not leave
                            uint32_t sp4 = lv & 0xFu;
                            sp4 = (sp4 | (sp4 << 6)) & 0x0303u;
                            sp4 = (sp4 | (sp4 << 3)) & 0x1111u;
                            const uint32_t stay4 = ((1u << (4u * occ)) - 1u) & 0x1111u & ~sp4;
                            out.w[0] = assemble4(own.w[0], occ, stay4, arr & b4_below(nlo4),
                                                 nlo4, b4_shr(arr, nlo4), tot - nlo4, cnt);
                            done = true;
                        }
                    }
                }
                if (!done) {
                    Slots<S> lo = zero_slots<S>(), hi = zero_slots<S>();
                    uint32_t nlo = 0u, nhi = 0u;
                    if (!over) {
                        // arrivals in the canonical order: from -x, -y, -z (lower-index sources), then from +z,
                        // +y, +x, each in source-slot order. The chain holds a voxel's events newest first, so
                        // each event goes in front of the earlier events of its rank (rank of hop direction d
                        // in that order = nibble d of 0x231405).
                        BV arr{0ull, 0ull};
                        uint64_t cnts = 0ull;  // arrivals per rank, one byte each
                        uint32_t e = e0;
                        while (__ballot(e != kWEvEnd) != 0ull) {
                            if (e != kWEvEnd) {
                                const uint32_t c = L.ev[rq][e];
                                const uint32_t k = (0x231405u >> (4u * (c & 7u))) & 0xFu;
                                const uint32_t pos =
                                    k ? uint32_t(((cnts * 0x0101010101010101ull) >> (8u * (k - 1u))) & 0xFFull) : 0u;
                                if (pos < 16u)
                                    arr = bv_insert(arr, pos, (c >> 3) & 0xFFu);
                                cnts += 1ull << (8u * k);
                                e = c >> 11;
                            }
                        }
                        nlo = uint32_t((cnts * 0x010101ull) >> 16) & 0xFFu;
                        nhi = uint32_t(((cnts >> 24) * 0x010101ull) >> 16) & 0xFFu;
                        const BV lom = bytes_below(nlo);
                        lo = from_bv<S>(BV{arr.lo & lom.lo, arr.hi & lom.hi});
                        hi = from_bv<S>(shr_bytes(arr, nlo));
                    } else {
                        if (act)
                            gather_arrivals_rolled<S>(a, tb, Q, Y, Z, V, lo, nlo, hi, nhi);
                        vm_drain();  // the global loads above complete here, not at the join below
                    }
                    tight = occ + nlo + nhi > uint32_t(S);
                    if (act && tight) {
                        ++n_tight;
                        note_tight(a, V);
                    }
                    const uint64_t stay = first_nibbles(occ) & ~spread_nibbles(lv);
                    out = assemble<S>(own, occ, occ ? stay : 0ull, lo, min(nlo, 16u), hi, min(nhi, 16u), cnt);
                }
                const bool may = act && react_may_fire<S>(a, tb, out, cnt, uint32_t(Q), Y, Z);
                // the result goes back to the list entry; the voxel's own lane stores whole rows below, so no
                // output line is written in pieces at different times (a reaction still to run is stored
                // by its batch instead, flagged here)
                // an output with more particles than a head holds is stored here (rare)
                const bool wide = act && !may && cnt > uint32_t(kHeadSlots<S>);
                if (wide)
                    store_slots<S>(a.dst, V, out);
                if (act) {
                    L.hop_slots[i] = to_head<S>(out);
                    L.slow[i] = uint16_t((may || wide) ? 1u : 0u);
                }
                const uint64_t mr = __ballot(may);
                if (mr != 0ull) {
                    const uint32_t rank =
                        __builtin_amdgcn_mbcnt_hi(uint32_t(mr >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mr), 0u));
                    if (may) {
                        uint32_t idx = rb_head + rb_n + rank;
                        idx -= idx >= uint32_t(kWReact) ? uint32_t(kWReact) : 0u;
                        L.rb_slots[idx] = to_vec<S>(out);
                        L.rb_v[idx] = V;
                        L.rb_xc[idx] = uint32_t(Q) | (cnt << 24);
                        L.rb_yz[idx] = (Y << 16) | Z;
                    }
                    rb_n += uint32_t(__popcll(mr));
                    while (rb_n >= uint32_t(kWReactRun))
                        react_flush(min(rb_n, 64u));
                }
            }
            wave_sync();
        }
        // ---- the output rows of plane Q, one store per row and lane on every plane (side planes included): a
        // lane with nothing to store gets an offset past the plane, which the buffer store drops. With the same
        // number of stores on every path, the wait for the next plane's prefetch leaves them in flight; with
        // stores on some paths only, that wait had to cover this plane's stores too, a write round trip per plane.
        {
            // an empty output over an empty destination needs no store
            const bool old_q = summ && pold(Q);
            const bool st = t2 && (!summ || ns != 0u || old_q);
            HVec h[kWY];
            uint32_t sl[kWY];
#pragma unroll
            for (int r = 0; r < kWY; ++r) {
                h[r] = HVec{};
                sl[r] = 0u;
            }
            if (ns != 0u) {  // (a plane with nothing gathered stores zeros, if anything)
#pragma unroll
                for (int r = 0; r < kWY; ++r) {  // every list read first (in bounds on every lane), then one wait
                    const uint32_t i = ((needm >> r) & 1u) ? sidx[r] : 0u;
                    h[r] = L.hop_slots[i];
                    sl[r] = L.slow[i];
                }
            }
            const uint32_t Qc = uint32_t(min(max(Q, 0), R0 - 1));
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                uniform_ptr(a.dst) + size_t(Qc * a.R12) * uint32_t(S), 0, int(a.R12 * uint32_t(S)), kBufferRsrcWord3);
            bool nz = false;
#pragma unroll
            for (int r = 0; r < kWY; ++r) {
                const bool need = (needm >> r) & 1u;  // implies a lattice voxel of a gathered plane
                const Vec v = to_vec<S>(from_head<S>(need ? h[r] : HVec{}));
                const bool deferred = need && sl[r] != 0u;
                const bool put = st && ((inl >> r) & 1u) && !deferred;
                buffer_store_vec<S>(v, rs, put ? off[r] * uint32_t(S) : kBufferDrop);
                nz = nz || deferred || (put && vec_nonzero<S>(v));
            }
            if (summ && t2) {
                const bool new_q = __ballot(nz) != 0ull;
                if (new_q != old_q && l == 0)
                    own_sum[Q] = uint8_t(new_q);
            }
        }
        occ_prev = occ_cur;
        occ_cur = 0u;
#pragma unroll
        for (int r = 0; r < kWY; ++r)
            prev[r] = use.rec[r];
        if (xrow_ok) {
            xm = xm + 1u;
            xm = xm == uint32_t(xrow) ? 0u : xm;
        }
        wave_sync();
    };
    // planes xb-1 .. xe in pairs; an odd count ends with plane xe+1, which lists and gathers nothing
    for (int X = xb - 1; X <= xe; X += 2) {
        plane(X, st0, st1);
        plane(X + 1, st1, st0);
    }
    while (rb_n != 0u)
        react_flush(min(rb_n, 64u));
    tc.tight += n_tight;
    tc.near += n_near;
    tc.fired += n_fired;
    tc.rolled += n_rolled;
}

// The plain step on wave tiles, as a persistent kernel: every block takes groups (kWavesPerBlock tiles
// stacked in y in one x-chunk, one per wave) from a work queue until it is empty. The waves of a group
// march side by side on one CU, so the halo rows one wave loads are the interior rows of its neighbours,
// found in L1/L2. A group's time grows with its particles, and a dense region holds several times the
// particles of the average group, so groups are queued by decreasing particle count: the longest start
// first and the short ones fill in behind them. The queue is split per XCD: XCD k takes the groups of the k-th
// band of block rows, whose neighbours' halo lines are then in its own L2; an XCD whose queue is empty helps
// the others.
//
// TM, where the model tables are read: kTabLds all staged in LDS; kTabSplit the hop probabilities and the propensity
// bounds (every particle and voxel reads them) staged, the reaction tables (read only by the rare reactions) in global
// memory through the caches; kTabGlobal all in global memory. The last two serve tables larger than the LDS a block
// has beside its plane records.
constexpr int kTabLds = 0, kTabSplit = 1, kTabGlobal = 2;

template <int S, int TM>
__global__ __launch_bounds__(kWaveBlock) __attribute__((amdgpu_waves_per_eu(3)))
void rdme_step_wave(StepArgs a0, WaveArgs w)
{
    StepArgs a = a0;
    if (*a0.skew_n == 0u)
        a.live = nullptr;  // no skewed voxel: the live-bit loads (and their waits) are skipped
    __shared__ WaveLds<S> L[kWavesPerBlock];
    extern __shared__ uint32_t tab_lds[];
    const uint8_t *gbase = reinterpret_cast<const uint8_t *>(w.tab);
    const uint8_t *lbase = reinterpret_cast<const uint8_t *>(tab_lds);
    if constexpr (TM == kTabLds) {
        for (uint32_t i = threadIdx.x; i < w.tab_words; i += kWaveBlock)
            tab_lds[i] = w.tab[i];
    } else if constexpr (TM == kTabSplit) {
        // the hop probabilities (from offset 0) and then the propensity bounds with the pair table
        const uint32_t np2 = w.off_rates / 4u, nb = bound_words(uint32_t(S), w.pair_n1);
        for (uint32_t i = threadIdx.x; i < np2 + nb; i += kWaveBlock)
            tab_lds[i] = i < np2 ? w.tab[i] : w.tab[w.off_bound / 4u + (i - np2)];
    }
    __syncthreads();
    const uint8_t *hb = TM == kTabGlobal ? gbase : lbase;  // hop probabilities and bounds
    const uint8_t *rb = TM == kTabLds ? lbase : gbase;     // reaction tables
    const uint32_t off_bound = TM == kTabSplit ? w.off_rates : w.off_bound;
    const float *bnd = reinterpret_cast<const float *>(hb + off_bound);
    const Tab tb{reinterpret_cast<const float *>(hb + w.off_p2), reinterpret_cast<const float *>(rb + w.off_rates),
                 reinterpret_cast<const uint32_t *>(rb + w.off_ridx),
                 reinterpret_cast<const int8_t *>(rb + w.off_stoich),
                 reinterpret_cast<const uint64_t *>(rb + w.off_rmask), bnd,
                 w.pair_n1 != 0u ? bnd + pair_offset(uint32_t(S)) : nullptr};
    const uint32_t wv = threadIdx.x / 64u;
    // XCC_ID hardware register (id 20, bits 3:0): the XCD this block runs on
    uint32_t xcd = uint32_t(__builtin_amdgcn_s_getreg((31 << 11) | 20)) & 0xFu;
    xcd = __builtin_amdgcn_readfirstlane(xcd < 8u ? xcd : blockIdx.x % 8u);
    __shared__ uint32_t s_item;
    TileCounts tc;
    for (uint32_t probe = 0; probe < 8u; ++probe) {
        const uint32_t k = (xcd + probe) & 7u;
        const uint32_t ngk = w.nitems > k ? (w.nitems - k + 7u) / 8u : 0u;
        for (;;) {
            if (threadIdx.x == 0)
                s_item = atomicAdd(&w.queue[k], 1u);
            __syncthreads();
            const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
            __syncthreads();
            if (it >= ngk)
                break;
            const uint2 item = w.items[it * 8u + k];
            const uint32_t tz = item.x & 0xFFFFu, by = item.x >> 16;
            const uint32_t wy = by * kWavesPerBlock + wv;
            if (wy < w.wtiles_y)
                wave_march<S>(L[wv], tb, a, w, tz, wy, int(item.y & 0xFFFFu), int(item.y >> 16), tc);
        }
    }
    flush_counts(a.counters, tc);
}

// ---- well-mixed (CME) phase --------------------------------------------------------------------------
// No counterpart in the reference. Once per step, before the lattice step, the well-mixed counts of a
// simulation advance by an exact SSA (Gillespie's direct method) over tau:
//   a_r = k_r * f32(h_r), h_r = cme_h (binomials of the reactant counts);
//   cumulative sums in f32 with Kahan compensation, read as f64 (sum - compensation);
//   waiting time -log(u1) / a0 and the simulated time in f64; choice: first r with a_r > 0 and
//   u2 * a0 < cumulative_r (u1 in (0, 1], u2 in [0, 1), 53 bits each, from cme_bits).
// A chosen reaction's lattice products are placed at voxels of its region drawn with cme_place_draw,
// up to 100 draws each, after the voxel's particles (the first draw with a free slot wins).
constexpr uint32_t kCmeMaxEvents = 1u << 20;  // per step; more stop the step and count an exception

struct CmeArgs {
    uint32_t nc, ncr;
    const int32_t *reactants;    // [ncr][3], -1 = none
    const int32_t *delta;        // [ncr][nc]
    const uint8_t *emit;         // [ncr][4] lattice species, 0 = none
    const int32_t *emit_region;  // [ncr]
    const float *rates;          // [ncr]
    const uint32_t *lists;       // region voxel lists, concatenated
    const uint32_t *list_off;    // [nreg]
    const uint32_t *list_len;    // [nreg]
    unsigned long long *counters;
    unsigned long long *audit;   // [3]: max relative error (f64 bits), divergences, audited events
    uint32_t audit_on;
    float tau;
};

__device__ __forceinline__ void kahan_add(float &s, float &c, float x)
{
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

template <int S>
__device__ void cme_cell(const CmeArgs &a, int32_t *counts, uint8_t *lattice, const SumRef &sm, uint32_t seed,
                         uint32_t itime)
{
    const double tau = double(a.tau);
    double t = 0.0, max_rel = 0.0;
    uint32_t e = 0, near = 0, fails = 0, div = 0, bad = 0;
    for (;;) {
        float s = 0.0f, c = 0.0f;
        double s64 = 0.0;
        for (uint32_t r = 0; r < a.ncr; ++r) {
            const double h = cme_h(a.reactants + 3u * r, counts);
            kahan_add(s, c, a.rates[r] * float(h));
            if (a.audit_on)
                s64 = s64 + double(a.rates[r]) * h;
        }
        const double a0 = double(s) - double(c);
        if (!(a0 > 0.0))
            break;
        if (e >= kCmeMaxEvents) {
            bad = 1;
            break;
        }
        const double u1 = double(cme_bits(itime, seed, 2u * e) + 1ull) * 0x1p-53;
        t = t + (-log(u1)) / a0;
        if (fabs(t - tau) <= tau * 1e-12)
            ++near;
        if (t >= tau)
            break;
        const double u2 = double(cme_bits(itime, seed, 2u * e + 1u)) * 0x1p-53;
        const double target = u2 * a0, target64 = u2 * s64;
        float s2 = 0.0f, c2 = 0.0f;
        double cum64 = 0.0;
        uint32_t chosen = a.ncr, last = a.ncr, chosen64 = a.ncr, last64 = a.ncr;
        for (uint32_t r = 0; r < a.ncr; ++r) {
            const double h = cme_h(a.reactants + 3u * r, counts);
            const float ar = a.rates[r] * float(h);
            kahan_add(s2, c2, ar);
            if (ar > 0.0f) {
                last = r;
                if (chosen == a.ncr && target < double(s2) - double(c2))
                    chosen = r;
            }
            if (a.audit_on) {
                const double ar64 = double(a.rates[r]) * h;
                cum64 = cum64 + ar64;
                if (ar64 > 0.0) {
                    last64 = r;
                    if (chosen64 == a.ncr && target64 < cum64)
                        chosen64 = r;
                }
            }
        }
        if (chosen == a.ncr)
            chosen = last;
        if (a.audit_on) {
            if (chosen64 == a.ncr)
                chosen64 = last64;
            div += chosen64 != chosen;
            const double rel = fabs(a0 - s64) / s64;
            max_rel = rel > max_rel ? rel : max_rel;
        }
        for (uint32_t k = 0; k < a.nc; ++k)
            counts[k] += a.delta[chosen * a.nc + k];
        const int32_t rg = a.emit_region[chosen];
        for (uint32_t j = 0; j < 4u; ++j) {
            const uint32_t sp = a.emit[4u * chosen + j];
            if (sp == 0u || rg < 0)
                continue;
            const uint32_t n = a.list_len[rg];
            const uint32_t *list = a.lists + a.list_off[rg];
            bool placed = false;
            for (uint32_t att = 0; att < 100u && n > 0u && !placed; ++att) {
                const uint32_t u = cme_place_draw(itime, seed, e, j, att);
                const uint32_t v = list[uint32_t((uint64_t(u) * n) >> 32)];
                Slots<S> sl = load_slots<S>(lattice, v);
                const uint32_t occ = count_nonzero(sl);
                if (occ < uint32_t(S)) {
                    put_byte_dyn(sl, occ, sp);
                    store_slots<S>(lattice, v, sl);
                    sum_mark(sm, v);
                    placed = true;
                }
            }
            fails += !placed;
        }
        ++e;
    }
    if (e)
        atomicAdd(a.counters + CS_CNT_CME_EVENTS, (unsigned long long)e);
    if (near)
        atomicAdd(a.counters + CS_CNT_CME_NEAR, (unsigned long long)near);
    if (fails)
        atomicAdd(a.counters + CS_CNT_EMIT_FAIL, (unsigned long long)fails);
    if (bad)
        atomicAdd(a.counters + CS_CNT_EXCEPTION, 1ull);
    if (a.audit_on && e) {
        atomicMax(a.audit, (unsigned long long)__double_as_longlong(max_rel));  // non-negative: bits order like values
        atomicAdd(a.audit + 1, (unsigned long long)div);
        atomicAdd(a.audit + 2, (unsigned long long)e);
    }
}

// One simulation's CME phase: a single lane (the pool is one well-mixed compartment).
template <int S>
__global__ __launch_bounds__(64) void cme_kernel(CmeArgs a, int32_t *counts, uint8_t *lattice, SumRef sm,
                                                 uint32_t seed, uint32_t itime)
{
    if (threadIdx.x == 0)
        cme_cell<S>(a, counts, lattice, sm, seed, itime);
}

// Per-region species histogram of the lattice (statistics; the reference keeps global counts only,
// src/simulation.rs:361-388 and cme.wgsl:87,105).
template <int S>
__global__ __launch_bounds__(kBlock) void region_count_kernel(const uint8_t *lat, const uint8_t *reg, uint32_t N,
                                                            uint32_t ns, uint32_t nreg, int32_t *hist)
{
    extern __shared__ int32_t sh[];
    const uint32_t nbins = nreg * (ns + 1u);
    for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x)
        sh[i] = 0;
    __syncthreads();
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
        const Slots<S> s = load_slots<S>(lat, v);
        if (count_nonzero(s) == 0u)
            continue;
        const uint32_t base = uint32_t(reg[v]) * (ns + 1u);
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const uint32_t sp = get_byte(s, k);
            if (sp != 0u && sp <= ns)
                atomicAdd(&sh[base + sp], 1);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x)
        if (sh[i])
            atomicAdd(&hist[i], sh[i]);
}


__global__ __launch_bounds__(kBlock) void fill_i32_kernel(int32_t *p, uint32_t n, int32_t v)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// Adds the count offsets of the skewed voxels (what their counts hold beyond their slots) to a region x species
// histogram.
__global__ __launch_bounds__(kBlock) void skew_region_kernel(SkewTab k, const uint8_t *reg, uint32_t ns, int32_t *hist)
{
    const uint32_t ne = k.ctl[0];
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < ne; i += gridDim.x * kBlock) {
        const SkewEntry e = k.ent[i];
        for (int q = 0; q < kSkewSp; ++q) {
            const uint32_t sp = (e.sp >> (8 * q)) & 0xFFu;
            if (sp != 0u && sp <= ns && e.dc[q] != 0)
                atomicAdd(hist + uint32_t(reg[e.v]) * (ns + 1u) + sp, e.dc[q]);
        }
    }
}

// ---- batched ensembles (see ensemble.hpp) ------------------------------------------------------------
struct EnsCell {
    uint64_t lat_off;          // byte offset of the cell's slots in each lattice buffer
    uint32_t vox_off;          // index of the cell's voxel 0 among all voxels (labels, tight list)
    uint32_t R0, R1, R2;
    uint32_t tiles_z, wtiles_y, btiles_y, chunks, first_item;  // wave tiles in z and y, groups per x-chunk, x-chunks
    uint32_t regp_off, PY, PZ4;  // the cell's padded labels: word offset, rows per plane, words per row
    uint32_t tab_off, tab_words;  // the cell's table blob, in 32-bit words of the tables buffer
    uint32_t off_p2, off_rates, off_ridx, off_stoich, off_rmask, off_bound, pair_n1;
    uint32_t ns, nr, nreg, nrw, has_zero_order, mod_zero;
    float tau;
    uint32_t itime_seed;       // itime(step) = low 32 bits of SplitMix64((itime_seed << 32) + step)
    int32_t cme;               // the cell's well-mixed pool (index of its CmeArgs), -1 = none
    uint32_t cme_counts_off, cme_seed;
    uint32_t sum_off, sum_planes;  // the cell's occupancy summary: byte offset in each summary buffer, planes per
                                   // tile (R0); 0: none (cells with zero-order reactions)
};

struct EnsArgs {
    const EnsCell *cells;
    uint32_t n_cells, total_items, xc, batch;
    uint8_t *src, *dst;
    const uint8_t *reg;
    const uint32_t *regp;      // padded labels of every cell (words)
    const uint32_t *tab;
    uint32_t *itime;           // [n_cells], written by the prologue of each step
    uint32_t *queue;           // [0] next item, [1] step counter
    unsigned long long *counters;
    uint32_t *fix_ctl, *fix_list;
    uint32_t max_tight;
    const uint32_t *live;         // skewed voxels (bit per voxel of every cell)
    const uint32_t *skew_n;       // entries of the skewed-voxel table
    uint8_t *sum_src, *sum_dst;   // occupancy summaries of src and dst (every cell's, at EnsCell::sum_off)
    const uint32_t *item_cell;    // [2 * total_items] the cell of each queue entry and its item within the cell
};

__device__ __forceinline__ SumRef cell_sum(uint8_t *bytes, const EnsCell &c)
{
    return SumRef{c.sum_planes != 0u ? bytes + c.sum_off : nullptr, c.sum_planes, c.tiles_z, c.R1, c.R2};
}

__device__ __forceinline__ uint32_t splitmix_low32(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return uint32_t(z ^ (z >> 31));
}

__device__ __forceinline__ StepArgs cell_step_args(const EnsArgs &e, const EnsCell &c, uint32_t itime)
{
    StepArgs a;
    a.src = e.src + c.lat_off;
    a.dst = e.dst + c.lat_off;
    a.reg = e.reg + c.vox_off;
    const uint8_t *tb = reinterpret_cast<const uint8_t *>(e.tab + c.tab_off);
    a.p2 = reinterpret_cast<const float *>(tb + c.off_p2);
    a.rates = reinterpret_cast<const float *>(tb + c.off_rates);
    a.ridx = tb + c.off_ridx;
    a.stoich = reinterpret_cast<const int8_t *>(tb + c.off_stoich);
    a.rmask = reinterpret_cast<const uint64_t *>(tb + c.off_rmask);
    a.bound = reinterpret_cast<const float *>(tb + c.off_bound);
    a.counters = e.counters;
    a.fix_ctl = e.fix_ctl;
    a.fix_list = e.fix_list;
    a.max_tight = e.max_tight;
    a.vox_base = c.vox_off;
    a.R0 = c.R0;
    a.R1 = c.R1;
    a.R2 = c.R2;
    a.R12 = c.R1 * c.R2;
    a.N = c.R0 * c.R1 * c.R2;
    a.ns = c.ns;
    a.nr = c.nr;
    a.nreg = c.nreg;
    a.nrw = c.nrw;
    a.itime = itime;
    a.mod_zero = c.mod_zero;
    a.tau = c.tau;
    a.live = e.live;
    a.skew_n = e.skew_n;
    a.sum_dst = cell_sum(e.sum_dst, c);
    return a;
}

__device__ __forceinline__ WaveArgs cell_wave_args(const EnsArgs &e, const EnsCell &c)
{
    WaveArgs w;
    w.tiles_z = c.tiles_z;
    w.wtiles_y = c.wtiles_y;
    w.btiles_y = c.btiles_y;
    w.chunks = c.chunks;
    w.xc = e.xc;
    w.has_zero_order = c.has_zero_order;
    w.ev_cap = uint32_t(kWEv);
    w.tab_words = c.tab_words;
    w.off_p2 = c.off_p2;
    w.off_rates = c.off_rates;
    w.off_ridx = c.off_ridx;
    w.off_stoich = c.off_stoich;
    w.off_rmask = c.off_rmask;
    w.off_bound = c.off_bound;
    w.pair_n1 = c.pair_n1;
    w.tab = e.tab + c.tab_off;
    w.regp = e.regp + c.regp_off;
    w.PY = c.PY;
    w.PZ4 = c.PZ4;
    w.nitems = 0u;
    w.queue = nullptr;
    w.items = nullptr;
    w.sum_src = c.sum_planes != 0u ? e.sum_src + c.sum_off : nullptr;
    return w;
}

// Per step, one lane per cell: the cell's itime from the device step counter, then its well-mixed phase.
template <int S>
__global__ __launch_bounds__(64) void ens_prologue(EnsArgs e, const CmeArgs *cme, int32_t *cme_counts)
{
    const uint32_t c = blockIdx.x * 64u + threadIdx.x;
    if (c >= e.n_cells)
        return;
    const EnsCell &cell = e.cells[c];
    const uint32_t it = splitmix_low32((uint64_t(cell.itime_seed) << 32) + e.queue[1]);
    e.itime[c] = it;
    if (cell.cme >= 0)
        cme_cell<S>(cme[cell.cme], cme_counts + cell.cme_counts_off, e.src + cell.lat_off, cell_sum(e.sum_src, cell),
                    cell.cme_seed, it);
}

// The lattice step of every cell: persistent blocks take (cell, group, x-chunk) items from the queue
// (`batch` consecutive items per access) or, with batch 0, in a fixed round-robin. A group is
// kWavesPerBlock wave tiles stacked in y; the block stages the cell's tables in LDS and each wave marches
// its tile.
template <int S>
__global__ __launch_bounds__(kWaveBlock) __attribute__((amdgpu_waves_per_eu(3))) void rdme_step_ensemble(EnsArgs e)
{
    __shared__ WaveLds<S> L[kWavesPerBlock];
    __shared__ uint32_t next;
    extern __shared__ uint32_t tab_lds[];
    TileCounts tc;
    const uint32_t wv = threadIdx.x / 64u;
    const uint32_t per = e.batch ? e.batch : 1u;
    uint32_t rr = blockIdx.x;  // round-robin position when there is no queue
    for (;;) {
        if (e.batch) {
            if (threadIdx.x == 0)
                next = atomicAdd(e.queue, e.batch);
            __syncthreads();
        } else if (threadIdx.x == 0) {
            next = rr;
        }
        if (!e.batch)
            __syncthreads();
        const uint32_t first = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();  // every lane has read `next` before it is written again
        if (first >= e.total_items)
            break;
        rr += gridDim.x;
        const uint32_t last = min(first + per, e.total_items);
        for (uint32_t item = first; item < last; ++item) {
            const uint32_t lo = __builtin_amdgcn_readfirstlane(e.item_cell[2u * item]);
            const uint32_t local = __builtin_amdgcn_readfirstlane(e.item_cell[2u * item + 1u]);
            const EnsCell &c = e.cells[lo];
            StepArgs a = cell_step_args(e, c, e.itime[lo]);
            if (*e.skew_n == 0u)
                a.live = nullptr;  // no skewed voxel: the live-bit loads are skipped
            const WaveArgs w = cell_wave_args(e, c);
            // the group's waves take consecutive wave tiles of the cell, y fastest (a column of tiles, then the
            // next column), so that only the last group of an x-chunk may have waves without a tile
            const uint32_t chunk = local / c.btiles_y;
            const uint32_t t = (local % c.btiles_y) * kWavesPerBlock + wv;
            const uint32_t tz = t / c.wtiles_y, wy = t % c.wtiles_y;
            const int xb = int(chunk * e.xc), xe = min(int(c.R0), int((chunk + 1u) * e.xc));
            // a wave whose march has nothing to do stops here; a group of such waves skips the table staging
            bool busy = tz < c.tiles_z;
            if (busy && w.sum_src != nullptr && c.has_zero_order == 0u && xe - xb <= 60) {
                const MarchSummary ms = march_summary(a, w, tz, wy, xb, xe);
                busy = (ms.act | ms.old) != 0ull;
                if (!busy)
                    count_idle_fires(a, tz, wy, xb, xe, tc);
            }
            if (!__syncthreads_or(busy ? 1 : 0))
                continue;
            for (uint32_t i = threadIdx.x; i < w.tab_words; i += kWaveBlock)
                tab_lds[i] = w.tab[i];
            __syncthreads();
            const uint8_t *tbase = reinterpret_cast<const uint8_t *>(tab_lds);
            const Tab tb{reinterpret_cast<const float *>(tbase + w.off_p2),
                         reinterpret_cast<const float *>(tbase + w.off_rates),
                         reinterpret_cast<const uint32_t *>(tbase + w.off_ridx),
                         reinterpret_cast<const int8_t *>(tbase + w.off_stoich),
                         reinterpret_cast<const uint64_t *>(tbase + w.off_rmask),
                         reinterpret_cast<const float *>(tbase + w.off_bound),
                         w.pair_n1 != 0u ? reinterpret_cast<const float *>(tbase + w.off_bound) + pair_offset(uint32_t(S))
                                         : nullptr};
            if (busy)
                wave_march<S>(L[wv], tb, a, w, tz, wy, xb, xe, tc);
            __syncthreads();  // the tables stay until every wave of the group is done
        }
    }
    flush_counts(e.counters, tc);
}

// The cells of an ensemble as the replay sees them: a voxel index among all cells' voxels -> its cell.
struct EnsCells {
    EnsArgs e;
    __device__ uint32_t find(uint32_t v) const
    {
        uint32_t lo = 0, hi = e.n_cells;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (e.cells[mid].vox_off <= v)
                lo = mid;
            else
                hi = mid;
        }
        return lo;
    }
    __device__ StepArgs args(uint32_t c) const { return cell_step_args(e, e.cells[c], e.itime[c]); }
};

// After each step: the replay of every cell's noted and skewed voxels (tight_replay_kernel's work, clusters
// never spanning two cells), then the queue re-armed and the step counter advanced.
template <int S>
__global__ __launch_bounds__(kReplayThreads) void ens_epilogue(EnsArgs e, SkewTab k, ReplayBuf b)
{
    if (b.ctl[0] != 0u || k.ctl[0] != 0u)
        replay_step<S>(EnsCells{e}, k, b, e.counters);
    if (threadIdx.x == 0) {
        e.queue[0] = 0u;
        e.queue[1] += 1u;
    }
}
}  // namespace

// Upper bound of a voxel's total propensity sum_r k_r * prod(counts of r's reactants) over every voxel
// holding c particles, c = 0..S: per reaction the largest product of integer counts that sum to at most
// c (found by enumeration; at most three reactants), times k_r. The sum gets a relative margin of 1e-4,
// far above the rounding of the device's f32 products and sums (a few hundred terms at 2^-24 each), and is
// rounded up to f32. With a negative rate no bound holds and every entry is +inf.
static std::vector<float> propensity_bounds(const Model &m, uint32_t S)
{
    const uint32_t nr = m.num_reactions();
    std::vector<double> b(S + 1u, 0.0);
    bool valid = true;
    for (uint32_t r = 1; r <= nr; ++r) {
        const double k = m.rates()[r];
        if (!(k >= 0.0)) {
            valid = false;
            break;
        }
        int sp[3], mult[3], d = 0;
        for (int j = 0; j < 3; ++j) {
            const int32_t x = m.reactants()[size_t(r) * 3u + j];
            if (x <= 0)
                continue;
            int f = 0;
            while (f < d && sp[f] != x)
                ++f;
            if (f == d) {
                sp[d] = x;
                mult[d++] = 0;
            }
            ++mult[f];
        }
        for (uint32_t c = 0; c <= S; ++c) {
            double best = 0.0;
            if (d == 0) {
                best = 1.0;  // zero-order: C[v][0] = 1
            } else {
                const int lim[3] = {int(c), d > 1 ? int(c) : 0, d > 2 ? int(c) : 0};
                for (int n0 = 0; n0 <= lim[0]; ++n0)
                    for (int n1 = 0; n1 <= lim[1] && n0 + n1 <= int(c); ++n1)
                        for (int n2 = 0; n2 <= lim[2] && n0 + n1 + n2 <= int(c); ++n2) {
                            const int n[3] = {n0, n1, n2};
                            double prod = 1.0;
                            for (int j = 0; j < d; ++j)
                                for (int e = 0; e < mult[j]; ++e)
                                    prod *= double(n[j]);
                            best = std::max(best, prod);
                        }
            }
            b[c] += k * best;
        }
    }
    std::vector<float> out(S + 1u, INFINITY);
    if (!valid)
        return out;
    for (uint32_t c = 0; c <= S; ++c) {
        const double v = b[c] * (1.0 + 1e-4);
        float f = float(v);
        if (double(f) < v)
            f = std::nextafter(f, INFINITY);
        out[c] = f;
    }
    return out;
}

// Pair table of the reaction reject test: entry [a][b] (a, b species, 0 for no particle) bounds from above the total
// propensity sum_r k_r * C[r0] * C[r1] * C[r2] (cme.wgsl:42-44, C[0] = 1) of a voxel holding exactly the particles a
// and b, rounded up. For models whose reactions have at most two reactants every term of that sum is a sum over
// single particles and pairs of particles (k * n^2 of a homodimer row counts each particle once alone and each
// pair twice), so for a voxel of c >= 2 particles the sum of the entries over its c(c-1)/2 pairs is at least its
// total. Empty (no table) when some reaction has three reactants, a rate is negative or not finite, or the table
// would pass 4 KB.
static std::vector<float> pair_bounds(const Model &m)
{
    const uint32_t ns = m.num_species(), nr = m.num_reactions(), n1 = ns + 1u;
    if (size_t(n1) * n1 * sizeof(float) > 4096u)
        return {};
    for (uint32_t r = 1; r <= nr; ++r) {
        const double k = m.rates()[r];
        if (!(k >= 0.0) || !std::isfinite(k))
            return {};
        int order = 0;
        for (int j = 0; j < 3; ++j)
            order += m.reactants()[size_t(r) * 3u + j] > 0 ? 1 : 0;
        if (order > 2)
            return {};
    }
    std::vector<float> out(size_t(n1) * n1);
    for (uint32_t a = 0; a < n1; ++a)
        for (uint32_t b = 0; b < n1; ++b) {
            double cnt[256] = {0.0};
            cnt[0] = 1.0;
            if (a > 0u)
                cnt[a] += 1.0;
            if (b > 0u)
                cnt[b] += 1.0;
            double tot = 0.0;
            for (uint32_t r = 1; r <= nr; ++r) {
                double prod = m.rates()[r];
                for (int j = 0; j < 3; ++j) {
                    const int32_t sp = m.reactants()[size_t(r) * 3u + j];
                    prod *= cnt[sp > 0 ? sp : 0];
                }
                tot += prod;
            }
            const double v = tot * (1.0 + 1e-4);
            float f = float(v);
            if (double(f) < v)
                f = std::nextafter(f, INFINITY);
            out[size_t(a) * n1 + b] = f;
        }
    return out;
}

// Host form of the occupancy summary (SumRef) of one compacted lattice: `out` (zeroed, tiles x R0 bytes) gets a 1
// for every wave tile and plane holding a particle.
static void summary_bytes(const uint8_t *lat, const uint32_t res[3], uint32_t S, uint32_t tiles_z, uint8_t *out)
{
    for (uint32_t x = 0; x < res[0]; ++x)
        for (uint32_t y = 0; y < res[1]; ++y) {
            const uint8_t *row = lat + (size_t(x) * res[1] + y) * res[2] * S;
            uint8_t *base = out + size_t(y / kSumTY) * tiles_z * res[0] + x;
            for (uint32_t z = 0; z < res[2]; ++z)
                if (row[size_t(z) * S] != 0u)
                    base[size_t(z / kSumTZ) * res[0]] = 1u;
        }
}

// Occupancy summaries are kept when at least a fifth of the (wave tile, plane) entries start empty: below that
// the tests cost more than the planes they skip.
static bool summary_pays(const std::vector<uint8_t> &bytes)
{
    size_t empty = 0;
    for (uint8_t b : bytes)
        empty += b == 0u;
    return !bytes.empty() && empty * 5u >= bytes.size();
}

// Checks that the model fits the device encodings (one-byte labels and slots, int8 stoichiometry).
static void check_device_model(const Model &m)
{
    if (m.num_regions() > 255)
        throw Error(CS_ERR_UNSUPPORTED, "at most 255 regions fit the one-byte label");
    if (m.num_reactions() > 255)
        throw Error(CS_ERR_UNSUPPORTED, "at most 255 reactions");
    if (m.n_voxels() * m.slots() >= (uint64_t(1) << 32))
        throw Error(CS_ERR_UNSUPPORTED, "lattice too large for u32 slot indices");
    for (float k : m.rates())
        if (!std::isfinite(k))
            throw Error(CS_ERR_UNSUPPORTED, "reaction rates must be finite");
    for (int32_t st : m.stoich())
        if (st < -127 || st > 127)
            throw Error(CS_ERR_UNSUPPORTED, "stoichiometry entries must fit int8");
}

static size_t align8(size_t n) { return (n + 7u) & ~size_t(7); }

// The lattice-step tables of a model in one blob (8-byte aligned sections) that the tiled kernel copies
// into LDS as is: hop probabilities (D*tau)/(lambda*lambda) in f32 as rdme.wgsl:135 evaluates them,
// rates, reactant bytes, candidate masks, int8 stoichiometry and the propensity bounds.
struct TableBlob {
    std::vector<uint8_t> bytes;
    uint32_t off_p2 = 0, off_rates = 0, off_ridx = 0, off_rmask = 0, off_stoich = 0, off_bound = 0;
    uint32_t nrw = 0;
    uint32_t pair_n1 = 0;  // ns + 1 when the pair table follows the bounds
    bool has_zero_order = false;
};

static TableBlob make_table_blob(const Model &m)
{
    const LatticeParams &lp = m.params();
    const uint32_t ns = m.num_species(), nr = m.num_reactions(), n1 = ns + 1u;
    TableBlob b;
    b.nrw = (nr + 1u + 63u) / 64u;
    std::vector<float> p2(m.diffusion().size());
    const float den = lp.lambda * lp.lambda;
    for (size_t i = 0; i < p2.size(); ++i) {
        const float num = m.diffusion()[i] * lp.tau;
        p2[i] = num / den;
    }
    std::vector<uint8_t> ridx(size_t(nr + 1u) * 4u, 0);
    std::vector<int8_t> st8(size_t(nr + 1u) * n1, 0);
    std::vector<uint64_t> rmask(size_t(n1) * b.nrw, 0ull);
    for (uint32_t r = 0; r <= nr; ++r) {
        bool any = false;
        for (int k = 0; k < 3; ++k) {
            const int32_t sp = m.reactants()[size_t(r) * 3u + k];
            ridx[size_t(r) * 4u + k] = uint8_t(sp);
            if (sp > 0) {
                any = true;
                rmask[size_t(sp) * b.nrw + r / 64u] |= 1ull << (r % 64u);
            }
        }
        if (!any && r > 0)
            rmask[r / 64u] |= 1ull << (r % 64u);  // zero-order: a candidate in every voxel
        for (uint32_t k = 0; k < n1; ++k)
            st8[size_t(r) * n1 + k] = int8_t(m.stoich()[size_t(r) * n1 + k]);
    }
    for (uint32_t w = 0; w < b.nrw; ++w)
        b.has_zero_order = b.has_zero_order || rmask[w] != 0ull;
    std::vector<float> bound = propensity_bounds(m, m.slots());
    const std::vector<float> pair = pair_bounds(m);
    if (!pair.empty()) {
        b.pair_n1 = n1;
        bound.resize(pair_offset(m.slots()), 0.0f);
        bound.insert(bound.end(), pair.begin(), pair.end());
    }
    b.off_p2 = 0;
    b.off_rates = uint32_t(align8(b.off_p2 + p2.size() * sizeof(float)));
    b.off_ridx = uint32_t(align8(b.off_rates + m.rates().size() * sizeof(float)));
    b.off_rmask = uint32_t(align8(b.off_ridx + ridx.size()));
    b.off_stoich = uint32_t(align8(b.off_rmask + rmask.size() * sizeof(uint64_t)));
    b.off_bound = uint32_t(align8(b.off_stoich + st8.size()));
    b.bytes.assign(align8(b.off_bound + bound.size() * sizeof(float)), 0);
    std::memcpy(b.bytes.data() + b.off_p2, p2.data(), p2.size() * sizeof(float));
    std::memcpy(b.bytes.data() + b.off_rates, m.rates().data(), m.rates().size() * sizeof(float));
    std::memcpy(b.bytes.data() + b.off_ridx, ridx.data(), ridx.size());
    std::memcpy(b.bytes.data() + b.off_rmask, rmask.data(), rmask.size() * sizeof(uint64_t));
    std::memcpy(b.bytes.data() + b.off_stoich, st8.data(), st8.size());
    std::memcpy(b.bytes.data() + b.off_bound, bound.data(), bound.size() * sizeof(float));
    return b;
}

// The well-mixed tables of a model in one blob, plus the voxel lists of the regions that receive
// lattice products (offsets and lengths indexed by region).
struct CmeBlob {
    std::vector<uint8_t> bytes;
    uint32_t off_reac = 0, off_delta = 0, off_emit = 0, off_region = 0, off_rates = 0;
    uint32_t off_lists = 0, off_loff = 0, off_llen = 0;
};

static CmeBlob make_cme_blob(const Model &m)
{
    CmeBlob b;
    const uint32_t ncr = m.num_cme_reactions(), nreg = m.num_regions();
    std::vector<uint32_t> lists, loff(nreg, 0u), llen(nreg, 0u);
    for (uint32_t r = 0; r < ncr; ++r) {
        const int32_t rg = m.cme_emit_region()[r];
        if (rg < 0 || llen[uint32_t(rg)] > 0u)
            continue;
        const std::vector<uint32_t> &vox = m.region_voxels(rg);
        loff[uint32_t(rg)] = uint32_t(lists.size());
        llen[uint32_t(rg)] = uint32_t(vox.size());
        lists.insert(lists.end(), vox.begin(), vox.end());
    }
    auto put = [&](const void *src, size_t bytes) {
        const uint32_t off = uint32_t(align8(b.bytes.size()));
        b.bytes.resize(off + bytes);
        if (bytes)
            std::memcpy(b.bytes.data() + off, src, bytes);
        return off;
    };
    b.off_reac = put(m.cme_reactants().data(), m.cme_reactants().size() * sizeof(int32_t));
    b.off_delta = put(m.cme_delta().data(), m.cme_delta().size() * sizeof(int32_t));
    b.off_emit = put(m.cme_emit().data(), m.cme_emit().size());
    b.off_region = put(m.cme_emit_region().data(), m.cme_emit_region().size() * sizeof(int32_t));
    b.off_rates = put(m.cme_rates().data(), m.cme_rates().size() * sizeof(float));
    b.off_lists = put(lists.data(), lists.size() * sizeof(uint32_t));
    b.off_loff = put(loff.data(), loff.size() * sizeof(uint32_t));
    b.off_llen = put(llen.data(), llen.size() * sizeof(uint32_t));
    b.bytes.resize(align8(b.bytes.size()));
    return b;
}

// Device view of a CME blob at `base`.
static CmeArgs cme_args_at(const uint8_t *base, const CmeBlob &b, const Model &m, unsigned long long *counters,
                           unsigned long long *audit, bool audit_on)
{
    CmeArgs c;
    c.nc = m.num_cme_species();
    c.ncr = m.num_cme_reactions();
    c.reactants = reinterpret_cast<const int32_t *>(base + b.off_reac);
    c.delta = reinterpret_cast<const int32_t *>(base + b.off_delta);
    c.emit = base + b.off_emit;
    c.emit_region = reinterpret_cast<const int32_t *>(base + b.off_region);
    c.rates = reinterpret_cast<const float *>(base + b.off_rates);
    c.lists = reinterpret_cast<const uint32_t *>(base + b.off_lists);
    c.list_off = reinterpret_cast<const uint32_t *>(base + b.off_loff);
    c.list_len = reinterpret_cast<const uint32_t *>(base + b.off_llen);
    c.counters = counters;
    c.audit = audit;
    c.audit_on = audit_on ? 1u : 0u;
    c.tau = m.params().tau;
    return c;
}

// Device table of skewed voxels (SkewTab) and the replay kernel's scratch (ReplayBuf) of one device state: a
// plain simulation, or every cell of an ensemble (voxel indices among all cells' voxels). The table starts with
// the reservoir voxels, rebuilt from the host copy whenever the lattice is replaced.
struct SkewStore {
    SkewTab k{};
    ReplayBuf b{};
    uint64_t nvox = 0;
    std::vector<uint32_t> keys0, idx0;  // the table as it starts: reservoir voxels only
    std::vector<SkewEntry> ent0;
    std::vector<uint32_t> live0;        // their bits

    static uint32_t pow2_at_least(uint64_t n)
    {
        uint64_t p = 1;
        while (p < n)
            p <<= 1;
        return uint32_t(p);
    }

    // Host-side insert of a reservoir voxel's entry, in the probing order of skew_insert.
    void add_static(uint32_t v, uint32_t R, uint32_t K)
    {
        uint32_t h = skew_hash(v) & k.mask;
        while (keys0[h] != kSkewFree && keys0[h] != v)
            h = (h + 1u) & k.mask;
        if (keys0[h] == kSkewFree) {
            keys0[h] = v;
            idx0[h] = uint32_t(ent0.size());
            SkewEntry e{v, 0, 0u, kSkewStatic | K, {0, 0, 0, 0}};
            ent0.push_back(e);
        }
        SkewEntry &e = ent0[idx0[h]];
        e.kf = kSkewStatic | K;
        if (R != 0u) {  // the reservoir unit: C[v][R] = 1 with no slot (lattice.rs:297)
            int32_t dc[kSkewSp];
            std::memcpy(dc, e.dc, sizeof(dc));
            uint32_t sp = e.sp;
            bool placed = false;
            for (int q = 0; q < kSkewSp && !placed; ++q)
                if (((sp >> (8 * q)) & 0xFFu) == R || ((sp >> (8 * q)) & 0xFFu) == 0u) {
                    sp |= R << (8 * q);
                    dc[q] += 1;
                    placed = true;
                }
            e.sp = sp;
            std::memcpy(e.dc, dc, sizeof(dc));
        }
        live0[v >> 5] |= 1u << (v & 31u);
    }

    // Sizes the table for every voxel of a state of up to 2^20 voxels (a voxel keeps its entry once it has been
    // skewed), 2^20 entries above that (at least twice the reservoir voxels), and the replay scratch: up to tcap
    // noted voxels per step (every voxel twice, at most 2^18), at most `max_clusters` clusters (any two are
    // kClusterGap voxels apart along some axis, so no box of kClusterGap^3 voxels holds the first voxels of two),
    // and their F sets (a cluster's F is disjoint from every other's, so all of them hold at most every voxel
    // once, plus kEscMax voxels each for the voxels they may add).
    void alloc(uint64_t nvox_, uint32_t S, uint64_t n_static, uint32_t *fix_ctl, uint64_t max_clusters)
    {
        nvox = nvox_;
        uint64_t want = std::max<uint64_t>(2 * n_static + 8192, std::min<uint64_t>(nvox, uint64_t(1) << 20));
        want = std::min<uint64_t>(want, nvox + 64) + kSkewMargin;
        k.ecap = uint32_t(want);
        const uint32_t cap = pow2_at_least(2 * want);
        k.mask = cap - 1u;
        keys0.assign(cap, kSkewFree);
        idx0.assign(cap, 0u);
        ent0.clear();
        live0.assign((nvox + 31u) / 32u, 0u);
        hip_check(hipMalloc(&k.keys, size_t(cap) * sizeof(uint32_t)), "hipMalloc skew keys");
        hip_check(hipMalloc(&k.idx, size_t(cap) * sizeof(uint32_t)), "hipMalloc skew index");
        hip_check(hipMalloc(&k.ent, size_t(k.ecap) * sizeof(SkewEntry)), "hipMalloc skew entries");
        hip_check(hipMalloc(&k.ctl, 4 * sizeof(uint32_t)), "hipMalloc skew control");
        hip_check(hipMalloc(&k.live, std::max<size_t>(live0.size(), 1) * sizeof(uint32_t)), "hipMalloc skew bits");
        b.ctl = fix_ctl;
        b.tcap = pow2_at_least(std::max<uint64_t>(std::min<uint64_t>(2 * nvox + 64, uint64_t(1) << 18), 1024));
        b.ccap = uint32_t(std::min<uint64_t>(max_clusters + 1, b.tcap));
        b.fcap = uint32_t(std::min<uint64_t>(12ull * b.tcap, nvox) + (kEscMax + 8) * uint64_t(b.ccap));
        b.nvox = uint32_t(nvox);
        b.lds_sort_max = kSortLds;
        if (const char *x = std::getenv("CS_REPLAY_LDS_SORT"))  // test hook: longer lists take the bitmap sort
            b.lds_sort_max = std::min<uint32_t>(uint32_t(std::strtoul(x, nullptr, 10)), kSortLds);
        b.wave_max = kWaveReplayT;
        b.block_max = kBlockReplayT;
        if (const char *x = std::getenv("CS_REPLAY_FORMS")) {  // test hook: "block" (no wave form), "lane" (one lane)
            const std::string f(x);
            if (f == "block" || f == "lane")
                b.wave_max = 0u;
            if (f == "lane")
                b.block_max = 0u;
        }
        const size_t lo_words = (nvox + 31u) / 32u, hi_words = ((nvox + 1023u) / 1024u + 31u) / 32u;
        hip_check(hipMalloc(&b.T, size_t(b.tcap) * sizeof(uint32_t)), "hipMalloc replay list");
        hip_check(hipMalloc(&b.xclus, size_t(b.ccap) * sizeof(uint4)), "hipMalloc replay runs");
        hip_check(hipMalloc(&b.clus, size_t(b.ccap) * sizeof(uint4)), "hipMalloc replay clusters");
        hip_check(hipMalloc(&b.mark_lo, (lo_words + 32u) * sizeof(uint32_t)), "hipMalloc replay marks");
        hip_check(hipMalloc(&b.mark_hi, hi_words * sizeof(uint32_t)), "hipMalloc replay marks");
        hip_check(hipMemset(b.mark_lo, 0, (lo_words + 32u) * sizeof(uint32_t)), "clear replay marks");
        hip_check(hipMemset(b.mark_hi, 0, hi_words * sizeof(uint32_t)), "clear replay marks");
        hip_check(hipMalloc(&b.F, size_t(b.fcap) * sizeof(uint32_t)), "hipMalloc replay F");
        hip_check(hipMalloc(&b.FL, size_t(b.fcap) * S), "hipMalloc replay slots");
        hip_check(hipMalloc(&b.FO, size_t(b.fcap) * sizeof(uint32_t)), "hipMalloc replay occupancy");
        hip_check(hipMalloc(&b.FW, size_t(b.fcap) * sizeof(uint32_t)), "hipMalloc replay walks");
        hip_check(hipMalloc(&b.FS, size_t(b.fcap) * sizeof(uint32_t)), "hipMalloc replay offsets");
        hip_check(hipMalloc(&b.FD, size_t(b.fcap) * kSkewSp * sizeof(int32_t)), "hipMalloc replay offsets");
    }

    // Resets the device table to its starting entries.
    void upload(hipStream_t st)
    {
        const uint32_t ctl[4] = {uint32_t(ent0.size()), 0u, 0u, 0u};
        hip_check(hipMemcpyAsync(k.keys, keys0.data(), keys0.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st),
                  "upload skew keys");
        hip_check(hipMemcpyAsync(k.idx, idx0.data(), idx0.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st),
                  "upload skew index");
        if (!ent0.empty())
            hip_check(hipMemcpyAsync(k.ent, ent0.data(), ent0.size() * sizeof(SkewEntry), hipMemcpyHostToDevice, st),
                      "upload skew entries");
        hip_check(hipMemcpyAsync(k.ctl, ctl, sizeof(ctl), hipMemcpyHostToDevice, st), "upload skew control");
        if (!live0.empty())
            hip_check(hipMemcpyAsync(k.live, live0.data(), live0.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st),
                      "upload skew bits");
        hip_check(hipStreamSynchronize(st), "upload skew table");  // ctl lives on this stack frame
    }

    // The live entries (voxel indices among all voxels of the state), in insertion order.
    void read(hipStream_t st, std::vector<SkewVoxel> &out) const
    {
        uint32_t ctl[4];
        hip_check(hipMemcpyAsync(ctl, k.ctl, sizeof(ctl), hipMemcpyDeviceToHost, st), "read skew control");
        hip_check(hipStreamSynchronize(st), "read skew control");
        std::vector<SkewEntry> e(ctl[0]);
        if (!e.empty())
            hip_check(hipMemcpyAsync(e.data(), k.ent, e.size() * sizeof(SkewEntry), hipMemcpyDeviceToHost, st),
                      "read skew entries");
        hip_check(hipStreamSynchronize(st), "read skew entries");
        out.clear();
        for (const SkewEntry &x : e) {
            SkewVoxel s;
            s.v = x.v;
            s.off = x.off;
            s.K = x.kf & 0xFFu;
            for (int q = 0; q < kSkewSp; ++q) {
                s.sp[q] = (x.sp >> (8 * q)) & 0xFFu;
                s.dc[q] = s.sp[q] ? x.dc[q] : 0;
            }
            if (x.off != 0 || x.sp != 0u || s.K != 0u || (x.kf & kSkewStatic))
                out.push_back(s);
        }
    }

    uint32_t refused(hipStream_t st) const
    {
        uint32_t ctl[4];
        hip_check(hipMemcpyAsync(ctl, k.ctl, sizeof(ctl), hipMemcpyDeviceToHost, st), "read skew control");
        hip_check(hipStreamSynchronize(st), "read skew control");
        return ctl[1];
    }

    size_t table_bytes() const { return size_t(k.mask + 1u) * 2u * sizeof(uint32_t); }

    void release()
    {
        void *bufs[] = {k.keys, k.idx, k.ent, k.ctl, k.live, b.T, b.xclus, b.clus, b.mark_lo, b.mark_hi,
                        b.F, b.FL, b.FO, b.FW, b.FS, b.FD};
        for (void *p : bufs)
            if (p)
                (void)hipFree(p);
        k = SkewTab{};
        b = ReplayBuf{};
    }
};

// Clusters a lattice can hold at once: one per box of kClusterGap^3 voxels.
static uint64_t cluster_bound(const uint32_t res[3])
{
    uint64_t n = 1;
    for (int i = 0; i < 3; ++i)
        n *= (res[i] + kClusterGap - 1u) / kClusterGap;
    return n;
}

// The reservoir voxels of a model, as static entries with voxel indices offset by `base`: sites of a reservoir
// species R (one unit of C[v][R] in no slot) and voxels v whose reservoirs[v*S] names a species K (cme.wgsl:80
// indexes the reservoir buffer with the lattice index; an index past the end reads 0).
static void add_reservoir_entries(SkewStore &sk, const Model &m, uint32_t base)
{
    if (!m.has_reservoir())
        return;
    const std::vector<uint32_t> &rs = m.reservoir();
    const uint64_t N = m.n_voxels(), S = m.slots();
    for (uint64_t v = 0; v < N; ++v) {
        const uint32_t R = rs[v];
        const uint32_t K = v * S < N ? rs[v * S] : 0u;
        if (R == 0u && K == 0u)
            continue;
        if (R > 255u || K > 255u)
            throw Error(CS_ERR_UNSUPPORTED, "reservoir species index beyond one byte");
        sk.add_static(base + uint32_t(v), R, K);
    }
}

static uint64_t count_reservoir_voxels(const Model &m)
{
    if (!m.has_reservoir())
        return 0;
    const std::vector<uint32_t> &rs = m.reservoir();
    const uint64_t N = m.n_voxels(), S = m.slots();
    uint64_t n = 0;
    for (uint64_t v = 0; v < N; ++v)
        n += rs[v] != 0u || (v * S < N && rs[v * S] != 0u);
    return n;
}

// Each device's copy of `scratch`, looked up once per device with that device current.
// Simulations alive on each HIP device. The work planner cuts a small lattice into shorter items only when its
// simulation has the device to itself: beside others (an application stepping many cells, each on its own stream)
// the idle blocks are theirs, and the shorter items' extra side planes only add work.
static std::atomic<int> &live_sims(int dev)
{
    static std::atomic<int> n[256];
    return n[dev & 255];
}

static int32_t *noalloc_bins(int dev)
{
    static std::mutex mu;
    static void *addr[256] = {};
    if (dev < 0 || dev >= 256)
        throw Error(CS_ERR_ARG, "HIP device index out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        int cur = -1;
        hip_check(hipGetDevice(&cur), "hipGetDevice");
        if (cur != dev)
            throw Error(CS_ERR_STATE, "the fixed histogram is looked up with its device current");
        void *p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        addr[dev] = p;
    }
    return static_cast<int32_t *>(addr[dev]);
}

Device::Device(const Model &m, int hip_device) : dev_(hip_device)
{
    const LatticeParams &lp = m.params();
    for (int i = 0; i < 3; ++i)
        res_[i] = lp.res[i];
    N_ = m.n_voxels();
    S_ = m.slots();
    ns_ = m.num_species();
    nr_ = m.num_reactions();
    nreg_ = m.num_regions();
    tau_ = lp.tau;
    check_device_model(m);

    int ndev = 0;
    hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (hip_device < 0 || hip_device >= ndev)
        throw Error(CS_ERR_ARG, "no such HIP device");
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st;
    hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    stream_ = st;
    (void)noalloc_bins(dev_);  // this device's fixed histogram, for region_counts_noalloc

    const uint32_t n1 = ns_ + 1u;
    const TableBlob blob = make_table_blob(m);
    nrw_ = blob.nrw;
    pair_n1_ = blob.pair_n1;
    has_zero_order_ = blob.has_zero_order;
    std::vector<uint8_t> reg8(N_);
    for (uint64_t v = 0; v < N_; ++v)
        reg8[v] = uint8_t(m.regions()[v]);
    if (const char *k = std::getenv("CS_STEP_KERNEL")) {
        const std::string kk(k);
        simple_ = kk == "simple";
        tiled_ = kk == "tiled";
    }
    if (const char *x = std::getenv("CS_WAVE_EVENT_CAP")) {  // test hook: a smaller arrival-event list
        const long v = std::strtol(x, nullptr, 10);
        if (v >= 0 && v <= kWEv)
            ev_cap_ = uint32_t(v);
    }
    if (const char *x = std::getenv("CS_WAVE_PLAN"))  // tuning hook: "cost" (the cost-ordered planner) or "balanced"
        balanced_plan_ = std::string(x) != "cost";
    if (const char *x = std::getenv("CS_WAVE_ITEMS")) {  // tuning hook: work items per resident block
        const long v = std::strtol(x, nullptr, 10);
        if (v >= 1 && v <= 256)
            items_per_block_ = uint32_t(v);
    }
    if (const char *x = std::getenv("CS_WAVE_MIN_PLANES")) {  // tuning hook: shortest x-range of an item
        const long v = std::strtol(x, nullptr, 10);
        if (v >= 1 && v <= 4096) {
            xc_min_ = uint32_t(v);
            min_planes_set_ = true;
        }
    }
    if (const char *x = std::getenv("CS_TILE_PLANES")) {  // fixed chunk length (tests, tuning)
        const long v = std::strtol(x, nullptr, 10);
        if (v >= 1 && v <= 4096) {
            xc_ = xc_min_ = tiled_xc_ = uint32_t(v);
            balanced_plan_ = false;  // fixed chunks: the cost-ordered planner cuts every column into them
        }
    }

    hip_check(hipMalloc(&lat_[0], N_ * S_), "hipMalloc lattice");
    hip_check(hipMalloc(&lat_[1], N_ * S_), "hipMalloc lattice");
    hip_check(hipMalloc(&reg_, N_), "hipMalloc regions");
    // Labels padded to whole wave tiles plus two voxels on every side, 0xFF outside the lattice: the wave
    // kernel reads its label window as aligned words without bounds tests, and 0xFF marks a missing neighbour.
    wtiles_y_ = (res_[1] + kWY - 1) / kWY;
    wtiles_z_ = (res_[2] + kWZ - 1) / kWZ;
    pad_y_ = kWY * wtiles_y_ + 4u;
    pad_z_ = kWZ * wtiles_z_ + 4u;
    std::vector<uint8_t> regp(size_t(res_[0] + 4u) * pad_y_ * pad_z_, 0xFFu);
    for (uint32_t x = 0; x < res_[0]; ++x)
        for (uint32_t y = 0; y < res_[1]; ++y)
            std::memcpy(&regp[(size_t(x + 2u) * pad_y_ + (y + 2u)) * pad_z_ + 2u],
                        &reg8[(size_t(x) * res_[1] + y) * res_[2]], res_[2]);
    hip_check(hipMalloc(&regp_, regp.size()), "hipMalloc padded regions");

    off_p2_ = blob.off_p2;
    off_rates_ = blob.off_rates;
    off_ridx_ = blob.off_ridx;
    off_rmask_ = blob.off_rmask;
    off_stoich_ = blob.off_stoich;
    off_bound_ = blob.off_bound;
    const size_t blob_bytes = blob.bytes.size();
    tab_words_ = uint32_t(blob_bytes / 4u);
    hip_check(hipMalloc(&tab_, blob_bytes), "hipMalloc tables");
    uint8_t *tb = reinterpret_cast<uint8_t *>(tab_);
    p2_ = reinterpret_cast<float *>(tb + off_p2_);
    rates_ = reinterpret_cast<float *>(tb + off_rates_);
    ridx_ = tb + off_ridx_;
    rmask_ = reinterpret_cast<uint64_t *>(tb + off_rmask_);
    stoich_ = reinterpret_cast<int8_t *>(tb + off_stoich_);
    bound_ = reinterpret_cast<float *>(tb + off_bound_);
    if (nreg_ >= 255u)
        tiled_ = true;   // the wave kernel's padded labels use 0xFF for "outside the lattice"
    if (res_[1] > 4096u || res_[2] > 4096u || res_[0] > 65535u)
        tiled_ = true;   // the wave kernel forms y * R2 with 24-bit multiplies, packs x into 16 bits
    // tables too large to stage in LDS next to the plane records: the wave kernel stages the hop probabilities and
    // bounds and reads the reaction tables from global memory, or reads everything there when even those are too
    // large (the tiled kernel, which stages all of them, gives way to the one-lane-per-voxel kernel)
    const size_t split_bytes = off_rates_ + size_t(bound_words(S_, blob.pair_n1)) * sizeof(float);
    const char *gt = std::getenv("CS_WAVE_GLOBAL_TABLES");  // test hook: "1" the split form if it fits, "2" global
    if (blob_bytes > kTabLdsMax || gt != nullptr) {
        if (tiled_ && blob_bytes > kTabLdsMax)
            simple_ = true;
        const bool all = gt != nullptr && std::string(gt) == "2";
        tab_mode_ = !all && split_bytes <= kTabLdsMax ? kTabSplit : kTabGlobal;
    }
    {
        // resident blocks of the persistent wave kernel: every CU filled to its occupancy
        int cus = 0, per_cu = 0;
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_), "CU count");
        const size_t shm = wave_shm_bytes();
        hipError_t e = hipSuccess;
        switch (S_ * 4 + tab_mode_) {
#define CS_WAVE_OCC(s, tm) \
    case s * 4 + tm: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rdme_step_wave<s, tm>, kWaveBlock, shm); break;
        CS_WAVE_OCC(4, kTabLds) CS_WAVE_OCC(4, kTabSplit) CS_WAVE_OCC(4, kTabGlobal)
        CS_WAVE_OCC(8, kTabLds) CS_WAVE_OCC(8, kTabSplit) CS_WAVE_OCC(8, kTabGlobal)
        CS_WAVE_OCC(16, kTabLds) CS_WAVE_OCC(16, kTabSplit) CS_WAVE_OCC(16, kTabGlobal)
#undef CS_WAVE_OCC
        default: throw Error(CS_ERR_UNSUPPORTED, "slots per voxel must be 4, 8 or 16");
        }
        hip_check(e, "occupancy of the wave kernel");
        wave_blocks_ = uint32_t(std::max(cus, 1) * std::max(per_cu, 1));
        if (const char *x = std::getenv("CS_WAVE_BLOCKS")) {  // tuning hook: resident blocks
            const long v = std::strtol(x, nullptr, 10);
            if (v >= 1 && v <= 65536)
                wave_blocks_ = uint32_t(v);
        }
    }
    hip_check(hipMalloc(&counters_, kNumCounters * sizeof(unsigned long long)), "hipMalloc counters");
    hip_check(hipMalloc(&hist_, size_t(nreg_) * n1 * sizeof(int32_t)), "hipMalloc hist");
    hip_check(hipMalloc(&fix_ctl_, kFixCtl * sizeof(uint32_t)), "hipMalloc fix_ctl");
    hip_check(hipMemsetAsync(fix_ctl_, 0, kFixCtl * sizeof(uint32_t), st), "clear fix_ctl");
    // skewed voxels: the reservoir voxels to start with (their table and bits uploaded below)
    skew_ = new SkewStore();
    skew_->alloc(N_, S_, count_reservoir_voxels(m), fix_ctl_, cluster_bound(res_));
    add_reservoir_entries(*skew_, m, 0u);
    reg_host_ = reg8;

    // Well-mixed (CME) pool and its tables in one blob (8-byte aligned sections), plus the voxel lists of
    // the regions that receive lattice products.
    nc_ = m.num_cme_species();
    ncr_ = m.num_cme_reactions();
    cme_seed_ = m.cme_seed();
    has_cme_ = ncr_ > 0u;
    hip_check(hipMalloc(&cme_counts_, size_t(nc_ > 0u ? nc_ : 1u) * sizeof(int32_t)), "hipMalloc cme counts");
    hip_check(hipMalloc(&audit_, 3 * sizeof(unsigned long long)), "hipMalloc audit");
    hip_check(hipMemsetAsync(audit_, 0, 3 * sizeof(unsigned long long), st), "clear audit");
    if (nc_ > 0u)
        hip_check(hipMemcpyAsync(cme_counts_, m.cme_init().data(), nc_ * sizeof(int32_t), hipMemcpyHostToDevice, st),
                  "upload cme counts");
    CmeBlob cb;  // lives until the uploads below have completed (the stream is synchronised at the end)
    if (has_cme_) {
        cb = make_cme_blob(m);
        off_creac_ = cb.off_reac;
        off_cdelta_ = cb.off_delta;
        off_cemit_ = cb.off_emit;
        off_cregion_ = cb.off_region;
        off_crates_ = cb.off_rates;
        off_clists_ = cb.off_lists;
        off_cloff_ = cb.off_loff;
        off_cllen_ = cb.off_llen;
        hip_check(hipMalloc(&cme_tab_, cb.bytes.size()), "hipMalloc cme tables");
        hip_check(hipMemcpyAsync(cme_tab_, cb.bytes.data(), cb.bytes.size(), hipMemcpyHostToDevice, st),
                  "upload cme tables");
    }

    std::vector<uint8_t> rvol;
    if (m.has_reservoir()) {  // the reservoir view's labels (final_stage.wgsl:63-66)
        rvol.resize(N_);
        for (uint64_t v = 0; v < N_; ++v)
            rvol[v] = uint8_t(m.reservoir()[v]);
        hip_check(hipMalloc(&res_vol_, N_), "hipMalloc reservoir volume");
        hip_check(hipMemcpyAsync(res_vol_, rvol.data(), N_, hipMemcpyHostToDevice, st), "upload reservoir volume");
    }

    hip_check(hipMemcpyAsync(lat_[0], m.lattice().data(), N_ * S_, hipMemcpyHostToDevice, st), "upload lattice");
    hip_check(hipMemsetAsync(lat_[1], 0, N_ * S_, st), "clear lattice");
    skew_->upload(st);
    // occupancy summaries of both buffers for the wave kernel (none with zero-order reactions, which can
    // fire in any voxel)
    summ_ = !simple_ && !tiled_ && !has_zero_order_;
    std::vector<uint8_t> sum0;
    if (summ_) {
        const size_t nb = size_t(wtiles_y_) * wtiles_z_ * res_[0];
        sum0.assign(nb, 0u);
        summary_bytes(m.lattice().data(), res_, S_, wtiles_z_, sum0.data());
        // they pay where tile planes are empty; a lattice with particles nearly everywhere keeps none
        summ_ = summary_pays(sum0);
        if (const char *x = std::getenv("CS_OCC_SUMMARY"))  // "on" / "off": keep them or not regardless
            summ_ = std::string(x) == "on" ? true : (std::string(x) == "off" ? false : summ_);
    }
    if (summ_) {
        const size_t nb = sum0.size();
        for (int b = 0; b < 2; ++b)
            hip_check(hipMalloc(&sum_[b], nb), "hipMalloc occupancy summary");
        hip_check(hipMemcpyAsync(sum_[0], sum0.data(), nb, hipMemcpyHostToDevice, st), "upload occupancy summary");
        hip_check(hipMemsetAsync(sum_[1], 0, nb, st), "clear occupancy summary");
    }
    plan_wave_items(m.lattice().data());
    hip_check(hipMemcpyAsync(reg_, reg8.data(), N_, hipMemcpyHostToDevice, st), "upload regions");
    hip_check(hipMemcpyAsync(regp_, regp.data(), regp.size(), hipMemcpyHostToDevice, st), "upload padded regions");
    hip_check(hipMemcpyAsync(tab_, blob.bytes.data(), blob_bytes, hipMemcpyHostToDevice, st), "upload tables");
    hip_check(hipMemsetAsync(counters_, 0, kNumCounters * sizeof(unsigned long long), st), "clear counters");
    hip_check(hipStreamSynchronize(st), "upload");
    cur_ = 0;
    live_sims(dev_)++;
    counted_ = true;
}

Device::~Device()
{
    if (counted_)
        live_sims(dev_)--;
    hipSetDevice(dev_);
    if (stream_)
        hipStreamSynchronize(static_cast<hipStream_t>(stream_));
    hipFree(lat_[0]);
    hipFree(lat_[1]);
    hipFree(reg_);
    hipFree(regp_);
    hipFree(items_);
    hipFree(tab_);
    hipFree(counters_);
    hipFree(hist_);
    hipFree(fix_ctl_);
    if (skew_) {
        skew_->release();
        delete skew_;
    }
    hipFree(cme_counts_);
    hipFree(cme_tab_);
    hipFree(audit_);
    hipFree(res_vol_);
    hipFree(view_);
    hipFree(sum_[0]);
    hipFree(sum_[1]);
    if (stream_)
        hipStreamDestroy(static_cast<hipStream_t>(stream_));
}

void Device::step(uint32_t itime)
{
    StepArgs a;
    a.src = lat_[cur_];
    a.dst = lat_[cur_ ^ 1];
    a.reg = reg_;
    a.p2 = p2_;
    a.rates = rates_;
    a.ridx = ridx_;
    a.stoich = stoich_;
    a.rmask = rmask_;
    a.bound = bound_;
    a.counters = counters_;
    a.fix_ctl = fix_ctl_;
    a.fix_list = skew_->b.T;
    a.max_tight = skew_->b.tcap;
    a.vox_base = 0;
    a.R0 = res_[0];
    a.R1 = res_[1];
    a.R2 = res_[2];
    a.R12 = res_[1] * res_[2];
    a.N = uint32_t(N_);
    a.ns = ns_;
    a.nr = nr_;
    a.nreg = nreg_;
    a.nrw = nrw_;
    a.itime = itime;
    a.mod_zero = mod_zero_;
    a.tau = tau_;
    a.live = skew_->k.live;
    a.skew_n = skew_->k.ctl;
    a.sum_dst = SumRef{summ_ ? sum_[cur_ ^ 1] : nullptr, res_[0], wtiles_z_, res_[1], res_[2]};
    const SumRef sum_src{summ_ ? sum_[cur_] : nullptr, res_[0], wtiles_z_, res_[1], res_[2]};
    hipStream_t st = static_cast<hipStream_t>(stream_);
    if (has_cme_) {
        const uint8_t *cb = static_cast<const uint8_t *>(cme_tab_);
        CmeArgs c;
        c.nc = nc_;
        c.ncr = ncr_;
        c.reactants = reinterpret_cast<const int32_t *>(cb + off_creac_);
        c.delta = reinterpret_cast<const int32_t *>(cb + off_cdelta_);
        c.emit = cb + off_cemit_;
        c.emit_region = reinterpret_cast<const int32_t *>(cb + off_cregion_);
        c.rates = reinterpret_cast<const float *>(cb + off_crates_);
        c.lists = reinterpret_cast<const uint32_t *>(cb + off_clists_);
        c.list_off = reinterpret_cast<const uint32_t *>(cb + off_cloff_);
        c.list_len = reinterpret_cast<const uint32_t *>(cb + off_cllen_);
        c.counters = counters_;
        c.audit = audit_;
        c.audit_on = audit_on_ ? 1u : 0u;
        c.tau = tau_;
        switch (S_) {
        case 4: hipLaunchKernelGGL(cme_kernel<4>, dim3(1), dim3(64), 0, st, c, cme_counts_, lat_[cur_], sum_src, cme_seed_, itime); break;
        case 8: hipLaunchKernelGGL(cme_kernel<8>, dim3(1), dim3(64), 0, st, c, cme_counts_, lat_[cur_], sum_src, cme_seed_, itime); break;
        default: hipLaunchKernelGGL(cme_kernel<16>, dim3(1), dim3(64), 0, st, c, cme_counts_, lat_[cur_], sum_src, cme_seed_, itime); break;
        }
        hip_check(hipGetLastError(), "cme launch");
    }
    if (simple_) {
        const uint32_t blocks = uint32_t((N_ + kBlock - 1) / kBlock);
        switch (S_) {
        case 4: hipLaunchKernelGGL(rdme_step_kernel<4>, dim3(blocks), dim3(kBlock), 0, st, a); break;
        case 8: hipLaunchKernelGGL(rdme_step_kernel<8>, dim3(blocks), dim3(kBlock), 0, st, a); break;
        case 16: hipLaunchKernelGGL(rdme_step_kernel<16>, dim3(blocks), dim3(kBlock), 0, st, a); break;
        default: throw Error(CS_ERR_UNSUPPORTED, "slots per voxel must be 4, 8 or 16");
        }
    } else if (!tiled_) {
        WaveArgs w;
        w.tiles_z = wtiles_z_;
        w.wtiles_y = wtiles_y_;
        w.btiles_y = (wtiles_y_ + kWavesPerBlock - 1) / kWavesPerBlock;
        w.xc = xc_;
        w.chunks = (res_[0] + xc_ - 1) / xc_;
        w.has_zero_order = has_zero_order_ ? 1u : 0u;
        w.ev_cap = ev_cap_;
        w.tab_words = tab_words_;
        w.off_p2 = off_p2_;
        w.off_rates = off_rates_;
        w.off_ridx = off_ridx_;
        w.off_stoich = off_stoich_;
        w.off_rmask = off_rmask_;
        w.off_bound = off_bound_;
        w.pair_n1 = pair_n1_;
        w.tab = static_cast<const uint32_t *>(tab_);
        w.regp = reinterpret_cast<const uint32_t *>(regp_);
        w.PY = pad_y_;
        w.PZ4 = pad_z_ / 4u;
        w.nitems = n_items_;
        w.queue = fix_ctl_ + kFixQueue;
        w.items = static_cast<const uint2 *>(items_);
        w.sum_src = sum_src.bytes;
        const uint32_t blocks = std::min(n_items_, wave_blocks_);
        const size_t shm = wave_shm_bytes();
        // with a timing pair set, the dispatch itself stamps the events at the kernel's start and end
        hipEvent_t e0 = static_cast<hipEvent_t>(kernel_ev_[0]), e1 = static_cast<hipEvent_t>(kernel_ev_[1]);
        const dim3 g(blocks), b(kWaveBlock);
        switch (S_ * 4 + tab_mode_) {
#define CS_WAVE_LAUNCH(s, tm) \
    case s * 4 + tm: hipExtLaunchKernelGGL(rdme_step_wave<s, tm>, g, b, shm, st, e0, e1, 0u, a, w); break;
        CS_WAVE_LAUNCH(4, kTabLds) CS_WAVE_LAUNCH(4, kTabSplit) CS_WAVE_LAUNCH(4, kTabGlobal)
        CS_WAVE_LAUNCH(8, kTabLds) CS_WAVE_LAUNCH(8, kTabSplit) CS_WAVE_LAUNCH(8, kTabGlobal)
        CS_WAVE_LAUNCH(16, kTabLds) CS_WAVE_LAUNCH(16, kTabSplit) CS_WAVE_LAUNCH(16, kTabGlobal)
#undef CS_WAVE_LAUNCH
        default: throw Error(CS_ERR_UNSUPPORTED, "slots per voxel must be 4, 8 or 16");
        }
    } else {
        TileArgs t;
        t.tiles_z = (res_[2] + kTZ - 1) / kTZ;
        t.tiles_y = (res_[1] + kTY - 1) / kTY;
        t.xc = tiled_xc_;
        t.chunks = (res_[0] + tiled_xc_ - 1) / tiled_xc_;
        t.has_zero_order = has_zero_order_ ? 1u : 0u;
        t.tab_words = tab_words_;
        t.off_p2 = off_p2_;
        t.off_rates = off_rates_;
        t.off_ridx = off_ridx_;
        t.off_stoich = off_stoich_;
        t.off_rmask = off_rmask_;
        t.off_bound = off_bound_;
        t.tab = static_cast<const uint32_t *>(tab_);
        const uint32_t blocks = t.tiles_z * t.tiles_y * t.chunks;
        const size_t shm = size_t(tab_words_) * 4u;
        switch (S_) {
        case 4: hipLaunchKernelGGL(rdme_step_tiled<4>, dim3(blocks), dim3(kBlock), shm, st, a, t); break;
        case 8: hipLaunchKernelGGL(rdme_step_tiled<8>, dim3(blocks), dim3(kBlock), shm, st, a, t); break;
        case 16: hipLaunchKernelGGL(rdme_step_tiled<16>, dim3(blocks), dim3(kBlock), shm, st, a, t); break;
        default: throw Error(CS_ERR_UNSUPPORTED, "slots per voxel must be 4, 8 or 16");
        }
    }
    hip_check(hipGetLastError(), "rdme step launch");
    const SkewTab k = skew_->k;
    const ReplayBuf rb = skew_->b;
    switch (S_) {
    case 4: hipLaunchKernelGGL(tight_replay_kernel<4>, dim3(1), dim3(kReplayThreads), 0, st, a, k, rb); break;
    case 8: hipLaunchKernelGGL(tight_replay_kernel<8>, dim3(1), dim3(kReplayThreads), 0, st, a, k, rb); break;
    default: hipLaunchKernelGGL(tight_replay_kernel<16>, dim3(1), dim3(kReplayThreads), 0, st, a, k, rb); break;
    }
    hip_check(hipGetLastError(), "tight replay launch");
    ++steps_;
    cur_ ^= 1;
}

void Device::render_view(uint32_t view, uint32_t axis, uint32_t slice, float *out)
{
    if (view > 2u || axis > 2u)
        throw Error(CS_ERR_ARG, "view must be 0, 1 or 2 and the slice axis 0, 1 or 2");
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    if (!view_)
        hip_check(hipMalloc(&view_, N_ * sizeof(float)), "hipMalloc view");
    StepArgs a{};
    a.src = lat_[cur_];
    a.reg = reg_;
    a.R0 = res_[0];
    a.R1 = res_[1];
    a.R2 = res_[2];
    a.R12 = res_[1] * res_[2];
    a.N = uint32_t(N_);
    a.vox_base = 0u;
    a.live = skew_->k.live;
    a.skew_n = skew_->k.ctl;
    const SkewTab k = skew_->k;
    const uint32_t blocks = uint32_t((N_ + kBlock - 1) / kBlock);
    switch (S_) {
    case 4: hipLaunchKernelGGL(view_kernel<4>, dim3(blocks), dim3(kBlock), 0, st, a, k, res_vol_, view, axis, slice, view_); break;
    case 8: hipLaunchKernelGGL(view_kernel<8>, dim3(blocks), dim3(kBlock), 0, st, a, k, res_vol_, view, axis, slice, view_); break;
    default: hipLaunchKernelGGL(view_kernel<16>, dim3(blocks), dim3(kBlock), 0, st, a, k, res_vol_, view, axis, slice, view_); break;
    }
    hip_check(hipGetLastError(), "view launch");
    hip_check(hipMemcpyAsync(out, view_, N_ * sizeof(float), hipMemcpyDeviceToHost, st), "read view");
    hip_check(hipStreamSynchronize(st), "read view");
}

void Device::read_skew(std::vector<SkewVoxel> &out)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    skew_->read(static_cast<hipStream_t>(stream_), out);
}

size_t Device::wave_shm_bytes() const
{
    switch (tab_mode_) {
    case kTabLds: return size_t(tab_words_) * 4u;
    case kTabSplit: return size_t(off_rates_) + size_t(bound_words(S_, pair_n1_)) * sizeof(float);
    default: return 0u;
    }
}

float Device::run_timed(const uint32_t *itime, uint32_t nsteps)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    std::vector<hipEvent_t> ev(2 * size_t(nsteps));
    for (auto &e : ev)
        hip_check(hipEventCreate(&e), "hipEventCreate");
    // The wave kernel's own duration: its dispatch stamps the pair. The other kernels (tiled, one lane per voxel)
    // are timed by events recorded around the step, which then also hold the tight replay and any CME phase.
    const bool wave = !simple_ && !tiled_;
    for (uint32_t t = 0; t < nsteps; ++t) {
        if (wave) {
            kernel_ev_[0] = ev[2 * t];
            kernel_ev_[1] = ev[2 * t + 1];
            try {
                step(itime[t]);
            } catch (...) {
                kernel_ev_[0] = kernel_ev_[1] = nullptr;
                throw;
            }
            kernel_ev_[0] = kernel_ev_[1] = nullptr;
        } else {
            hip_check(hipEventRecord(ev[2 * t], st), "hipEventRecord");
            step(itime[t]);
            hip_check(hipEventRecord(ev[2 * t + 1], st), "hipEventRecord");
        }
    }
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    float total = 0.f;
    for (uint32_t t = 0; t < nsteps; ++t) {
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, ev[2 * t], ev[2 * t + 1]), "hipEventElapsedTime");
        total += ms;
    }
    for (auto &e : ev)
        hipEventDestroy(e);
    return total;
}

void Device::synchronize()
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hip_check(hipStreamSynchronize(static_cast<hipStream_t>(stream_)), "hipStreamSynchronize");
}

void Device::read_lattice(uint8_t *out)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    hip_check(hipMemcpyAsync(out, lat_[cur_], N_ * S_, hipMemcpyDeviceToHost, st), "read lattice");
    hip_check(hipStreamSynchronize(st), "read lattice");
}

void Device::write_lattice(const uint8_t *in)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    hip_check(hipMemcpyAsync(lat_[cur_], in, N_ * S_, hipMemcpyHostToDevice, st), "write lattice");
    skew_->upload(st);  // the new lattice starts with the reference's counters on its slots (reservoir units aside)
    std::vector<uint8_t> bytes;
    if (summ_) {
        bytes.assign(size_t(wtiles_y_) * wtiles_z_ * res_[0], 0u);
        summary_bytes(in, res_, S_, wtiles_z_, bytes.data());
        hip_check(hipMemcpyAsync(sum_[cur_], bytes.data(), bytes.size(), hipMemcpyHostToDevice, st),
                  "upload occupancy summary");
    }
    plan_wave_items(in);
    hip_check(hipStreamSynchronize(st), "write lattice");
}

// A work item of the wave kernel: planes [x0, x1) of column col (block row col / tiles_z, z tile col % tiles_z).
struct WaveItem {
    double cost;
    uint32_t col, x0, x1;
};

// Work items of the wave kernel. By default plan_balanced_items (below); the cost-ordered plan (CS_WAVE_PLAN=cost,
// and fixed chunk lengths): each column of kWavesPerBlock wave tiles stacked in y is cut along x into
// ranges of about equal expected time, queued longest first. A tile's time is a fixed cost per plane plus a
// cost per particle (about 70 particles' worth per tile and plane, measured), and every range also marches
// the plane on each side of it; so ranges through dense regions are short (a long one would finish last)
// and ranges through empty space long (fewer side planes). The particle counts come from the given
// lattice: the spatial density changes slowly, so the plan of one state serves the steps after it.
void Device::plan_wave_items(const uint8_t *lat)
{
    const uint32_t tz_n = wtiles_z_, by_n = (wtiles_y_ + kWavesPerBlock - 1) / kWavesPerBlock;
    const uint32_t R0 = res_[0], gy = uint32_t(kWY * kWavesPerBlock);
    std::vector<uint64_t> cnt(size_t(tz_n) * by_n * R0, 0);  // particles per (column, plane)
    for (uint32_t x = 0; x < R0; ++x)
        for (uint32_t y = 0; y < res_[1]; ++y) {
            const uint8_t *row = lat + (size_t(x) * res_[1] + y) * res_[2] * S_;
            uint64_t *c = &cnt[(size_t(y / gy) * tz_n) * R0 + x];
            for (uint32_t z = 0; z < res_[2]; ++z) {
                const uint8_t *v = row + size_t(z) * S_;
                uint32_t k = 0;
                while (k < S_ && v[k])
                    ++k;
                c[size_t(z / uint32_t(kWZ)) * R0] += k;
            }
        }
    const double plane_cost = 70.0 * kWavesPerBlock;
    double total = 0.0;
    for (uint64_t v : cnt)
        total += double(v) + plane_cost;
    if (balanced_plan_) {
        plan_balanced_items(cnt, plane_cost, total);
        return;
    }
    // about `items_per_block_` items per resident block, so the longest ones end with the rest
    const uint32_t per_block = items_per_block_ != 0u ? items_per_block_ : 8u;
    const double target = std::max(total / (double(per_block) * std::max(wave_blocks_, 1u)), 2.0 * plane_cost);
    const uint32_t min_len = std::min<uint32_t>(xc_min_, R0), max_len = std::max(xc_, min_len);
    std::vector<WaveItem> items;
    for (uint32_t col = 0; col < tz_n * by_n; ++col) {
        const uint64_t *c = &cnt[size_t(col) * R0];
        uint32_t x0 = 0;
        while (x0 < R0) {
            double cost = 2.0 * plane_cost;  // the side planes
            uint32_t x1 = x0;
            while (x1 < R0 && (x1 - x0 < min_len || (x1 - x0 < max_len && cost + double(c[x1]) + plane_cost <= target))) {
                cost += double(c[x1]) + plane_cost;
                ++x1;
            }
            items.push_back({cost, col, x0, x1});
            x0 = x1;
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const WaveItem &p, const WaveItem &q) { return p.cost > q.cost; });
    // Queue k (XCD k) takes entries k, k + 8, ...: the columns of the k-th band of block rows (all their z tiles and
    // x ranges, cost order within the band), so the z and x neighbours of an item, whose halo lines it reads, run on
    // the same XCD and share its L2 (256^3: 10% fewer bytes read and 4.5% less time than dealing the cost-ordered
    // items round the XCDs, which CS_WAVE_PLACEMENT=rank restores; bands along x measured slower than either).
    // Shorter queues are padded with empty entries; an XCD whose queue is done helps the others.
    std::vector<std::vector<const WaveItem *>> q(8);
    const char *placement = std::getenv("CS_WAVE_PLACEMENT");
    if (placement != nullptr && std::string(placement) == "rank") {
        for (size_t i = 0; i < items.size(); ++i)
            q[i % 8].push_back(&items[i]);
    } else {
        for (const WaveItem &it : items)
            q[std::min<uint32_t>(7u, (it.col / tz_n) * 8u / std::max(by_n, 1u))].push_back(&it);
    }
    upload_items(q);
}

// Packs the per-XCD queues (entry e of the item array belongs to queue e % 8) and uploads them.
void Device::upload_items(const std::vector<std::vector<const WaveItem *>> &q)
{
    const uint32_t tz_n = wtiles_z_;
    size_t qlen = 0;
    for (const auto &l : q)
        qlen = std::max(qlen, l.size());
    std::vector<uint32_t> packed(2 * 8 * qlen);
    for (size_t i = 0; i < qlen; ++i)
        for (uint32_t k = 0; k < 8u; ++k) {
            const size_t e = i * 8u + k;
            if (i < q[k].size()) {
                const WaveItem &it = *q[k][i];
                packed[2 * e] = (it.col % tz_n) | ((it.col / tz_n) << 16);
                packed[2 * e + 1] = it.x0 | (it.x1 << 16);
            } else {
                packed[2 * e] = 0xFFFFu << 16;  // an empty entry: its block row lies past the lattice
                packed[2 * e + 1] = 0u;
            }
        }
    // trailing empty entries are never taken
    size_t n_packed = 8 * qlen;
    while (n_packed > 0 && packed[2 * (n_packed - 1)] == (0xFFFFu << 16))
        --n_packed;
    packed.resize(2 * n_packed);
    if (items_)
        hipFree(items_);
    items_ = nullptr;
    hip_check(hipMalloc(&items_, std::max<size_t>(packed.size(), 2) * sizeof(uint32_t)), "hipMalloc wave items");
    n_items_ = uint32_t(n_packed);
    hipStream_t st = static_cast<hipStream_t>(stream_);
    hip_check(hipMemcpyAsync(items_, packed.data(), packed.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st),
              "upload wave items");
    hip_check(hipStreamSynchronize(st), "upload wave items");
}

// Items of about equal cost, `items_per_block_` per resident block in all (default one each, so every block marches
// one item and no block waits for a second round). Column c (a block row of tiles at one z tile) gets a number of
// items in proportion to its cost (particles plus a fixed cost per tile plane, as above), largest remainders first,
// at least one and at most R0 / xc_min_, and is cut into that many x-ranges of equal cost. Queues: the columns in
// order (block rows, then z tiles) are dealt to the XCDs in runs of about equal item count, so the y and z neighbours
// of an item, whose halo lines it reads, mostly run on the same XCD and share its L2; an XCD whose queue is done
// helps the others.
void Device::plan_balanced_items(const std::vector<uint64_t> &cnt, double plane_cost, double total)
{
    const uint32_t tz_n = wtiles_z_, by_n = (wtiles_y_ + kWavesPerBlock - 1) / kWavesPerBlock;
    const uint32_t R0 = res_[0], ncol = tz_n * by_n;
    // at least 6 planes per item (shorter marches end sooner where a few blocks hold all the work: a populated 64^3
    // cell takes 47 us per step with 6-plane items, 72 with 12); 12 on an empty lattice, whose items the occupancy
    // summaries skip whole, so that only their fixed cost counts (an empty 128^3 cell: 43 us with 6, 26 with 12)
    uint64_t particles = 0;
    for (uint64_t v : cnt)
        particles += v;
    uint32_t min_default = particles == 0u ? 12u : 6u;
    // a small lattice alone on its device, whose 6-plane items would occupy under a quarter of the resident blocks,
    // takes 2-plane items instead: its step is then the time of a few plane passes on mostly idle CUs, and shorter
    // marches end sooner (a 64^3 cell: 45 -> 29 us per step; a 128^3 cell keeps 6, 0.376 against 0.406 ms with 2;
    // 512 cells of 32^3..96^3 stepped side by side, each on its own stream: 6.7 -> 8.4 ms per step with 2)
    const bool alone = live_sims(dev_).load() == (counted_ ? 1 : 0);
    if (particles != 0u && alone && uint64_t(ncol) * (R0 / 6u) * 4u < uint64_t(std::max(wave_blocks_, 1u)))
        min_default = 2u;
    const uint32_t min_len = std::max(1u, std::min<uint32_t>(min_planes_set_ ? xc_min_ : min_default, R0));
    const uint32_t max_per_col = std::max(1u, R0 / min_len);
    // items per resident block: one while an item holds up to about 24 planes per tile column, more on larger
    // lattices (an item of hundreds of planes would leave the cost model's error as a tail)
    const double blocks = double(std::max(wave_blocks_, 1u));
    const uint32_t per_block =
        items_per_block_ != 0u ? items_per_block_
                               : std::max(1u, uint32_t(std::lround(double(ncol) * double(R0) / (blocks * 24.0))));
    const double want = double(per_block) * blocks;
    std::vector<double> ccost(ncol, 0.0);
    for (uint32_t col = 0; col < ncol; ++col)
        for (uint32_t x = 0; x < R0; ++x)
            ccost[col] += double(cnt[size_t(col) * R0 + x]) + plane_cost;
    // items per column: floor of the share, then the largest remainders
    std::vector<uint32_t> n(ncol);
    std::vector<std::pair<double, uint32_t>> rem;
    double assigned = 0.0;
    for (uint32_t col = 0; col < ncol; ++col) {
        const double share = ccost[col] / total * want;
        n[col] = std::min(max_per_col, std::max(1u, uint32_t(share)));
        assigned += n[col];
        rem.push_back({share - double(n[col]), col});
    }
    std::stable_sort(rem.begin(), rem.end(), [](const auto &p, const auto &q) { return p.first > q.first; });
    for (size_t i = 0; i < rem.size() && assigned + 1.0 <= want; ++i)
        if (rem[i].first > 0.0 && n[rem[i].second] < max_per_col) {
            ++n[rem[i].second];
            assigned += 1.0;
        }
    std::vector<WaveItem> items;
    for (uint32_t col = 0; col < ncol; ++col) {
        const uint64_t *c = &cnt[size_t(col) * R0];
        const double step = ccost[col] / double(n[col]);
        double acc = 0.0;
        uint32_t x0 = 0, x = 0;
        for (uint32_t k = 1; k <= n[col]; ++k) {
            if (k == n[col]) {
                x = R0;
            } else {
                const double cut = step * double(k);
                // at least min_len planes per range, and min_len left for each range after it
                while (x < R0 && (x - x0 < min_len || acc + (double(c[x]) + plane_cost) * 0.5 <= cut) &&
                       R0 - (x + 1u) >= (n[col] - k) * min_len) {
                    acc += double(c[x]) + plane_cost;
                    ++x;
                }
            }
            if (x > x0)
                items.push_back({0.0, col, x0, x});
            x0 = x;
        }
    }
    // deal to the XCD queues in column order, runs of about equal length
    std::vector<std::vector<const WaveItem *>> q(8);
    for (size_t i = 0; i < items.size(); ++i)
        q[std::min<size_t>(7u, i * 8u / std::max<size_t>(items.size(), 1u))].push_back(&items[i]);
    upload_items(q);
}

void Device::read_cme_counts(int32_t *out)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    if (nc_ > 0u)
        hip_check(hipMemcpyAsync(out, cme_counts_, nc_ * sizeof(int32_t), hipMemcpyDeviceToHost, st), "read cme counts");
    hip_check(hipStreamSynchronize(st), "read cme counts");
}

void Device::read_cme_audit(double *max_rel_error, uint64_t *divergences, uint64_t *events)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    unsigned long long h[3] = {0, 0, 0};
    hip_check(hipMemcpyAsync(h, audit_, sizeof(h), hipMemcpyDeviceToHost, st), "read audit");
    hip_check(hipStreamSynchronize(st), "read audit");
    double rel;
    std::memcpy(&rel, &h[0], sizeof(rel));
    if (max_rel_error)
        *max_rel_error = rel;
    if (divergences)
        *divergences = h[1];
    if (events)
        *events = h[2];
}

void Device::region_counts(int32_t *out)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    const uint32_t nbins = nreg_ * (ns_ + 1u);
    hip_check(hipMemsetAsync(hist_, 0, nbins * sizeof(int32_t), st), "clear hist");
    const uint32_t blocks = 1024;
    const size_t shmem = nbins * sizeof(int32_t);
    if (shmem > 64 * 1024)
        throw Error(CS_ERR_UNSUPPORTED, "region x species histogram exceeds 64 KiB of LDS");
    const uint32_t N = uint32_t(N_);
    switch (S_) {
    case 4: hipLaunchKernelGGL(region_count_kernel<4>, dim3(blocks), dim3(kBlock), shmem, st, lat_[cur_], reg_, N, ns_, nreg_, hist_); break;
    case 8: hipLaunchKernelGGL(region_count_kernel<8>, dim3(blocks), dim3(kBlock), shmem, st, lat_[cur_], reg_, N, ns_, nreg_, hist_); break;
    case 16: hipLaunchKernelGGL(region_count_kernel<16>, dim3(blocks), dim3(kBlock), shmem, st, lat_[cur_], reg_, N, ns_, nreg_, hist_); break;
    default: throw Error(CS_ERR_UNSUPPORTED, "slots per voxel must be 4, 8 or 16");
    }
    hip_check(hipGetLastError(), "region_count_kernel launch");
    hip_check(hipMemcpyAsync(out, hist_, nbins * sizeof(int32_t), hipMemcpyDeviceToHost, st), "read hist");
    hip_check(hipStreamSynchronize(st), "region counts");
    // the reference's counts hold what the slots of skewed voxels do not (reservoir units, slot-S aliases,
    // counts below zero)
    std::vector<SkewVoxel> sk;
    read_skew(sk);
    for (const SkewVoxel &e : sk)
        for (int q = 0; q < 4; ++q)
            if (e.sp[q] != 0u && e.sp[q] <= ns_)
                out[size_t(reg_host_[e.v]) * (ns_ + 1u) + e.sp[q]] += e.dc[q];
}

void Device::region_counts_noalloc(int32_t *out)
{
    const uint32_t n1 = ns_ + 1u, nbins = nreg_ * n1;
    if (nbins > kNoAllocBins)
        throw Error(CS_ERR_ARG, "regions x (species + 1) exceeds the " + std::to_string(kNoAllocBins) +
                                    " bins of the fixed histogram");
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    int32_t *bins = noalloc_bins(dev_);
    // the array keeps the previous call's bins: cleared by a kernel (no memset node for graph capture)
    hipLaunchKernelGGL(fill_i32_kernel, dim3((nbins + kBlock - 1) / kBlock), dim3(kBlock), 0, st, bins, nbins, 0);
    const uint32_t N = uint32_t(N_);
    const size_t shmem = nbins * sizeof(int32_t);
    switch (S_) {
    case 4: hipLaunchKernelGGL(region_count_kernel<4>, dim3(1024), dim3(kBlock), shmem, st, lat_[cur_], reg_, N, ns_, nreg_, bins); break;
    case 8: hipLaunchKernelGGL(region_count_kernel<8>, dim3(1024), dim3(kBlock), shmem, st, lat_[cur_], reg_, N, ns_, nreg_, bins); break;
    default: hipLaunchKernelGGL(region_count_kernel<16>, dim3(1024), dim3(kBlock), shmem, st, lat_[cur_], reg_, N, ns_, nreg_, bins); break;
    }
    hipLaunchKernelGGL(skew_region_kernel, dim3(64), dim3(kBlock), 0, st, skew_->k, reg_, ns_, bins);
    hip_check(hipGetLastError(), "region count launches");
    hip_check(hipMemcpyAsync(out, bins, nbins * sizeof(int32_t), hipMemcpyDeviceToHost, st), "read bins");
    hip_check(hipStreamSynchronize(st), "region counts");
}

void Device::read_counters(uint64_t *out)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    hip_check(hipMemcpyAsync(out, counters_, kNumCounters * sizeof(uint64_t), hipMemcpyDeviceToHost, st),
              "read counters");
    hip_check(hipStreamSynchronize(st), "read counters");
}

void Device::reset_counters()
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    hip_check(hipMemsetAsync(counters_, 0, kNumCounters * sizeof(unsigned long long), st), "clear counters");
    hip_check(hipStreamSynchronize(st), "clear counters");
}


// ---- Ensemble (host) -------------------------------------------------------------------------------------
// Descriptor fields that do not depend on the tiling are filled once; tiles, chunks and item offsets are
// refilled by build_descriptors() whenever the tiling changes.

Ensemble::Ensemble(const std::vector<const Model *> &cells, const std::vector<uint32_t> &itime_seeds,
                   int hip_device)
    : dev_(hip_device)
{
    if (cells.empty())
        throw Error(CS_ERR_ARG, "an ensemble needs at least one cell");
    if (itime_seeds.size() != cells.size())
        throw Error(CS_ERR_ARG, "one itime seed per cell");
    S_ = cells[0]->slots();
    for (const Model *m : cells) {
        if (!m)
            throw Error(CS_ERR_ARG, "null cell");
        if (m->slots() != S_)
            throw Error(CS_ERR_ARG, "all cells of an ensemble have the same slots per voxel");
        check_device_model(*m);
    }
    int ndev = 0;
    hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (hip_device < 0 || hip_device >= ndev)
        throw Error(CS_ERR_ARG, "no such HIP device");
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    int cus = 0;
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_), "CU count");
    n_cu_ = uint32_t(cus > 0 ? cus : 256);
    hipStream_t st;
    hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    stream_ = st;

    const uint32_t n = uint32_t(cells.size());
    n_cells_ = n;
    std::vector<uint32_t> tab_words;  // all table blobs, 8-byte aligned
    std::vector<uint8_t> cme_bytes;
    std::vector<CmeBlob> cme_blobs;
    std::vector<uint32_t> cme_blob_off;
    std::vector<int32_t> cme_init;
    std::vector<EnsCell> desc(n);
    lat_off_.resize(n);
    vox_off_.resize(n);
    res_.resize(3 * size_t(n));
    cell_nc_.resize(n);
    cell_ns_.resize(n);
    cme_counts_off_.resize(n);
    uint64_t vox = 0;
    int32_t n_cme = 0;
    for (uint32_t c = 0; c < n; ++c) {
        const Model &m = *cells[c];
        const LatticeParams &lp = m.params();
        EnsCell &d = desc[c];
        std::memset(&d, 0, sizeof(d));
        lat_off_[c] = vox * S_;
        vox_off_[c] = vox;
        d.lat_off = vox * S_;
        d.vox_off = uint32_t(vox);
        d.R0 = lp.res[0];
        d.R1 = lp.res[1];
        d.R2 = lp.res[2];
        for (int i = 0; i < 3; ++i)
            res_[3 * size_t(c) + i] = lp.res[i];
        vox += m.n_voxels();
        if (vox >= (uint64_t(1) << 32) / S_)
            throw Error(CS_ERR_UNSUPPORTED, "ensemble too large for u32 voxel indices");
        const TableBlob b = make_table_blob(m);
        if (b.bytes.size() > kTabLdsMax)
            throw Error(CS_ERR_UNSUPPORTED, "cell tables too large to stage in LDS");
        if (lp.res[1] > 4096u || lp.res[2] > 4096u || m.num_regions() >= 255u)
            throw Error(CS_ERR_UNSUPPORTED, "ensemble cells need y and z extents up to 4096 and fewer than 255 regions");
        d.tab_off = uint32_t(tab_words.size());
        d.tab_words = uint32_t(b.bytes.size() / 4u);
        tab_words.resize(tab_words.size() + d.tab_words);
        std::memcpy(tab_words.data() + d.tab_off, b.bytes.data(), b.bytes.size());
        tab_words_max_ = std::max(tab_words_max_, d.tab_words);
        d.off_p2 = b.off_p2;
        d.off_rates = b.off_rates;
        d.off_ridx = b.off_ridx;
        d.off_stoich = b.off_stoich;
        d.off_rmask = b.off_rmask;
        d.off_bound = b.off_bound;
        d.pair_n1 = b.pair_n1;
        d.ns = m.num_species();
        d.nr = m.num_reactions();
        d.nreg = m.num_regions();
        d.nrw = b.nrw;
        d.has_zero_order = b.has_zero_order ? 1u : 0u;
        d.mod_zero = 0xFFFFFFFFu;
        d.tau = lp.tau;
        d.itime_seed = itime_seeds[c];
        cell_nc_[c] = m.num_cme_species();
        cell_ns_[c] = m.num_species();
        cme_counts_off_[c] = uint32_t(cme_init.size());
        d.cme_counts_off = cme_counts_off_[c];
        cme_init.insert(cme_init.end(), m.cme_init().begin(), m.cme_init().end());
        d.cme = -1;
        d.cme_seed = m.cme_seed();
        if (m.num_cme_reactions() > 0) {
            d.cme = n_cme++;
            cme_blobs.push_back(make_cme_blob(m));
            cme_blob_off.push_back(uint32_t(cme_bytes.size()));
            cme_bytes.insert(cme_bytes.end(), cme_blobs.back().bytes.begin(), cme_blobs.back().bytes.end());
        }
        if (tab_words.size() % 2u)
            tab_words.push_back(0u);  // keep every blob 8-byte aligned
        // the cell's occupancy summaries (none with zero-order reactions, which can fire in any voxel)
        const uint64_t nb = uint64_t((d.R1 + kWY - 1) / kWY) * ((d.R2 + kWZ - 1) / kWZ) * d.R0;
        if (sum_total_ + nb >= (uint64_t(1) << 32))
            throw Error(CS_ERR_UNSUPPORTED, "occupancy summaries of the ensemble exceed 2^32 bytes");
        d.sum_off = uint32_t(sum_total_);
        d.sum_planes = b.has_zero_order ? 0u : d.R0;
        sum_total_ += nb;
    }
    total_vox_ = vox;
    cells_.resize(sizeof(EnsCell) * n);
    std::memcpy(cells_.data(), desc.data(), cells_.size());

    hip_check(hipMalloc(&lat_[0], total_vox_ * S_), "hipMalloc ensemble lattice");
    hip_check(hipMalloc(&lat_[1], total_vox_ * S_), "hipMalloc ensemble lattice");
    hip_check(hipMalloc(&reg_, total_vox_), "hipMalloc ensemble regions");
    // padded labels per cell (see Device): two voxels of 0xFF around the lattice, rows padded to wave tiles
    regp_off_.assign(n, 0);
    uint64_t regp_words = 0;
    {
        EnsCell *d = desc.data();
        for (uint32_t c = 0; c < n; ++c) {
            d[c].PY = kWY * ((d[c].R1 + kWY - 1) / kWY) + 4u;
            d[c].PZ4 = (kWZ * ((d[c].R2 + kWZ - 1) / kWZ) + 4u) / 4u;
            regp_off_[c] = regp_words;
            if (regp_words >= (uint64_t(1) << 32))
                throw Error(CS_ERR_UNSUPPORTED, "padded labels of the ensemble exceed 2^32 words");
            d[c].regp_off = uint32_t(regp_words);
            regp_words += uint64_t(d[c].R0 + 4u) * d[c].PY * d[c].PZ4;
        }
        std::memcpy(cells_.data(), desc.data(), cells_.size());
    }
    hip_check(hipMalloc(&regp_, std::max<uint64_t>(regp_words, 1) * 4u), "hipMalloc ensemble padded regions");
    hip_check(hipMalloc(&tab_, tab_words.size() * sizeof(uint32_t)), "hipMalloc ensemble tables");
    hip_check(hipMalloc(&cells_dev_, cells_.size()), "hipMalloc ensemble cells");
    hip_check(hipMalloc(&itime_, size_t(n) * sizeof(uint32_t)), "hipMalloc itime");
    hip_check(hipMalloc(&queue_, 2 * sizeof(uint32_t)), "hipMalloc queue");
    hip_check(hipMalloc(&counters_, kNumCounters * sizeof(unsigned long long)), "hipMalloc counters");
    hip_check(hipMalloc(&audit_, 3 * sizeof(unsigned long long)), "hipMalloc audit");
    hip_check(hipMalloc(&cme_counts_, std::max<size_t>(cme_init.size(), 1) * sizeof(int32_t)), "hipMalloc cme counts");
    hip_check(hipMalloc(&fix_ctl_, kFixCtl * sizeof(uint32_t)), "hipMalloc fix_ctl");
    {
        // skewed voxels of every cell, starting with the cells' reservoir voxels
        uint64_t n_static = 0, max_clusters = 0;
        for (const Model *m : cells) {
            n_static += count_reservoir_voxels(*m);
            max_clusters += cluster_bound(m->params().res);
        }
        skew_ = new SkewStore();
        skew_->alloc(total_vox_, S_, n_static, fix_ctl_, max_clusters);
        for (uint32_t c = 0; c < n; ++c)
            add_reservoir_entries(*skew_, *cells[c], uint32_t(vox_off_[c]));
    }
    if (!cme_bytes.empty())
        hip_check(hipMalloc(&cme_tab_, cme_bytes.size()), "hipMalloc ensemble cme tables");
    if (n_cme > 0)
        hip_check(hipMalloc(&cme_args_dev_, size_t(n_cme) * sizeof(CmeArgs)), "hipMalloc cme args");

    std::vector<CmeArgs> cargs;
    for (uint32_t c = 0, k = 0; c < n; ++c) {
        if (desc[c].cme < 0)
            continue;
        cargs.push_back(cme_args_at(static_cast<const uint8_t *>(cme_tab_) + cme_blob_off[k], cme_blobs[k], *cells[c],
                                    counters_, audit_, false));
        ++k;
    }
    const size_t sum_bytes = std::max<uint64_t>(sum_total_, 1);
    for (int b = 0; b < 2; ++b)
        hip_check(hipMalloc(&sum_[b], sum_bytes), "hipMalloc occupancy summary");
    hip_check(hipMemsetAsync(lat_[1], 0, total_vox_ * S_, st), "clear ensemble lattice");
    hip_check(hipMemsetAsync(sum_[1], 0, sum_bytes, st), "clear occupancy summary");
    std::vector<uint8_t> sum0(sum_bytes, 0u);
    for (uint32_t c = 0; c < n; ++c)
        if (desc[c].sum_planes != 0u) {
            const uint32_t tzn = (desc[c].R2 + kWZ - 1) / kWZ;
            const size_t nb = size_t((desc[c].R1 + kWY - 1) / kWY) * tzn * desc[c].R0;
            std::vector<uint8_t> cs(nb, 0u);
            summary_bytes(cells[c]->lattice().data(), &res_[3 * size_t(c)], S_, tzn, cs.data());
            if (summary_pays(cs))
                std::memcpy(sum0.data() + desc[c].sum_off, cs.data(), nb);
            else
                desc[c].sum_planes = 0u;  // the cell keeps no summary (see summary_pays)
        }
    std::memcpy(cells_.data(), desc.data(), cells_.size());
    hip_check(hipMemcpyAsync(sum_[0], sum0.data(), sum_bytes, hipMemcpyHostToDevice, st), "upload occupancy summary");
    for (uint32_t c = 0; c < n; ++c) {
        const Model &m = *cells[c];
        hip_check(hipMemcpyAsync(lat_[0] + lat_off_[c], m.lattice().data(), m.n_voxels() * S_, hipMemcpyHostToDevice, st),
                  "upload ensemble lattice");
        std::vector<uint8_t> reg8(m.n_voxels());
        for (uint64_t v = 0; v < m.n_voxels(); ++v)
            reg8[v] = uint8_t(m.regions()[v]);
        hip_check(hipMemcpyAsync(reg_ + vox_off_[c], reg8.data(), reg8.size(), hipMemcpyHostToDevice, st),
                  "upload ensemble regions");
        const EnsCell &dc = desc[c];
        const uint32_t PZ = dc.PZ4 * 4u;
        std::vector<uint8_t> regp(size_t(dc.R0 + 4u) * dc.PY * PZ, 0xFFu);
        for (uint32_t x = 0; x < dc.R0; ++x)
            for (uint32_t y = 0; y < dc.R1; ++y)
                std::memcpy(&regp[(size_t(x + 2u) * dc.PY + (y + 2u)) * PZ + 2u],
                            &reg8[(size_t(x) * dc.R1 + y) * dc.R2], dc.R2);
        hip_check(hipMemcpyAsync(static_cast<uint32_t *>(regp_) + regp_off_[c], regp.data(), regp.size(),
                                 hipMemcpyHostToDevice, st), "upload ensemble padded regions");
        hip_check(hipStreamSynchronize(st), "upload ensemble cell");  // reg8 is freed next iteration
    }
    hip_check(hipMemcpyAsync(tab_, tab_words.data(), tab_words.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st),
              "upload ensemble tables");
    if (!cme_bytes.empty())
        hip_check(hipMemcpyAsync(cme_tab_, cme_bytes.data(), cme_bytes.size(), hipMemcpyHostToDevice, st),
                  "upload ensemble cme tables");
    if (!cargs.empty())
        hip_check(hipMemcpyAsync(cme_args_dev_, cargs.data(), cargs.size() * sizeof(CmeArgs), hipMemcpyHostToDevice, st),
                  "upload cme args");
    if (!cme_init.empty())
        hip_check(hipMemcpyAsync(cme_counts_, cme_init.data(), cme_init.size() * sizeof(int32_t), hipMemcpyHostToDevice, st),
                  "upload cme counts");
    hip_check(hipMemsetAsync(queue_, 0, 2 * sizeof(uint32_t), st), "clear queue");
    hip_check(hipMemsetAsync(counters_, 0, kNumCounters * sizeof(unsigned long long), st), "clear counters");
    hip_check(hipMemsetAsync(audit_, 0, 3 * sizeof(unsigned long long), st), "clear audit");
    hip_check(hipMemsetAsync(fix_ctl_, 0, kFixCtl * sizeof(uint32_t), st), "clear fix_ctl");
    skew_->upload(st);
    build_descriptors();
    hip_check(hipStreamSynchronize(st), "ensemble upload");
}

Ensemble::~Ensemble()
{
    hipSetDevice(dev_);
    if (stream_)
        hipStreamSynchronize(static_cast<hipStream_t>(stream_));
    destroy_graph();
    void *bufs[] = {lat_[0], lat_[1], reg_, regp_, tab_, cells_dev_, cme_args_dev_, cme_tab_, cme_counts_, itime_, queue_,
                    counters_, audit_, fix_ctl_, sum_[0], sum_[1], item_cell_};
    for (void *b : bufs)
        if (b)
            (void)hipFree(b);
    if (skew_) {
        skew_->release();
        delete skew_;
    }
    if (stream_)
        (void)hipStreamDestroy(static_cast<hipStream_t>(stream_));
}

void Ensemble::build_descriptors()
{
    const uint32_t n = n_cells();
    EnsCell *d = reinterpret_cast<EnsCell *>(cells_.data());
    uint64_t items = 0;
    for (uint32_t c = 0; c < n; ++c) {
        d[c].tiles_z = (d[c].R2 + kWZ - 1) / kWZ;
        d[c].wtiles_y = (d[c].R1 + kWY - 1) / kWY;
        d[c].btiles_y = (d[c].tiles_z * d[c].wtiles_y + kWavesPerBlock - 1) / kWavesPerBlock;
        d[c].chunks = (d[c].R0 + tiling_.xc - 1) / tiling_.xc;
        d[c].first_item = uint32_t(items);
        items += uint64_t(d[c].btiles_y) * d[c].chunks;
    }
    if (items >= (uint64_t(1) << 31))
        throw Error(CS_ERR_UNSUPPORTED, "too many work items");
    total_items_ = uint32_t(items);
    // the queue holds the full-length items first and the shorter last chunks of the cells behind them (each
    // group in cell order), so that the items still running when the queue empties are the short ones
    std::vector<uint32_t> order(items);
    for (uint32_t c = 0; c < n; ++c)
        for (uint32_t i = 0; i < d[c].btiles_y * d[c].chunks; ++i)
            order[d[c].first_item + i] = c;
    auto item_len = [&](uint32_t c, uint32_t local) {
        const uint32_t chunk = local / d[c].btiles_y;
        return std::min(tiling_.xc, d[c].R0 - chunk * tiling_.xc);
    };
    std::vector<uint32_t> item_cell(2 * std::max<uint64_t>(items, 1), 0u);
    {
        std::vector<uint32_t> idx(items);
        for (uint32_t i = 0; i < uint32_t(items); ++i)
            idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t p, uint32_t q) {
            return item_len(order[p], p - d[order[p]].first_item) > item_len(order[q], q - d[order[q]].first_item);
        });
        for (uint32_t i = 0; i < uint32_t(items); ++i) {
            item_cell[2 * i] = order[idx[i]];
            item_cell[2 * i + 1] = idx[i] - d[order[idx[i]]].first_item;
        }
    }
    hipStream_t st = static_cast<hipStream_t>(stream_);
    if (item_cell_)
        hip_check(hipFree(item_cell_), "hipFree item cells");
    item_cell_ = nullptr;
    hip_check(hipMalloc(&item_cell_, item_cell.size() * sizeof(uint32_t)), "hipMalloc item cells");
    hip_check(hipMemcpyAsync(item_cell_, item_cell.data(), item_cell.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st),
              "upload item cells");
    hip_check(hipMemcpyAsync(cells_dev_, cells_.data(), cells_.size(), hipMemcpyHostToDevice, st), "upload cells");
    hip_check(hipStreamSynchronize(st), "upload cells");  // cells_ and item_cell may change again before the copy ran
}

void Ensemble::set_tiling(const EnsembleTiling &t)
{
    if (t.xc < 1 || t.xc > 4096 || t.blocks_per_cu < 1 || t.blocks_per_cu > 16 || t.batch > 64)
        throw Error(CS_ERR_ARG, "tiling out of range");
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    synchronize();
    tiling_ = t;
    destroy_graph();
    build_descriptors();
}

void Ensemble::destroy_graph()
{
    if (graph_exec_) {
        (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec_));
        graph_exec_ = nullptr;
        graph_block_ = 0;
    }
}

void Ensemble::enqueue_step()
{
    hipStream_t st = static_cast<hipStream_t>(stream_);
    EnsArgs e;
    e.cells = static_cast<const EnsCell *>(cells_dev_);
    e.n_cells = n_cells();
    e.total_items = total_items_;
    e.xc = tiling_.xc;
    e.batch = tiling_.batch;
    e.src = lat_[cur_];
    e.dst = lat_[cur_ ^ 1];
    e.sum_src = sum_[cur_];
    e.sum_dst = sum_[cur_ ^ 1];
    e.item_cell = item_cell_;
    e.reg = reg_;
    e.regp = static_cast<const uint32_t *>(regp_);
    e.tab = tab_;
    e.itime = itime_;
    e.queue = queue_;
    e.counters = counters_;
    e.fix_ctl = fix_ctl_;
    e.fix_list = skew_->b.T;
    e.max_tight = skew_->b.tcap;
    e.live = skew_->k.live;
    e.skew_n = skew_->k.ctl;
    const SkewTab k = skew_->k;
    const ReplayBuf rb = skew_->b;
    const uint32_t pro_blocks = (n_cells() + 63u) / 64u;
    const CmeArgs *ca = static_cast<const CmeArgs *>(cme_args_dev_);
    const uint32_t grid = std::min(tiling_.blocks_per_cu * n_cu_, std::max(total_items_, 1u));
    const size_t shm = size_t(tab_words_max_) * 4u;
    switch (S_) {
    case 4:
        hipLaunchKernelGGL(ens_prologue<4>, dim3(pro_blocks), dim3(64), 0, st, e, ca, cme_counts_);
        hipLaunchKernelGGL(rdme_step_ensemble<4>, dim3(grid), dim3(kWaveBlock), shm, st, e);
        hipLaunchKernelGGL(ens_epilogue<4>, dim3(1), dim3(kReplayThreads), 0, st, e, k, rb);
        break;
    case 8:
        hipLaunchKernelGGL(ens_prologue<8>, dim3(pro_blocks), dim3(64), 0, st, e, ca, cme_counts_);
        hipLaunchKernelGGL(rdme_step_ensemble<8>, dim3(grid), dim3(kWaveBlock), shm, st, e);
        hipLaunchKernelGGL(ens_epilogue<8>, dim3(1), dim3(kReplayThreads), 0, st, e, k, rb);
        break;
    default:
        hipLaunchKernelGGL(ens_prologue<16>, dim3(pro_blocks), dim3(64), 0, st, e, ca, cme_counts_);
        hipLaunchKernelGGL(rdme_step_ensemble<16>, dim3(grid), dim3(kWaveBlock), shm, st, e);
        hipLaunchKernelGGL(ens_epilogue<16>, dim3(1), dim3(kReplayThreads), 0, st, e, k, rb);
        break;
    }
    hip_check(hipGetLastError(), "ensemble step launch");
    cur_ ^= 1;
}

void Ensemble::run(uint32_t nsteps)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    for (uint32_t t = 0; t < nsteps; ++t)
        enqueue_step();
    steps_ += nsteps;
}

void Ensemble::run_graph(uint32_t nsteps, uint32_t block)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    if (block < 2)
        return run(nsteps);
    block += block & 1u;  // even, so the ping-pong parity is the same at the start of every replay
    hipStream_t st = static_cast<hipStream_t>(stream_);
    uint32_t done = 0;
    if (nsteps >= block)
        capture(block);
    for (; done + block <= nsteps; done += block)
        hip_check(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec_), st), "graph launch");
    steps_ += done;
    run(nsteps - done);
}

void Ensemble::capture(uint32_t block)
{
    hipStream_t st = static_cast<hipStream_t>(stream_);
    if (graph_exec_ && (graph_block_ != block || graph_parity_ != cur_))
        destroy_graph();
    if (!graph_exec_) {
        hipGraph_t g = nullptr;
        const int parity = cur_;
        hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            for (uint32_t k = 0; k < block; ++k)
                enqueue_step();
        } catch (...) {
            // leave capture mode before reporting, so the stream stays usable
            hipGraph_t partial = nullptr;
            if (hipStreamEndCapture(st, &partial) == hipSuccess && partial)
                (void)hipGraphDestroy(partial);
            cur_ = parity;
            throw;
        }
        hip_check(hipStreamEndCapture(st, &g), "end capture");
        cur_ = parity;  // capture recorded the launches; nothing ran
        hipGraphExec_t ex;
        hip_check(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0), "graph instantiate");
        (void)hipGraphDestroy(g);
        graph_exec_ = ex;
        graph_block_ = block;
        graph_parity_ = parity;
    }
}

float Ensemble::run_timed(uint32_t nsteps, uint32_t block)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    if (block >= 2 && nsteps >= block + (block & 1u))
        capture(block + (block & 1u));  // outside the timed window
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    hip_check(hipEventRecord(e0, st), "hipEventRecord");
    if (block >= 2)
        run_graph(nsteps, block);
    else
        run(nsteps);
    hip_check(hipEventRecord(e1, st), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

float Ensemble::autotune(uint32_t trial_steps)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    synchronize();
    // snapshot: current lattice, well-mixed counts, step counter, event counters
    const size_t lat_bytes = total_vox_ * S_;
    const size_t nc_total = std::max<size_t>(1, cme_counts_off_.empty() ? 1 : cme_counts_off_.back() + cell_nc_.back());
    const size_t sum_bytes = std::max<uint64_t>(sum_total_, 1);
    uint8_t *lat_snap = nullptr;
    int32_t *cme_snap = nullptr;
    uint32_t *q_snap = nullptr;
    unsigned long long *cnt_snap = nullptr;
    uint8_t *sum_snap = nullptr;
    hip_check(hipMalloc(&lat_snap, 2 * lat_bytes), "hipMalloc snapshot");
    hip_check(hipMalloc(&sum_snap, 2 * sum_bytes), "hipMalloc snapshot");
    hip_check(hipMalloc(&cme_snap, nc_total * sizeof(int32_t)), "hipMalloc snapshot");
    hip_check(hipMalloc(&q_snap, 2 * sizeof(uint32_t)), "hipMalloc snapshot");
    hip_check(hipMalloc(&cnt_snap, kNumCounters * sizeof(unsigned long long)), "hipMalloc snapshot");
    // the skew table: keys and their entry indices, entries, control words and live bits
    const SkewTab &kt = skew_->k;
    const size_t key_bytes = size_t(kt.mask + 1u) * sizeof(uint32_t), ent_bytes = size_t(kt.ecap) * sizeof(SkewEntry);
    const size_t live_bytes = std::max<size_t>((total_vox_ + 31u) / 32u, 1) * sizeof(uint32_t);
    uint8_t *sk_snap = nullptr;
    hip_check(hipMalloc(&sk_snap, 2 * key_bytes + ent_bytes + live_bytes + 4 * sizeof(uint32_t)), "hipMalloc snapshot");
    struct Part {
        void *dev;
        size_t bytes;
    };
    const Part parts[5] = {{kt.keys, key_bytes}, {kt.idx, key_bytes}, {kt.ent, ent_bytes}, {kt.live, live_bytes},
                           {kt.ctl, 4 * sizeof(uint32_t)}};
    const int cur0 = cur_;
    const uint64_t steps0 = steps_;
    // both buffers: the one a step writes is also the one whose emptiness its summary vouches for
    auto save = [&] {
        for (int b = 0; b < 2; ++b) {
            hip_check(hipMemcpyAsync(lat_snap + b * lat_bytes, lat_[b], lat_bytes, hipMemcpyDeviceToDevice, st), "snapshot");
            hip_check(hipMemcpyAsync(sum_snap + b * sum_bytes, sum_[b], sum_bytes,
                                     hipMemcpyDeviceToDevice, st), "snapshot");
        }
        hip_check(hipMemcpyAsync(cme_snap, cme_counts_, nc_total * sizeof(int32_t), hipMemcpyDeviceToDevice, st), "snapshot");
        hip_check(hipMemcpyAsync(q_snap, queue_, 2 * sizeof(uint32_t), hipMemcpyDeviceToDevice, st), "snapshot");
        hip_check(hipMemcpyAsync(cnt_snap, counters_, kNumCounters * sizeof(unsigned long long), hipMemcpyDeviceToDevice, st),
                  "snapshot");
        size_t o = 0;
        for (const Part &p : parts) {
            hip_check(hipMemcpyAsync(sk_snap + o, p.dev, p.bytes, hipMemcpyDeviceToDevice, st), "snapshot");
            o += p.bytes;
        }
    };
    auto restore = [&] {
        synchronize();
        cur_ = cur0;
        steps_ = steps0;
        for (int b = 0; b < 2; ++b) {
            hip_check(hipMemcpyAsync(lat_[b], lat_snap + b * lat_bytes, lat_bytes, hipMemcpyDeviceToDevice, st), "restore");
            hip_check(hipMemcpyAsync(sum_[b], sum_snap + b * sum_bytes, sum_bytes,
                                     hipMemcpyDeviceToDevice, st), "restore");
        }
        hip_check(hipMemcpyAsync(cme_counts_, cme_snap, nc_total * sizeof(int32_t), hipMemcpyDeviceToDevice, st), "restore");
        hip_check(hipMemcpyAsync(queue_, q_snap, 2 * sizeof(uint32_t), hipMemcpyDeviceToDevice, st), "restore");
        hip_check(hipMemcpyAsync(counters_, cnt_snap, kNumCounters * sizeof(unsigned long long), hipMemcpyDeviceToDevice, st),
                  "restore");
        size_t o = 0;
        for (const Part &p : parts) {
            hip_check(hipMemcpyAsync(p.dev, sk_snap + o, p.bytes, hipMemcpyDeviceToDevice, st), "restore");
            o += p.bytes;
        }
        synchronize();
    };
    save();
    const EnsembleTiling before = tiling_;
    EnsembleTiling best = before;
    float best_ms = -1.f;
    const uint32_t xcs[] = {8, 16, 32, 48, 60};  // up to 60 planes keep the occupancy summaries usable
    const uint32_t bpcs[] = {2, 3, 4};
    const uint32_t batches[] = {0, 1, 4};
    for (uint32_t xc : xcs)
        for (uint32_t bpc : bpcs)
            for (uint32_t batch : batches) {
                set_tiling(EnsembleTiling{xc, bpc, batch});
                run(1);  // warm-up
                const float ms = run_timed(trial_steps, 0) / float(trial_steps);
                restore();
                if (best_ms < 0.f || ms < best_ms) {
                    best_ms = ms;
                    best = tiling_;
                }
            }
    set_tiling(best);
    (void)hipFree(lat_snap);
    (void)hipFree(sum_snap);
    (void)hipFree(cme_snap);
    (void)hipFree(q_snap);
    (void)hipFree(cnt_snap);
    (void)hipFree(sk_snap);
    return best_ms;
}

void Ensemble::synchronize()
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hip_check(hipStreamSynchronize(static_cast<hipStream_t>(stream_)), "hipStreamSynchronize");
}

uint64_t Ensemble::cell_voxels(uint32_t c) const
{
    if (c >= n_cells())
        throw Error(CS_ERR_ARG, "cell index out of range");
    return uint64_t(res_[3 * size_t(c)]) * res_[3 * size_t(c) + 1] * res_[3 * size_t(c) + 2];
}

void Ensemble::cell_res(uint32_t c, uint32_t res[3]) const
{
    if (c >= n_cells())
        throw Error(CS_ERR_ARG, "cell index out of range");
    for (int i = 0; i < 3; ++i)
        res[i] = res_[3 * size_t(c) + i];
}

void Ensemble::read_lattice(uint32_t cell, uint8_t *out)
{
    const uint64_t nv = cell_voxels(cell);
    synchronize();
    hip_check(hipMemcpy(out, lat_[cur_] + lat_off_[cell], nv * S_, hipMemcpyDeviceToHost), "read cell lattice");
}

void Ensemble::read_skew(uint32_t cell, std::vector<SkewVoxel> &out)
{
    const uint64_t nv = cell_voxels(cell), base = vox_off_[cell];
    synchronize();
    std::vector<SkewVoxel> all;
    skew_->read(static_cast<hipStream_t>(stream_), all);
    out.clear();
    for (SkewVoxel e : all)
        if (e.v >= base && e.v < base + nv) {
            e.v -= uint32_t(base);
            out.push_back(e);
        }
}

void Ensemble::read_cme_counts(uint32_t cell, int32_t *out)
{
    if (cell >= n_cells())
        throw Error(CS_ERR_ARG, "cell index out of range");
    synchronize();
    if (cell_nc_[cell])
        hip_check(hipMemcpy(out, cme_counts_ + cme_counts_off_[cell], cell_nc_[cell] * sizeof(int32_t),
                            hipMemcpyDeviceToHost),
                  "read cell cme counts");
}

void Ensemble::read_counters(uint64_t *out)
{
    synchronize();
    hip_check(hipMemcpy(out, counters_, kNumCounters * sizeof(uint64_t), hipMemcpyDeviceToHost), "read counters");
}

void Ensemble::reset_counters()
{
    synchronize();
    hip_check(hipMemset(counters_, 0, kNumCounters * sizeof(unsigned long long)), "clear counters");
}

}  // namespace cs
