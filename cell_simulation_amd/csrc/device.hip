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

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cell_sim.h"
#include "cs_rng.h"
#include "device.hpp"

namespace cs {

namespace {

constexpr int kMaxRW = 4;  // reaction bitmask words: up to 255 reactions
constexpr int kBlock = 256;

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw Error(CS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
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
    unsigned long long *counters;
    uint32_t R0, R1, R2, R12;
    uint32_t N;
    uint32_t ns, nr, nreg, nrw;
    uint32_t itime;
    uint32_t mod_zero;
    float tau;
};

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

// number of slots holding species sp (sp >= 1)
template <int S>
__device__ __forceinline__ uint32_t count_species(const Slots<S> &s, uint32_t sp)
{
    const uint32_t pat = sp * 0x01010101u;
    uint32_t n = 0;
#pragma unroll
    for (int q = 0; q < S / 4; ++q)
        n += __builtin_popcount(nonzero_bytes(s.w[q] ^ pat) ^ 0x80808080u);
    return n;
}

template <int S>
__device__ __forceinline__ uint32_t get_byte(const Slots<S> &s, int i)
{
    return (s.w[i >> 2] >> ((i & 3) * 8)) & 0xFFu;
}

// byte at a run-time position j < S
template <int S>
__device__ __forceinline__ uint32_t get_byte_dyn(const Slots<S> &s, uint32_t j)
{
    uint32_t w = s.w[0];
#pragma unroll
    for (int q = 1; q < S / 4; ++q)
        w = ((j >> 2) == uint32_t(q)) ? s.w[q] : w;
    return (w >> ((j & 3u) * 8u)) & 0xFFu;
}

// OR a byte into a zero slot at run-time position j; positions >= S are dropped
template <int S>
__device__ __forceinline__ void put_byte_dyn(Slots<S> &s, uint32_t j, uint32_t val)
{
    const uint32_t sh = (j & 3u) * 8u;
#pragma unroll
    for (int q = 0; q < S / 4; ++q)
        s.w[q] |= ((j >> 2) == uint32_t(q)) ? (val << sh) : 0u;
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

// Direction 0..5 of an existing neighbour the particle hops to, or 6 if it stays (stay draw or
// reflecting boundary). The cumulative sum is the reference's p[0..5] in f32, in the same order.
__device__ __forceinline__ uint32_t hop(const StepArgs &a, const Hood &h, uint32_t sp, uint32_t slot)
{
    const float r = uniform_float(pcg_hash(h.seed0 + slot));
    const uint32_t n1 = a.ns + 1u;
    const float *row = a.p2 + size_t(n1) * a.nreg * h.src_reg + sp;
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (h.inside & (1u << d))
            acc = acc + row[size_t(n1) * h.rg[d]];
        if (!(r > acc))
            return ((h.inside >> d) & 1u) ? uint32_t(d) : 6u;
    }
    return 6u;
}

// Append the particles of neighbour u that hop in direction `want` (toward the voxel being gathered),
// in source-slot order. Returns the number found (may exceed the list capacity S).
template <int S>
__device__ __forceinline__ uint32_t gather_from(const StepArgs &a, uint32_t ux, uint32_t uy, uint32_t uz,
                                                uint32_t u, uint32_t want, Slots<S> &list, uint32_t n)
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
            if (hop(a, h, sp, s) == want) {
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
__device__ __forceinline__ uint32_t species_count(const Slots<S> &s, uint32_t sp)
{
    return sp == 0u ? 1u : count_species(s, sp);  // C[v][0] = 1 (cme.wgsl:31)
}

template <int S>
__device__ __forceinline__ float propensity(const StepArgs &a, const Slots<S> &s, uint32_t r)
{
    const uchar4 ix = *reinterpret_cast<const uchar4 *>(a.ridx + 4u * r);
    const uint32_t prod = species_count(s, ix.x) * species_count(s, ix.y) * species_count(s, ix.z);
    return a.rates[r] * float(int32_t(prod));
}

// Applies the reaction step to `out` (cnt particles, compacted). Reactions with a reactant absent from
// the voxel have propensity +0 and leave every partial sum unchanged, so only candidate reactions
// (zero-order ones and those with a reactant present) are visited; the f32 sums match the reference's
// sequential sums bit for bit.
template <int S>
__device__ void react(const StepArgs &a, Slots<S> &out, uint32_t cnt, uint32_t x, uint32_t y, uint32_t z,
                      bool &fired, bool &rolled, bool &near, bool &bad)
{
    fired = rolled = near = bad = false;
    const uint32_t P = pcg_hash(a.itime + x * y + z);
    const float r1 = uniform_float(P);

    uint64_t cand[kMaxRW];
#pragma unroll
    for (int w = 0; w < kMaxRW; ++w)
        cand[w] = (uint32_t(w) < a.nrw) ? a.rmask[w] : 0ull;
    for (uint32_t j = 0; j < cnt; ++j) {
        const uint32_t sp = get_byte_dyn(out, j);
#pragma unroll
        for (int w = 0; w < kMaxRW; ++w)
            if (uint32_t(w) < a.nrw)
                cand[w] |= a.rmask[sp * a.nrw + w];
    }

    float tot = 0.0f;
#pragma unroll
    for (int w = 0; w < kMaxRW; ++w) {
        uint64_t m = cand[w];
        while (m) {
            const uint32_t r = uint32_t(w) * 64u + uint32_t(__ffsll((long long)m) - 1);
            m &= m - 1ull;
            tot = tot + propensity(a, out, r);
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
            cum = cum + propensity(a, out, r);
            i = r;
            if (r >= a.nr || !(r2 > cum / tot))
                chosen = true;
        }
    }
    if (!chosen)
        i = a.nr;

    // Stoichiometry and slot rewrite in species order (cme.wgsl:76-119). The rewrite is undone when it
    // reaches S slots (cme.wgsl:95-108), leaving the pre-reaction slots.
    const int8_t *row = a.stoich + size_t(i) * (a.ns + 1u);
    Slots<S> res = zero_slots<S>();
    uint32_t j = 0;
    for (uint32_t sp = 1; sp <= a.ns; ++sp) {
        const int32_t c = int32_t(count_species(out, sp)) + int32_t(row[sp]);
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

// Canonical slot picture of the voxel after diffusion (see the header comment).
template <int S>
__device__ __forceinline__ Slots<S> assemble(const Slots<S> &own, uint32_t occ, uint64_t stay_mask,
                                             const Slots<S> &lo, uint32_t nlo, const Slots<S> &hi, uint32_t nhi,
                                             uint32_t &cnt)
{
    Slots<S> out = zero_slots<S>();
    uint32_t j = 0, h = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if (uint32_t(s) < occ) {
            if ((stay_mask >> s) & 1ull) {
                put_byte_dyn(out, j++, get_byte(own, s));
            } else if (h < nhi) {
                put_byte_dyn(out, j++, get_byte_dyn(hi, h));
                ++h;
            }
        }
    }
    for (uint32_t q = 0; q < nlo; ++q)
        put_byte_dyn(out, j++, get_byte_dyn(lo, q));
    for (; h < nhi; ++h)
        put_byte_dyn(out, j++, get_byte_dyn(hi, h));
    cnt = j;
    return out;
}

__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
{
    // Consecutive dispatch IDs go round-robin over the 8 XCDs; give each XCD a contiguous range of
    // lattice blocks so neighbouring planes share its L2.
    const uint32_t q = nb / 8u, rem = nb % 8u, xcd = b % 8u, idx = b / 8u;
    return xcd * q + min(xcd, rem) + idx;
}

template <int S>
__global__ __launch_bounds__(kBlock) void rdme_step_kernel(StepArgs a)
{
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t v = blk * kBlock + threadIdx.x;
    const bool live = v < a.N;
    bool tight = false, fired = false, rolled = false, near = false, bad = false;
    if (live) {
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
                if (uint32_t(s) < occ && hop(a, h, get_byte(own, s), s) == 6u)
                    stay |= 1ull << s;
        }
        Slots<S> lo = zero_slots<S>(), hi = zero_slots<S>();
        uint32_t nlo = 0, nhi = 0;
        if (x > 0u) nlo = gather_from<S>(a, x - 1u, y, z, v - a.R12, 1u, lo, nlo);      // from -x, hopped +x
        if (y > 0u) nlo = gather_from<S>(a, x, y - 1u, z, v - a.R2, 3u, lo, nlo);       // from -y, hopped +y
        if (z > 0u) nlo = gather_from<S>(a, x, y, z - 1u, v - 1u, 5u, lo, nlo);         // from -z, hopped +z
        if (z + 1u < a.R2) nhi = gather_from<S>(a, x, y, z + 1u, v + 1u, 4u, hi, nhi);  // from +z, hopped -z
        if (y + 1u < a.R1) nhi = gather_from<S>(a, x, y + 1u, z, v + a.R2, 2u, hi, nhi);  // from +y, hopped -y
        if (x + 1u < a.R0) nhi = gather_from<S>(a, x + 1u, y, z, v + a.R12, 0u, hi, nhi);  // from +x, hopped -x
        tight = occ + nlo + nhi > uint32_t(S);

        uint32_t cnt = 0;
        Slots<S> out = assemble<S>(own, occ, stay, lo, nlo, hi, nhi, cnt);
        react<S>(a, out, cnt, x, y, z, fired, rolled, near, bad);
        store_slots<S>(a.dst, v, out);
    }
    count_event(a.counters, CS_CNT_TIGHT, tight);
    count_event(a.counters, CS_CNT_EXCEPTION, bad);
    count_event(a.counters, CS_CNT_NEAR_THRESHOLD, near);
    count_event(a.counters, CS_CNT_FIRED, fired);
    count_event(a.counters, CS_CNT_ROLLBACK, rolled);
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

}  // namespace

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
    if (m.has_reservoir())
        throw Error(CS_ERR_UNSUPPORTED, "reservoir particles are not supported by the HIP path");
    if (nreg_ > 255)
        throw Error(CS_ERR_UNSUPPORTED, "at most 255 regions fit the one-byte label");
    if (nr_ > 255)
        throw Error(CS_ERR_UNSUPPORTED, "at most 255 reactions");
    if (N_ * S_ >= (uint64_t(1) << 32))
        throw Error(CS_ERR_UNSUPPORTED, "lattice too large for u32 slot indices");
    for (float k : m.rates())
        if (!std::isfinite(k))
            throw Error(CS_ERR_UNSUPPORTED, "reaction rates must be finite");
    for (int32_t st : m.stoich())
        if (st < -127 || st > 127)
            throw Error(CS_ERR_UNSUPPORTED, "stoichiometry entries must fit int8");
    nrw_ = (nr_ + 1u + 63u) / 64u;

    int ndev = 0;
    hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (hip_device < 0 || hip_device >= ndev)
        throw Error(CS_ERR_ARG, "no such HIP device");
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st;
    hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    stream_ = st;

    // Hop probabilities (D*tau)/(lambda*lambda) in f32, as rdme.wgsl:135 evaluates them.
    const uint32_t n1 = ns_ + 1u;
    std::vector<float> p2(m.diffusion().size());
    const float den = lp.lambda * lp.lambda;
    for (size_t i = 0; i < p2.size(); ++i) {
        const float num = m.diffusion()[i] * lp.tau;
        p2[i] = num / den;
    }
    std::vector<uint8_t> ridx(size_t(nr_ + 1u) * 4u, 0);
    std::vector<int8_t> st8(size_t(nr_ + 1u) * n1, 0);
    std::vector<uint64_t> rmask(size_t(n1) * nrw_, 0ull);
    for (uint32_t r = 0; r <= nr_; ++r) {
        bool any = false;
        for (int k = 0; k < 3; ++k) {
            const int32_t sp = m.reactants()[size_t(r) * 3u + k];
            ridx[size_t(r) * 4u + k] = uint8_t(sp);
            if (sp > 0) {
                any = true;
                rmask[size_t(sp) * nrw_ + r / 64u] |= 1ull << (r % 64u);
            }
        }
        if (!any && r > 0)
            rmask[r / 64u] |= 1ull << (r % 64u);  // zero-order: a candidate in every voxel
        for (uint32_t k = 0; k < n1; ++k)
            st8[size_t(r) * n1 + k] = int8_t(m.stoich()[size_t(r) * n1 + k]);
    }
    std::vector<uint8_t> reg8(N_);
    for (uint64_t v = 0; v < N_; ++v)
        reg8[v] = uint8_t(m.regions()[v]);

    hip_check(hipMalloc(&lat_[0], N_ * S_), "hipMalloc lattice");
    hip_check(hipMalloc(&lat_[1], N_ * S_), "hipMalloc lattice");
    hip_check(hipMalloc(&reg_, N_), "hipMalloc regions");
    hip_check(hipMalloc(&p2_, p2.size() * sizeof(float)), "hipMalloc p2");
    hip_check(hipMalloc(&rates_, m.rates().size() * sizeof(float)), "hipMalloc rates");
    hip_check(hipMalloc(&ridx_, ridx.size()), "hipMalloc ridx");
    hip_check(hipMalloc(&stoich_, st8.size()), "hipMalloc stoich");
    hip_check(hipMalloc(&rmask_, rmask.size() * sizeof(uint64_t)), "hipMalloc rmask");
    hip_check(hipMalloc(&counters_, kNumCounters * sizeof(unsigned long long)), "hipMalloc counters");
    hip_check(hipMalloc(&hist_, size_t(nreg_) * n1 * sizeof(int32_t)), "hipMalloc hist");

    hip_check(hipMemcpyAsync(lat_[0], m.lattice().data(), N_ * S_, hipMemcpyHostToDevice, st), "upload lattice");
    hip_check(hipMemcpyAsync(reg_, reg8.data(), N_, hipMemcpyHostToDevice, st), "upload regions");
    hip_check(hipMemcpyAsync(p2_, p2.data(), p2.size() * sizeof(float), hipMemcpyHostToDevice, st), "upload p2");
    hip_check(hipMemcpyAsync(rates_, m.rates().data(), m.rates().size() * sizeof(float), hipMemcpyHostToDevice, st),
              "upload rates");
    hip_check(hipMemcpyAsync(ridx_, ridx.data(), ridx.size(), hipMemcpyHostToDevice, st), "upload ridx");
    hip_check(hipMemcpyAsync(stoich_, st8.data(), st8.size(), hipMemcpyHostToDevice, st), "upload stoich");
    hip_check(hipMemcpyAsync(rmask_, rmask.data(), rmask.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st),
              "upload rmask");
    hip_check(hipMemsetAsync(counters_, 0, kNumCounters * sizeof(unsigned long long), st), "clear counters");
    hip_check(hipStreamSynchronize(st), "upload");
    cur_ = 0;
}

Device::~Device()
{
    hipSetDevice(dev_);
    if (stream_)
        hipStreamSynchronize(static_cast<hipStream_t>(stream_));
    hipFree(lat_[0]);
    hipFree(lat_[1]);
    hipFree(reg_);
    hipFree(p2_);
    hipFree(rates_);
    hipFree(ridx_);
    hipFree(stoich_);
    hipFree(rmask_);
    hipFree(counters_);
    hipFree(hist_);
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
    a.counters = counters_;
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
    const uint32_t blocks = uint32_t((N_ + kBlock - 1) / kBlock);
    hipStream_t st = static_cast<hipStream_t>(stream_);
    switch (S_) {
    case 4: hipLaunchKernelGGL(rdme_step_kernel<4>, dim3(blocks), dim3(kBlock), 0, st, a); break;
    case 8: hipLaunchKernelGGL(rdme_step_kernel<8>, dim3(blocks), dim3(kBlock), 0, st, a); break;
    case 16: hipLaunchKernelGGL(rdme_step_kernel<16>, dim3(blocks), dim3(kBlock), 0, st, a); break;
    default: throw Error(CS_ERR_UNSUPPORTED, "slots per voxel must be 4, 8 or 16");
    }
    hip_check(hipGetLastError(), "rdme_step_kernel launch");
    cur_ ^= 1;
}

float Device::run_timed(const uint32_t *itime, uint32_t nsteps)
{
    hip_check(hipSetDevice(dev_), "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream_);
    std::vector<hipEvent_t> ev(2 * size_t(nsteps));
    for (auto &e : ev)
        hip_check(hipEventCreate(&e), "hipEventCreate");
    for (uint32_t t = 0; t < nsteps; ++t) {
        hip_check(hipEventRecord(ev[2 * t], st), "hipEventRecord");
        step(itime[t]);
        hip_check(hipEventRecord(ev[2 * t + 1], st), "hipEventRecord");
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
    hip_check(hipStreamSynchronize(st), "write lattice");
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

}  // namespace cs
