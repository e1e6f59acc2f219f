/*
 * cs_oracle.h -- CPU restatement of cell_simulation's RDME step (TEST INFRASTRUCTURE ONLY).
 *
 * This is the parity checker. It is loaded only by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg. The product (cell_simulation_amd) never links or calls it.
 *
 * The step is executed in the canonical serial order: voxels in ascending linear index
 * v = z + y*R2 + x*R2*R1, particle slots in ascending order. The reference kernels race (spin locks
 * and first-fit CAS, res/shader/rdme.wgsl:25-88); this order is the serialisation both the oracle and
 * the HIP path are defined against.
 *
 * Followed files (reference, read-only):
 *   res/shader/random.wgsl:14-49        PCG hash, xorshift32, u32 -> f32
 *   res/shader/functions.wgsl:42-51     voxel / lattice index math
 *   res/shader/rdme.wgsl:25-229         diffusion sweep (lock + first-fit claim)
 *   res/shader/cme.wgsl:25-120          per-voxel reaction sampling, slot rewrite, rollback
 *   res/shader/final_stage.wgsl:25-33   stable slot compaction
 *   src/simulation.rs:99-150            pass order rdme -> cme -> final_stage -> copy dest->src
 */
#ifndef CS_ORACLE_H
#define CS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t res[3];          /* lattice resolution (x, y, z); x is the slowest index */
    uint32_t S;               /* particle slots per voxel (MAX_PARTICLES_SITE), 1..16 */
    uint32_t ns;              /* species count excluding void (species 0) */
    uint32_t nr;              /* reaction count excluding the dummy row 0 */
    uint32_t nreg;            /* region count */
    float tau;                /* time step (s) */
    float lambda;             /* voxel edge (m) */
    uint32_t mod_zero;        /* value taken for X % 0 in the RDME seed (hazard H4) */
    const float *diffusion;   /* [nreg][nreg][ns+1]: index sp + (ns+1)*dst + (ns+1)*nreg*src */
    const int32_t *stoich;    /* [nr+1][ns+1] */
    const int32_t *reactants; /* [nr+1][3], padded with species 0 */
    const float *rates;       /* [nr+1] */
    const uint32_t *regions;  /* [N] region label per voxel */
    const uint32_t *reservoir;/* [N] reservoir species per voxel (0 = none); may be NULL */
} cso_model;

typedef struct {
    uint32_t *lattice;        /* [N*S] species id per slot */
    uint32_t *occupancy;      /* [N] */
    int32_t *counts;          /* [N][ns+1] per-voxel species counts */
    int32_t *stat;            /* [ns+1] global species counts (concentrations_stat) */
} cso_state;

typedef struct {
    uint64_t moves;           /* successful hops */
    uint64_t failed_claims;   /* hops rejected because the destination was full */
    uint64_t alias_writes;    /* first-fit landed in slot S, i.e. the next voxel's slot 0 (hazard H2) */
    uint64_t fired;           /* voxels where a reaction fired (including rollbacks) */
    uint64_t rollbacks;       /* fired reactions undone because the rewrite reached S slots (H8) */
    uint64_t near_threshold;  /* |r1 - (1-exp(-a0*tau))| <= 2^-24: an exp ulp could flip the outcome */
} cso_counters;

/* Random number primitives (random.wgsl:14-20, 32-45, 27-29). */
uint32_t cso_pcg(uint32_t state);
uint32_t cso_xorshift(uint32_t state);
float cso_uint_to_float(uint32_t u);
float cso_uniform_float(uint32_t state);
/* The build's definition of the f32 exponential: exp evaluated in double, rounded once to f32. */
float cso_expf(float x);

/* One full step (phase R, phase C, phase F). Returns 0 on success, nonzero on bad arguments. */
int cso_step(const cso_model *m, cso_state *st, uint32_t itime, cso_counters *cnt);

/* The phases separately, for tests. Phase R needs copies of the source lattice and occupancy. */
void cso_phase_r(const cso_model *m, const uint32_t *lat_src, const uint32_t *occ_src,
                 cso_state *st, uint32_t itime, cso_counters *cnt);
void cso_phase_c(const cso_model *m, cso_state *st, uint32_t itime, cso_counters *cnt);
void cso_phase_f(const cso_model *m, cso_state *st);

/* Probability vector of rdme.wgsl:138-191 for one particle, exposed for tests. */
void cso_probability_vector(const cso_model *m, uint32_t x, uint32_t y, uint32_t z, uint32_t sp,
                            float p[7]);

/* Multi-step driver: steps t = 0..nsteps-1 use itime[t]. */
int cso_run(const cso_model *m, cso_state *st, const uint32_t *itime, uint32_t nsteps, cso_counters *cnt);

/* cso_run on `nthreads` host threads (the CPU baseline): hop draws, phase C and phase F in parallel over
 * x-planes, the claims of phase R applied serially in the canonical order. Same result as cso_run. */
int cso_run_mt(const cso_model *m, cso_state *st, const uint32_t *itime, uint32_t nsteps, cso_counters *cnt,
               uint32_t nthreads);

/* final_stage.wgsl:35-67: the r32float view of the current state, one value per voxel in voxel order
 * (out[v], v = z + y*R2 + x*R2*R1). Voxels with res[axis] - coord[axis] < slice are 0. view 0: the
 * species found most often among slots 0..occupancy-1 (lattice index order, so slots past S are the next
 * voxel's; ties keep the first in slot order; an index past the lattice reads 0), divided by 255; view 1:
 * region / 255; view 2: reservoir species / 255. An occupancy word above 2*S (only a word wrapped below
 * zero) gives 0: the reference's loop would walk the whole buffer. */
void cso_view(const cso_model *m, const cso_state *st, uint32_t view, uint32_t axis, uint32_t slice, float *out);

/* ---- well-mixed (CME) phase: the build's own extension, NO reference counterpart -------------------
 * Once per step, before the lattice step, the well-mixed counts advance by an exact SSA (Gillespie's
 * direct method) over tau. The arithmetic below is the definition the HIP path follows bit for bit:
 *   h_r = product over distinct reactant species of binomial(count, multiplicity), in double;
 *   a_r = k_r * (float)h_r (f32); cumulative sums in f32 with Kahan compensation, read in double as
 *   (double)sum - (double)compensation; a0 = that total;
 *   draw k: bits = ((u64)pcg(itime + pcg(seed + 2k)) << 32 | pcg(itime + pcg(seed + 2k + 1))) >> 11;
 *   event e: u1 = (bits(2e) + 1) * 2^-53, t += (-log(u1)) / a0, stop when t >= tau;
 *            u2 = bits(2e + 1) * 2^-53, choice = first r with a_r > 0 and u2 * a0 < cumulative_r
 *            (the last r with a_r > 0 if rounding leaves none);
 *   lattice product j of event e goes to region voxel list[(u64)u * len >> 32] with
 *            u = pcg(itime + pcg((seed ^ 0x9E3779B9) + ((4e + j) << 7) + attempt)), attempts 0..99,
 *            appended after the voxel's particles (slot = particles held, its slots being compacted) when it
 *            has a free slot; the voxel's occupancy word and count of that species go up by one. */
typedef struct {
    uint32_t nc, ncr;                      /* well-mixed species, reactions */
    const int32_t *reactants;              /* [ncr][3], -1 = none */
    const int32_t *delta;                  /* [ncr][nc] */
    const uint8_t *emit;                   /* [ncr][4] lattice species, 0 = none */
    const int32_t *emit_region;            /* [ncr], -1 = none */
    const float *rates;                    /* [ncr] */
    const uint32_t *const *region_voxels;  /* [nreg] ascending voxel lists (NULL where unused) */
    const uint32_t *region_len;            /* [nreg] */
    uint32_t seed;
} cso_cme_model;

typedef struct {
    uint64_t events;      /* reactions fired */
    uint64_t near;        /* steps whose last waiting time ended within 1e-12*tau of tau */
    uint64_t emit_fail;   /* lattice products dropped (no free slot in 100 draws) */
    uint64_t exceptions;  /* steps stopped at the event cap */
} cso_cme_counters;

/* Combinatorial factor h_r and the 53-bit draw, as defined above. */
double cso_cme_h(const int32_t *row3, const int32_t *counts);
uint64_t cso_cme_bits(uint32_t itime, uint32_t seed, uint32_t k);
/* The CME phase on the oracle state (counts[nc] updated in place; placements update lattice,
 * occupancy, per-voxel counts and stat). */
void cso_cme_phase(const cso_model *m, const cso_cme_model *c, cso_state *st, int32_t *counts, uint32_t itime,
                   cso_cme_counters *cnt);
/* Plain f64 reference of one event's total propensity and choice for counts and u2, for precision reports. */
double cso_cme_total64(const cso_cme_model *c, const int32_t *counts);
uint32_t cso_cme_choice64(const cso_cme_model *c, const int32_t *counts, double u2);
/* The f32-compensated total and choice (the definition above), for the same comparison. */
double cso_cme_total32(const cso_cme_model *c, const int32_t *counts);
uint32_t cso_cme_choice32(const cso_cme_model *c, const int32_t *counts, double u2);
/* A full hybrid step: CME phase, then cso_step. */
int cso_step_hybrid(const cso_model *m, const cso_cme_model *c, cso_state *st, int32_t *counts, uint32_t itime,
                    cso_counters *cnt, cso_cme_counters *ccnt);

#ifdef __cplusplus
}
#endif

#endif
