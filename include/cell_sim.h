/*
 * cell_sim.h -- C ABI of the MI355X RDME step (libcellsim.so).
 *
 * The entry points mirror the reference crate's `Simulation` API (src/simulation.rs) and `Setup`
 * (src/setup.rs); each declaration names the reference item it replaces. Arguments are plain pointers
 * and sizes. Every function returns CS_OK (0) or an error code; cs_last_error() gives the message of
 * the last failure on the calling thread. Nothing aborts the process.
 *
 * Threading: a handle is not thread-safe. Each handle owns one HIP stream; cs_sim_step/cs_sim_run
 * enqueue work on it asynchronously, and every read synchronises that stream first.
 *
 * Device state: one byte per particle slot (species id, 0 = empty), slots of a voxel contiguous,
 * voxel index v = z + y*R2 + x*R2*R1 (res/shader/functions.wgsl:42-51); one byte of region label per
 * voxel. Occupancy is the number of non-empty slots (slots are kept compacted, as
 * res/shader/final_stage.wgsl:25-33 leaves them).
 */
#ifndef CELL_SIM_H
#define CELL_SIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------------- */
#define CS_OK 0
#define CS_ERR_ARG 1          /* bad argument (null pointer, size mismatch, out of range) */
#define CS_ERR_NOT_FOUND 2    /* unknown region or particle name (the reference panics) */
#define CS_ERR_STATE 3        /* call out of order (e.g. step before cs_sim_prepare_device) */
#define CS_ERR_HIP 4          /* HIP runtime failure */
#define CS_ERR_UNSUPPORTED 5  /* model feature the HIP path does not run */
#define CS_ERR_FULL 6         /* placement could not find a free slot (lattice.rs:93-95 panics) */
#define CS_ERR_PARSE 7        /* malformed JSON model */

/* Message of the last failing call on this thread ("" if none). */
const char *cs_last_error(void);

/* ---- construction ------------------------------------------------------------------------------ */
typedef struct cs_sim cs_sim;

/* LatticeParams::new(dimensions, resolution, tau, lambda) (src/lattice_params.rs:61-81), plus the
 * slot count that the reference fixes at compile time (MAX_PARTICLES_SITE, src/lib.rs:15). */
typedef struct {
    float dims[3];               /* physical box size (m) */
    uint32_t res[3];             /* voxels along x, y, z */
    float tau;                   /* time step (s) */
    float lambda;                /* voxel edge used by the hop probability (m) */
    uint32_t max_particles_site; /* slots per voxel: 4, 8 or 16 */
} cs_lattice_params;

/* Simulation::new (src/simulation.rs:51-97): one background region (index 0). */
int cs_sim_create(const cs_lattice_params *params, cs_sim **out);
/* Simulation::from_file (src/simulation.rs:210-232) on a JSON file / string. `seed` drives particle
 * placement (the reference uses the unseeded rand::thread_rng, src/lattice.rs:65). `max_particles_site`
 * 0 means 8, the reference's constant. */
int cs_sim_from_json_file(const char *path, uint64_t seed, uint32_t max_particles_site, cs_sim **out);
int cs_sim_from_json_string(const char *json, uint64_t seed, uint32_t max_particles_site, cs_sim **out);
void cs_sim_destroy(cs_sim *sim);

/* Placement RNG seed for the following add_particle_* calls (no reference counterpart: see above). */
int cs_sim_set_seed(cs_sim *sim, uint64_t seed);

/* ---- regions: Simulation::add_region(RegionType, diffusion_rate) (src/simulation.rs:428-478) --- */
int cs_sim_add_region_cube(cs_sim *sim, const char *name, const float p0[3], const float pf[3],
                           float diffusion_rate);                        /* :627-658 */
int cs_sim_add_region_sphere(cs_sim *sim, const char *name, const float center[3], float radius,
                             float diffusion_rate);                      /* :660-699 */
int cs_sim_add_region_cylinder(cs_sim *sim, const char *name, const float p0[3], const float pf[3],
                               float radius, float diffusion_rate);      /* :550-625 */
int cs_sim_add_region_semisphere(cs_sim *sim, const char *name, const float center[3], float radius,
                                 const float direction[3], float diffusion_rate); /* :701-743 */
int cs_sim_add_region_spherical_shell(cs_sim *sim, const char *shell_name, const char *interior_name,
                                      const float center[3], float internal_radius,
                                      float external_radius, float diffusion_rate); /* :439-443 */
int cs_sim_add_region_cylindrical_shell(cs_sim *sim, const char *shell_name, const char *interior_name,
                                        const float p0[3], const float pf[3], float internal_radius,
                                        float external_radius, float diffusion_rate); /* :444-448 */
int cs_sim_add_region_capsid(cs_sim *sim, const char *shell_name, const char *interior_name,
                             const float center[3], const float dir[3], float internal_radius,
                             float external_radius, float total_length, float diffusion_rate); /* :449-476 */
/* Simulation::add_sparse_region (src/simulation.rs:483-535): spheres of `radius` (the base Sphere region;
 * its center is not used) drawn wholly inside to_region (src/region.rs:150-194) become region `name` until it
 * holds at least max_volume voxels. max_volume must be below to_region's volume. Placement draws from the
 * simulation's seeded generator. */
int cs_sim_add_sparse_region(cs_sim *sim, const char *name, float radius, const char *to_region,
                             uint32_t max_volume, float diffusion_rate);
/* Simulation::prepare_regions (src/simulation.rs:424; src/region.rs:129-148). */
int cs_sim_prepare_regions(cs_sim *sim);
/* find_region_index (src/simulation.rs:745-756); returns the index or -1. */
int cs_sim_region_index(const cs_sim *sim, const char *name);

/* ---- particles (src/simulation.rs:792-864) --------------------------------------------------------- */
int cs_sim_add_particle_count(cs_sim *sim, const char *name, const char *to_region, uint32_t count,
                              int logging, int is_reservoir);            /* :792-814 */
int cs_sim_add_particle_concentration(cs_sim *sim, const char *name, const char *to_region,
                                      float concentration, int logging, int is_reservoir); /* :816-823 */
int cs_sim_fill_region(cs_sim *sim, const char *name, const char *to_region, int logging); /* :825-840 */
/* Simulation::particle_random_walk (src/simulation.rs:842-864; src/lattice.rs:113-278): a new species laid
 * as a chain of cylinders (length block_length, radius `radius`, random directions) of total_length inside
 * to_region, stepping back 25 blocks when a block leaves the region or meets a full voxel. Lengths in the
 * lattice's units. A walk that fails 100000 blocks returns CS_ERR_FULL (the reference retries forever). */
int cs_sim_particle_random_walk(cs_sim *sim, const char *name, const char *to_region, float total_length,
                                float block_length, float radius, int logging);
/* Lattice::find_particle (src/lattice.rs:315-322); returns the species index (>= 1) or -1. */
int cs_sim_particle_index(const cs_sim *sim, const char *name);

/* ---- diffusion matrix (src/simulation.rs:867-902) -------------------------------------------------- */
int cs_sim_set_diffusion_rate(cs_sim *sim, const char *region, float rate);                /* :867-875 */
int cs_sim_set_diffusion_rate_particle(cs_sim *sim, const char *particle, const char *region,
                                       float rate);                                        /* :877-883 */
int cs_sim_set_transition_rate(cs_sim *sim, const char *from_region, const char *to_region,
                               float rate);                                                /* :885-894 */
int cs_sim_set_transition_rate_particle(cs_sim *sim, const char *particle, const char *from_region,
                                        const char *to_region, float rate);                /* :896-902 */

/* ---- reactions: Simulation::add_reaction (src/simulation.rs:904-947) ---------------------------- */
int cs_sim_add_reaction(cs_sim *sim, const char *const *reactants, uint32_t n_reactants,
                        const char *const *products, uint32_t n_products, float k);

/* ---- well-mixed (CME) species: NO reference counterpart ---------------------------------------------
 * The reference has only lattice (RDME) species. These entry points add a well-mixed pool per
 * simulation, advanced once per step, before the lattice step, by an exact SSA (Gillespie's direct
 * method) over tau: propensities k*binomial(counts) in f32 with Kahan-compensated cumulative sums,
 * simulated time and the total-propensity scale in f64. Reactants are well-mixed species; a product is a
 * well-mixed species ("name") or a lattice particle emitted into a region ("name@region"), placed at a
 * random voxel of that region that has a free slot (after the voxel's particles). Counts are per
 * simulation; rates in 1/s (per-molecule for first order). For models that need gene switching and
 * transcription into a compartment (a gene_on -> gene_on + mRNA@nucleoid row). */
int cs_sim_add_cme_species(cs_sim *sim, const char *name, int32_t count);
int cs_sim_add_cme_reaction(cs_sim *sim, const char *const *reactants, uint32_t n_reactants,
                            const char *const *products, uint32_t n_products, float k);
/* Seed of the CME draws (cs_sim_set_seed also derives one). */
int cs_sim_set_cme_seed(cs_sim *sim, uint32_t seed);
/* Current well-mixed counts, int32[num_cme_species]. */
int cs_sim_read_cme_counts(cs_sim *sim, int32_t *out, size_t n);
int cs_sim_cme_species_name(const cs_sim *sim, uint32_t index, char *out, size_t cap);
/* Precision audit: when on, every CME event also evaluates the propensities and the choice in plain f64;
 * the largest relative error of the f32 compensated total, the number of events whose f64 choice differs
 * and the number of audited events accumulate until read. Off by default. */
/* Well-mixed tables: reactants i32[ncr][3] (-1 = none), count changes i32[ncr][nc], lattice products
 * u8[ncr][4] (0 = none) and their region i32[ncr] (-1 = none), rates f32[ncr]; and the draw seed. */
int cs_sim_read_cme_tables(const cs_sim *sim, int32_t *reactants, size_t n_reactants, int32_t *delta,
                           size_t n_delta, uint8_t *emit, size_t n_emit, int32_t *emit_region,
                           size_t n_region, float *rates, size_t n_rates, uint32_t *seed);
int cs_sim_set_cme_audit(cs_sim *sim, int on);
int cs_sim_read_cme_audit(cs_sim *sim, double *max_rel_error, uint64_t *divergences, uint64_t *events);

/* ---- device: Setup::new_nowindow (src/setup.rs:94-151) + Simulation::prepare_for_gpu
 *      (src/simulation.rs:152-208): uploads lattice, regions and tables to `hip_device`. ---------- */
int cs_sim_prepare_device(cs_sim *sim, int hip_device);

/* ---- stepping: Simulation::step (src/simulation.rs:99-150) ----------------------------------------
 * `itime` is the RNG seed the reference takes from the wall clock (examples/no_window.rs:127-131).
 * When frame_num % write_freq == 0 and frame_num > 0, the logged species' global counts of the state
 * before this step are queued as statistics samples (src/simulation.rs:146-148, 391-418). */
int cs_sim_step(cs_sim *sim, uint32_t frame_num, uint32_t itime, uint32_t write_freq);
/* nsteps calls of cs_sim_step with frames first_frame.. and seeds itime[0..nsteps). */
int cs_sim_run(cs_sim *sim, uint32_t first_frame, const uint32_t *itime, uint32_t nsteps,
               uint32_t write_freq);
int cs_sim_synchronize(cs_sim *sim);
/* cs_sim_run without statistics sampling, with HIP events recorded on the simulation's stream around
 * every step kernel; `kernel_ms` receives their summed duration. Synchronises. No reference
 * counterpart: the reference has no timestamp queries (it prints FPS, examples/no_window.rs:151-162). */
int cs_sim_run_timed(cs_sim *sim, const uint32_t *itime, uint32_t nsteps, float *kernel_ms);

/* ---- inspection ---------------------------------------------------------------------------------- */
typedef struct {
    uint32_t res[3];
    uint64_t n_voxels;
    uint32_t max_particles_site;
    uint32_t num_species;        /* excluding void */
    uint32_t num_reactions;      /* excluding the dummy row 0 */
    uint32_t num_regions;
    float tau;
    float lambda;
    float dims[3];
    int device_ready;
    uint32_t num_cme_species;    /* well-mixed species (no reference counterpart) */
    uint32_t num_cme_reactions;
} cs_sim_info;
int cs_sim_get_info(const cs_sim *sim, cs_sim_info *out);

/* Current lattice (N*S bytes), region labels (N bytes). Before cs_sim_prepare_device these return the
 * host model; afterwards the device state. */
int cs_sim_read_lattice(cs_sim *sim, uint8_t *out, size_t nbytes);
int cs_sim_write_lattice(cs_sim *sim, const uint8_t *in, size_t nbytes);
int cs_sim_read_regions(const cs_sim *sim, uint8_t *out, size_t nbytes);
/* Global species counts, int32[ns+1] (the reference's concentrations_stat; entry 0 is 0). */
int cs_sim_read_species_counts(cs_sim *sim, int32_t *out, size_t n);
/* Per-region species counts, int32[nreg][ns+1] (no reference counterpart: the reference keeps only
 * global counts). */
int cs_sim_read_region_counts(cs_sim *sim, int32_t *out, size_t n);
/* cs_sim_read_region_counts without allocating anything. No counterpart in the reference: for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * histogram's bins live in one fixed device array of the library, 16384 int32 per GPU, held until the process
 * ends; a model with more than 16384 regions x (species + 1) bins gets CS_ERR_ARG. Every call on one GPU uses
 * that GPU's array, so calls that overlap on one GPU race. */
int cs_sim_read_region_counts_noalloc(cs_sim *sim, int32_t *out, size_t n);

/* Model tables in the reference's device layout: diffusion f32[nreg][nreg][ns+1] (src,dst,species),
 * stoichiometry i32[nr+1][ns+1], reactant indices i32[nr+1][3], rates f32[nr+1]. */
int cs_sim_read_tables(const cs_sim *sim, float *diffusion, size_t n_diffusion, int32_t *stoich,
                       size_t n_stoich, int32_t *reactants, size_t n_reactants, float *rates,
                       size_t n_rates);
/* Reservoir species per voxel, u32[N] (src/lattice.rs:292-300). */
int cs_sim_read_reservoir(const cs_sim *sim, uint32_t *out, size_t n);
/* Per-voxel species counts, int32[N][ns+1]: the reference's `concentrations` buffer (src/lattice.rs:14-23,
 * updated by rdme.wgsl:66,71 and cme.wgsl:84,104). They equal the slot contents except at reservoir
 * voxels; entry 0 of a voxel is 1 once a step has run (cme.wgsl:31), 0 before. */
int cs_sim_read_concentrations(cs_sim *sim, int32_t *out, size_t n);
/* Per-voxel occupancy word, u32[N]: the reference's `occupancy` buffer (src/lattice.rs:14-23, rdme.wgsl:61-80,
 * cme.wgsl:114). It equals the occupied slot count except at reservoir voxels, where the reaction's
 * skipped species leave it off by a signed amount (kept modulo 2^32, as the reference's u32 atomics). */
int cs_sim_read_occupancy(cs_sim *sim, uint32_t *out, size_t n);
/* The renderer's view of the current state (res/shader/final_stage.wgsl:35-67, written there into an r32float
 * 3D texture every step), computed on request: f32[N] in voxel order (v = z + y*R2 + x*R2*R1). view 0 = the
 * majority species of the voxel / 255, 1 = region label / 255, 2 = reservoir species / 255
 * (Uniforms.rendering_view); voxels with res[slice_axis] - coord[slice_axis] < slice are 0 (Uniforms.slice,
 * slice_axis). Needs the device (CS_ERR_STATE before cs_sim_prepare_device). */
int cs_sim_render_view(cs_sim *sim, uint32_t view, uint32_t slice_axis, uint32_t slice, float *out, size_t n);

/* A statistics sample (src/statistics.rs:4-9 SolverStatisticSample). Returns CS_OK and fills the
 * sample, or CS_ERR_NOT_FOUND when the queue is empty. */
typedef struct {
    char name[64];
    int32_t value;
    uint32_t iteration_count;
} cs_stat_sample;
int cs_sim_pop_stat(cs_sim *sim, cs_stat_sample *out);

/* Step diagnostics accumulated on the device since prepare (uint64[CS_NUM_COUNTERS]). */
#define CS_CNT_TIGHT 0           /* voxel-steps whose capacity could be reached (replayed in canonical order) */
#define CS_CNT_EXCEPTION 1       /* voxel-steps not followed exactly: replay scratch or skew table exhausted, more than
                                    4 species with count offsets in one voxel, or an occupancy word wrapped below
                                    zero at the first voxels (whose sweep walks the whole lattice) */
#define CS_CNT_NEAR_THRESHOLD 2  /* voxel-steps with |r1 - (1-exp(-a0*tau))| <= 2^-24 */
#define CS_CNT_FIRED 3           /* voxel-steps where a reaction fired */
#define CS_CNT_ROLLBACK 4        /* fired reactions undone because S slots would fill */
#define CS_CNT_CME_EVENTS 5      /* well-mixed reactions fired */
#define CS_CNT_CME_NEAR 6        /* CME steps whose last waiting time ended within 1e-12*tau of tau */
#define CS_CNT_EMIT_FAIL 7       /* lattice products of CME reactions dropped: no free slot in 100 draws */
#define CS_CNT_FAILED_CLAIM 8    /* hops rejected because the destination was full (rdme.wgsl:62, 73-81) */
#define CS_CNT_ALIAS 9           /* claims that landed in slot S, the next voxel's slot 0 (rdme.wgsl:30-40) */
#define CS_NUM_COUNTERS 10
int cs_sim_read_counters(cs_sim *sim, uint64_t *out, size_t n);
int cs_sim_reset_counters(cs_sim *sim);

/* Name of species `index` (0 = "void"); copies at most `cap` bytes including the terminator. */
int cs_sim_particle_name(const cs_sim *sim, uint32_t index, char *out, size_t cap);

/* ---- batched ensembles: NO reference counterpart -----------------------------------------------------
 * The reference steps one Simulation per dispatch chain. An ensemble holds many independent cells
 * (lattices of any size with their own tables, regions and well-mixed pools; one slot count) in one
 * device state and steps them together: per step, a prologue launch derives each cell's itime from a
 * device-side step counter and runs the well-mixed phases, one persistent launch takes (cell, tile,
 * x-chunk) items from a device-side queue, and an epilogue replays tight voxels and re-arms the queue.
 * Blocks of steps are captured as a hipGraph and replayed. Cell c's itime at step t is the low 32 bits of
 * SplitMix64((itime_seeds[c] << 32) + t) -- what default_itime_schedule(itime_seeds[c]) gives a plain
 * simulation -- so every cell evolves exactly as that simulation stepped alone would. The cells are
 * copied from the handles' host models (their state as built); the handles stay independent. */
typedef struct cs_ens cs_ens;
typedef struct {
    uint32_t xc;             /* x planes per work item */
    uint32_t blocks_per_cu;  /* persistent blocks per compute unit */
    uint32_t batch;          /* items taken per queue access; 0 = fixed round-robin without the queue */
} cs_ens_tiling;
typedef struct {
    uint32_t n_cells;
    uint32_t max_particles_site;
    uint64_t total_voxels;
    uint64_t steps_done;
    uint32_t items;          /* work items per step under the current tiling */
} cs_ens_info;
int cs_ens_create(cs_sim *const *cells, uint32_t n_cells, const uint32_t *itime_seeds, int hip_device,
                  cs_ens **out);
void cs_ens_destroy(cs_ens *ens);
/* nsteps steps with launches per step / through a graph of `block` steps (rounded up to even). */
int cs_ens_run(cs_ens *ens, uint32_t nsteps);
int cs_ens_run_graph(cs_ens *ens, uint32_t nsteps, uint32_t block);
/* As cs_ens_run_graph (block >= 2) or cs_ens_run (block < 2), timed with events on the ensemble's
 * stream (graph capture outside the timed window). Synchronises. */
int cs_ens_run_timed(cs_ens *ens, uint32_t nsteps, uint32_t block, float *ms);
int cs_ens_synchronize(cs_ens *ens);
int cs_ens_set_tiling(cs_ens *ens, const cs_ens_tiling *t);
int cs_ens_get_tiling(const cs_ens *ens, cs_ens_tiling *t);
/* Tries tilings (planes per item 8/16/32/48/60, 2-4 blocks per CU, queue batch 0/1/4) for `trial_steps`
 * steps each from a snapshot of the state (restored afterwards) and keeps the fastest. */
int cs_ens_autotune(cs_ens *ens, uint32_t trial_steps, float *ms_per_step);
int cs_ens_get_info(const cs_ens *ens, cs_ens_info *out);
int cs_ens_cell_res(const cs_ens *ens, uint32_t cell, uint32_t res[3]);
int cs_ens_read_lattice(cs_ens *ens, uint32_t cell, uint8_t *out, size_t nbytes);
int cs_ens_read_cme_counts(cs_ens *ens, uint32_t cell, int32_t *out, size_t n);
/* One cell's per-voxel counts int32[N][ns+1] and occupancy words uint32[N], as cs_sim_read_concentrations and
 * cs_sim_read_occupancy give them for a simulation. */
int cs_ens_read_concentrations(cs_ens *ens, uint32_t cell, int32_t *out, size_t n);
int cs_ens_read_occupancy(cs_ens *ens, uint32_t cell, uint32_t *out, size_t n);
int cs_ens_read_counters(cs_ens *ens, uint64_t *out, size_t n);
int cs_ens_reset_counters(cs_ens *ens);

#ifdef __cplusplus
}
#endif

#endif /* CELL_SIM_H */
