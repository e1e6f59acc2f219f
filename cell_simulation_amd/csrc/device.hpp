// device.hpp -- device state of one simulation on one MI355X (gfx950) and the step launcher.
//
// Replaces the reference's GPU half: Simulation::prepare_for_gpu (src/simulation.rs:152-208),
// RDME::step / CME::step dispatches (src/rdme.rs:46-76, src/cme.rs:47-79), the final_stage compaction
// pass and the dest->src copies (src/simulation.rs:99-150), and the statistics readback
// (src/simulation.rs:361-418).
#pragma once

#include <cstdint>
#include <vector>

#include "model.hpp"

namespace cs {

constexpr int kNumCounters = 10;

// A voxel whose reference counters differ from its slots: the occupancy word is the occupied slots + off, the
// count of species sp[q] the slots holding it + dc[q] (sp[q] = 0: unused), and the reaction skips species K.
struct SkewVoxel {
    uint32_t v;       // voxel index (among all cells' voxels for an ensemble)
    int32_t off;
    uint32_t sp[4];
    int32_t dc[4];
    uint32_t K;
};

struct SkewStore;  // the device table of skewed voxels and the replay scratch (device.hip)
struct WaveItem;   // a work item of the wave kernel (device.hip)

class Device {
public:
    Device(const Model &m, int hip_device);
    ~Device();
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;

    // Enqueue one full step (diffusion, reaction, compaction) seeded with `itime`.
    void step(uint32_t itime);
    // Steps with an event pair around each kernel; returns the summed kernel milliseconds.
    float run_timed(const uint32_t *itime, uint32_t nsteps);
    void synchronize();

    void read_lattice(uint8_t *out);
    void write_lattice(const uint8_t *in);
    // Species counts per region, int32[nreg][ns+1], reduced on the device.
    void region_counts(int32_t *out);
    // The same counts with no allocation: the bins live in the library's fixed per-GPU array (calls that overlap
    // on one GPU race); more than 16384 bins are refused.
    void region_counts_noalloc(int32_t *out);
    void read_counters(uint64_t *out);
    void reset_counters();

    uint64_t n_voxels() const { return N_; }
    uint32_t slots() const { return S_; }
    uint64_t steps() const { return steps_; }

    // The skewed voxels of the current state (reservoir voxels and those the capacity hazards or negative counts
    // of the reference made), with their offsets.
    void read_skew(std::vector<SkewVoxel> &out);
    // The renderer's view of the current state (final_stage.wgsl:35-67), f32[N] in voxel order.
    void render_view(uint32_t view, uint32_t axis, uint32_t slice, float *out);

    // Well-mixed (CME) pool: counts, seed of the draws and the f64 precision audit.
    void read_cme_counts(int32_t *out);
    void set_cme_seed(uint32_t seed) { cme_seed_ = seed; }
    void set_cme_audit(bool on) { audit_on_ = on; }
    void read_cme_audit(double *max_rel_error, uint64_t *divergences, uint64_t *events);

private:
    void plan_wave_items(const uint8_t *lat);
    void plan_balanced_items(const std::vector<uint64_t> &cnt, double plane_cost, double total);
    void upload_items(const std::vector<std::vector<const WaveItem *>> &q);

    int dev_ = 0;
    void *stream_ = nullptr;       // hipStream_t
    void *kernel_ev_[2] = {nullptr, nullptr};  // hipEvent_t pair the next wave-kernel dispatch stamps (run_timed)
    uint32_t res_[3] = {0, 0, 0};
    uint64_t N_ = 0;
    uint32_t S_ = 0, ns_ = 0, nr_ = 0, nreg_ = 0, nrw_ = 0;
    float tau_ = 0.f;
    uint32_t mod_zero_ = 0xFFFFFFFFu;
    int cur_ = 0;                  // index of the buffer holding the current lattice
    bool has_zero_order_ = false;  // some reaction has no reactants
    uint32_t pair_n1_ = 0;         // ns + 1 when the table blob carries the pair table of the reaction reject test
    bool simple_ = false;          // one-lane-per-voxel kernel without LDS staging
    bool tiled_ = false;           // block-shared plane-march kernel instead of the wave-march kernel
    int tab_mode_ = 0;             // where the wave-march kernel reads the model tables: all in LDS (0), the hop
                                   // tables in LDS and the reaction tables in global memory (1), all global (2)
    size_t wave_shm_bytes() const; // its dynamic LDS for the tables
    uint32_t xc_ = 32;             // planes marched per tile: the longest x-range of a wave-kernel work item
    uint32_t wtiles_y_ = 0, wtiles_z_ = 0, pad_y_ = 0, pad_z_ = 0;  // wave tiles; padded label volume
    uint32_t ev_cap_ = 96;         // arrival events a wave-tile plane may hold
    uint32_t wave_blocks_ = 256;   // resident blocks of the persistent wave kernel
    void *items_ = nullptr;        // wave kernel work items (uint2 each), longest first
    uint32_t n_items_ = 0;
    uint32_t xc_min_ = 12;         // shortest x-range of a work item (its two side planes cost as much as
                                   // 1/6 of it; a longer one through a dense column would finish last)
    uint32_t items_per_block_ = 0; // wave-kernel work items per resident block (planning target; 0: the planner's
                                   // default, 1 balanced, 8 cost-ordered)
    bool balanced_plan_ = true;    // items of equal cost, one per resident block (else cost-ordered, shortest 12)
    bool counted_ = false;         // counted among the simulations alive on dev_
    bool min_planes_set_ = false;  // xc_min_ given by the environment
    uint32_t tiled_xc_ = 16;       // planes marched per block by the block-tiled kernel

    uint8_t *lat_[2] = {nullptr, nullptr};
    uint8_t *reg_ = nullptr;
    uint8_t *regp_ = nullptr;      // labels padded by two voxels (and to whole wave tiles), 0xFF outside
    void *tab_ = nullptr;          // one blob with every table below, at the byte offsets off_*
    uint32_t tab_words_ = 0;
    uint32_t off_p2_ = 0, off_rates_ = 0, off_ridx_ = 0, off_rmask_ = 0, off_stoich_ = 0, off_bound_ = 0;
    float *p2_ = nullptr;          // (D*tau)/(lambda*lambda), [src][dst][sp]
    float *rates_ = nullptr;       // [nr+1]
    uint8_t *ridx_ = nullptr;      // [nr+1][4]
    int8_t *stoich_ = nullptr;     // [nr+1][ns+1]
    uint64_t *rmask_ = nullptr;    // [ns+1][nrw]: reactions having the species as a reactant; row 0: zero-order
    float *bound_ = nullptr;       // [S+1]: upper bound of the total propensity of a voxel with c particles
    unsigned long long *counters_ = nullptr;
    int32_t *hist_ = nullptr;      // [nreg][ns+1]
    uint32_t *fix_ctl_ = nullptr;  // replay control words and the wave kernel's queues
    SkewStore *skew_ = nullptr;    // skewed voxels and the replay scratch

    bool has_cme_ = false;         // some well-mixed reaction exists
    uint32_t nc_ = 0, ncr_ = 0, cme_seed_ = 0;
    bool audit_on_ = false;
    int32_t *cme_counts_ = nullptr;
    void *cme_tab_ = nullptr;      // reactants, delta, emit, emit_region, rates, region voxel lists
    uint32_t off_creac_ = 0, off_cdelta_ = 0, off_cemit_ = 0, off_cregion_ = 0, off_crates_ = 0;
    uint32_t off_clists_ = 0, off_cloff_ = 0, off_cllen_ = 0;
    unsigned long long *audit_ = nullptr;  // [3]

    uint64_t steps_ = 0;           // steps enqueued
    std::vector<uint8_t> reg_host_;  // region label per voxel (per-region counts of skewed voxels' offsets)
    uint8_t *res_vol_ = nullptr;   // [N] reservoir species (reservoir models only), for the view
    bool summ_ = false;            // occupancy summaries kept (wave kernel, no zero-order reactions)
    uint8_t *sum_[2] = {nullptr, nullptr};  // summary of lat_[0], lat_[1]: a byte per wave tile and plane
    float *view_ = nullptr;        // [N] view output, allocated on first use
};

}  // namespace cs
