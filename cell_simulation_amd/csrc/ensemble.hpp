// ensemble.hpp -- batched ensembles: many independent cell lattices stepped together on one device.
//
// No reference counterpart: the reference steps one Simulation per dispatch chain (src/simulation.rs:
// 99-150). Here hundreds of replicate cells of different sizes and rate tables share one device state.
// Per step, a prologue kernel derives every cell's itime from a device-side step counter and runs the
// cells' well-mixed phases; one persistent step kernel takes (cell, tile, x-chunk) items from a
// device-side queue; an epilogue replays tight voxels cell by cell, re-arms the queue and advances the
// counter. Kernel arguments do not change between steps except the ping-pong parity, so blocks of steps
// are captured once as a hipGraph and replayed.
#pragma once

#include <cstdint>
#include <vector>

#include "device.hpp"
#include "model.hpp"

namespace cs {

struct EnsembleTiling {
    uint32_t xc = 16;            // planes per work item
    uint32_t blocks_per_cu = 3;  // persistent blocks per compute unit
    uint32_t batch = 1;          // items taken per queue access (0: static round-robin, no queue)
};

class Ensemble {
public:
    // itime of cell c at step t = low 32 bits of SplitMix64((itime_seeds[c] << 32) + t), the schedule a
    // plain simulation gets from default_itime_schedule(itime_seeds[c]).
    Ensemble(const std::vector<const Model *> &cells, const std::vector<uint32_t> &itime_seeds, int hip_device);
    ~Ensemble();
    Ensemble(const Ensemble &) = delete;
    Ensemble &operator=(const Ensemble &) = delete;

    void run(uint32_t nsteps);                        // launches per step
    void run_graph(uint32_t nsteps, uint32_t block);  // hipGraph of `block` steps, replayed
    float run_timed(uint32_t nsteps, uint32_t block); // run_graph (block > 0) or run, timed with events
    void synchronize();
    void set_tiling(const EnsembleTiling &t);
    const EnsembleTiling &tiling() const { return tiling_; }
    // Times candidate tilings on `trial_steps` steps each from a snapshot of the state, restores the
    // snapshot and keeps the fastest. Returns its time per step (ms).
    float autotune(uint32_t trial_steps);

    uint32_t n_cells() const { return n_cells_; }
    uint32_t slots() const { return S_; }
    uint64_t total_voxels() const { return total_vox_; }
    uint64_t steps_done() const { return steps_; }
    uint32_t items() const { return total_items_; }
    uint64_t cell_voxels(uint32_t c) const;
    void cell_res(uint32_t c, uint32_t res[3]) const;
    uint32_t cell_cme_species(uint32_t c) const { return cell_nc_.at(c); }
    uint32_t cell_species(uint32_t c) const { return cell_ns_.at(c); }
    void read_lattice(uint32_t cell, uint8_t *out);
    // The skewed voxels of one cell (voxel indices within the cell), with their offsets.
    void read_skew(uint32_t cell, std::vector<SkewVoxel> &out);
    void read_cme_counts(uint32_t cell, int32_t *out);
    void read_counters(uint64_t *out);
    void reset_counters();

private:
    void build_descriptors();
    void enqueue_step();
    void capture(uint32_t block);  // instantiates the graph of `block` steps (even) from the current parity
    void destroy_graph();

    int dev_ = 0;
    void *stream_ = nullptr;
    uint32_t n_cells_ = 0;
    uint32_t S_ = 0;
    uint64_t total_vox_ = 0;
    uint64_t steps_ = 0;
    uint32_t total_items_ = 0;
    uint32_t n_cu_ = 256;
    int cur_ = 0;
    EnsembleTiling tiling_;
    uint32_t tab_words_max_ = 0;

    std::vector<uint64_t> lat_off_;   // per cell, bytes
    std::vector<uint64_t> vox_off_;   // per cell, voxels
    std::vector<uint64_t> regp_off_;  // per cell, words of the padded label buffer
    std::vector<uint32_t> res_;       // per cell, 3 entries
    std::vector<uint32_t> cell_nc_;   // well-mixed species per cell
    std::vector<uint32_t> cell_ns_;   // lattice species per cell
    std::vector<uint32_t> cme_counts_off_;
    std::vector<uint8_t> cells_;      // host copy of the device descriptors (n_cells_ EnsCell), rebuilt per tiling

    uint8_t *lat_[2] = {nullptr, nullptr};
    uint8_t *sum_[2] = {nullptr, nullptr};   // occupancy summaries of lat_[0], lat_[1] (every cell's)
    uint64_t sum_total_ = 0;                 // bytes per summary
    uint32_t *item_cell_ = nullptr;          // [2 * total items] cell and item within it per queue entry (rebuilt with the tiling)
    uint8_t *reg_ = nullptr;
    void *regp_ = nullptr;             // padded labels of every cell (the wave kernel's label windows)
    uint32_t *tab_ = nullptr;          // all cells' table blobs
    void *cells_dev_ = nullptr;        // EnsCell[n]
    void *cme_args_dev_ = nullptr;     // CmeArgs[n_cme]
    void *cme_tab_ = nullptr;          // all cells' CME tables and region lists
    int32_t *cme_counts_ = nullptr;
    uint32_t *itime_ = nullptr;        // [n] per-step itime of each cell
    uint32_t *queue_ = nullptr;        // [0] next item, [1] step counter
    unsigned long long *counters_ = nullptr;
    unsigned long long *audit_ = nullptr;
    uint32_t *fix_ctl_ = nullptr;
    SkewStore *skew_ = nullptr;        // skewed voxels of every cell and the replay scratch
    void *graph_exec_ = nullptr;       // hipGraphExec_t
    uint32_t graph_block_ = 0;
    int graph_parity_ = 0;
};

}  // namespace cs
