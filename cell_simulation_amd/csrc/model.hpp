// model.hpp -- host-side model of a cell simulation: lattice parameters, region labels, species,
// diffusion matrix, reaction tables and the initial particle placement.
//
// Mirrors the builder half of the reference's `Simulation` (src/simulation.rs:51-97, 424-947),
// `Lattice` placement (src/lattice.rs:63-111, 280-300) and `Regions` (src/region.rs:87-148). The
// device half (kernels and stepping) lives in device.hip.
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace cs {

// Error carrying a C-ABI status code; the C entry points translate it into a return value.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

// lattice_params.rs:15-24 (without the GPU padding) plus the per-build slot count.
struct LatticeParams {
    uint32_t max_particles_site = 8;
    uint32_t n_regions = 1;
    float lambda = 0.f;
    float tau = 0.f;
    float dims[3] = {0.f, 0.f, 0.f};
    uint32_t res[3] = {0, 0, 0};

    // lattice_params.rs:28-34
    void voxel_size(float out[3]) const
    {
        for (int i = 0; i < 3; ++i)
            out[i] = dims[i] / static_cast<float>(res[i]);
    }
    uint64_t n_voxels() const { return uint64_t(res[0]) * res[1] * res[2]; }
};

enum class RegionKind { Cube, Sphere, SemiSphere, Cylinder, Sparse };

// region.rs:9-20, primitives only: composite types are expanded into primitives when added. The geometry is
// what region.rs:204-248 draws random points from: a cube's corners p0 = a, pf = b; a sphere's or
// semisphere's center a and radius r; a cylinder's end centers a, b and radius r.
struct RegionDesc {
    RegionKind kind;
    std::string name;
    float a[3] = {0.f, 0.f, 0.f};
    float b[3] = {0.f, 0.f, 0.f};
    float r = 0.f;
};

// SplitMix64: the build's seeded placement generator (the reference draws from rand::thread_rng,
// which cannot be seeded; src/lattice.rs:65).
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed = 0) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // Uniform f32 in [0, 1) from the top 24 bits.
    float unit() { return float(next() >> 40) * (1.0f / 16777216.0f); }
    // Uniform integer in [0, n), n >= 1, without modulo bias (Lemire's multiply-and-reject).
    uint64_t below(uint64_t n)
    {
        for (;;) {
            uint64_t x = next();
            unsigned __int128 m = (unsigned __int128)x * n;
            uint64_t lo = (uint64_t)m;
            if (lo >= n || lo >= (0 - n) % n)
                return (uint64_t)(m >> 64);
        }
    }
};

class Model {
public:
    explicit Model(const LatticeParams &p);

    const LatticeParams &params() const { return params_; }
    uint32_t slots() const { return params_.max_particles_site; }
    uint64_t n_voxels() const { return params_.n_voxels(); }
    uint32_t num_species() const { return static_cast<uint32_t>(particle_names_.size()) - 1u; }
    uint32_t num_reactions() const { return static_cast<uint32_t>(rates_.size()) - 1u; }
    uint32_t num_regions() const { return static_cast<uint32_t>(region_types_.size()); }

    void set_seed(uint64_t seed)
    {
        rng_ = SplitMix64(seed);
        cme_seed_ = uint32_t(SplitMix64(seed ^ 0x434D4553ull).next());  // independent of the placement stream
    }

    // ---- regions (simulation.rs:428-790) ----
    void add_region_cube(const std::string &name, const float p0[3], const float pf[3], float diffusion_rate);
    void add_region_sphere(const std::string &name, const float center[3], float radius, float diffusion_rate);
    void add_region_cylinder(const std::string &name, const float p0[3], const float pf[3], float radius,
                             float diffusion_rate);
    void add_region_semisphere(const std::string &name, const float center[3], float radius,
                               const float direction[3], float diffusion_rate);
    void add_region_spherical_shell(const std::string &shell, const std::string &interior, const float center[3],
                                    float internal_radius, float external_radius, float diffusion_rate);
    void add_region_cylindrical_shell(const std::string &shell, const std::string &interior, const float p0[3],
                                      const float pf[3], float internal_radius, float external_radius,
                                      float diffusion_rate);
    void add_region_capsid(const std::string &shell, const std::string &interior, const float center[3],
                           const float dir[3], float internal_radius, float external_radius, float total_length,
                           float diffusion_rate);
    // simulation.rs:483-535: spheres of `radius` drawn inside `to_region` (each wholly inside it) become the
    // new region `name` until it holds at least max_volume voxels.
    void add_sparse_region(const std::string &name, float radius, const std::string &to_region, uint32_t max_volume,
                           float diffusion_rate);
    void prepare_regions();
    int find_region_index(const std::string &name) const;  // -1 if absent
    uint32_t region_volume(uint32_t idx) const { return volumes_.at(idx); }

    // ---- particles (simulation.rs:792-840) ----
    void add_particle_count(const std::string &name, const std::string &to_region, uint32_t count, bool logging,
                            bool is_reservoir);
    void add_particle_concentration(const std::string &name, const std::string &to_region, float concentration,
                                    bool logging, bool is_reservoir);
    void fill_region(const std::string &name, const std::string &to_region, bool logging);
    // simulation.rs:842-864 with lattice.rs:113-278: a new species laid as a chain of cylinders (blocks of
    // block_length, radius `radius`, random directions) of total length total_length inside to_region,
    // stepping back 25 blocks whenever a block leaves the region or meets a full voxel. Returns the species.
    uint32_t particle_random_walk(const std::string &name, const std::string &to_region, float total_length,
                                  float block_length, float radius, bool logging);
    int find_particle(const std::string &name) const;  // lattice.rs:315-322; -1 if absent
    const std::string &particle_name(uint32_t i) const { return particle_names_.at(i); }
    const std::vector<uint32_t> &logging_particles() const { return logging_; }

    // ---- diffusion (simulation.rs:867-902) ----
    void set_diffusion_rate(const std::string &region, float rate);
    void set_diffusion_rate_particle(const std::string &particle, const std::string &region, float rate);
    void set_transition_rate(const std::string &from, const std::string &to, float rate);
    void set_transition_rate_particle(const std::string &particle, const std::string &from, const std::string &to,
                                      float rate);

    // ---- reactions (simulation.rs:904-947) ----
    void add_reaction(const std::vector<std::string> &reactants, const std::vector<std::string> &products, float k);

    // ---- well-mixed (CME) species: an extension without a counterpart in the reference ----
    // Counts per simulation (one well-mixed pool), advanced once per step by an exact SSA over tau before
    // the lattice step. Reactants are CME species; a product is a CME species ("name") or an RDME
    // particle emitted into a region ("name@region"), placed at a random voxel of the region that has a
    // free slot.
    void add_cme_species(const std::string &name, int32_t count);
    void add_cme_reaction(const std::vector<std::string> &reactants, const std::vector<std::string> &products,
                          float k);
    int find_cme(const std::string &name) const;  // -1 if absent
    uint32_t num_cme_species() const { return static_cast<uint32_t>(cme_names_.size()); }
    uint32_t num_cme_reactions() const { return static_cast<uint32_t>(cme_rates_.size()); }
    const std::string &cme_name(uint32_t i) const { return cme_names_.at(i); }
    const std::vector<int32_t> &cme_init() const { return cme_init_; }            // [nc]
    const std::vector<int32_t> &cme_reactants() const { return cme_reactants_; }  // [ncr][3], -1 = none
    const std::vector<int32_t> &cme_delta() const { return cme_delta_; }          // [ncr][nc]
    const std::vector<uint8_t> &cme_emit() const { return cme_emit_; }            // [ncr][4], RDME species, 0 = none
    const std::vector<int32_t> &cme_emit_region() const { return cme_emit_region_; }  // [ncr], -1 = none
    const std::vector<float> &cme_rates() const { return cme_rates_; }            // [ncr]
    uint32_t cme_seed() const { return cme_seed_; }
    void set_cme_seed(uint32_t s) { cme_seed_ = s; }
    // Voxels of a region in ascending order (the placement lists of region.rs:129-148).
    const std::vector<uint32_t> &region_voxels(int region_idx) const { return index_list(region_idx); }

    // ---- tables in the reference's device layout ----
    // diffusion: [nreg][nreg][ns+1] (src, dst, species); stoich: [nr+1][ns+1]; reactants: [nr+1][3];
    // rates: [nr+1]. Row 0 of the reaction tables is the dummy all-zero reaction (simulation.rs:74-76).
    const std::vector<float> &diffusion() const { return diffusion_; }
    const std::vector<int32_t> &stoich() const { return stoich_; }
    const std::vector<int32_t> &reactants() const { return reactants_; }
    const std::vector<float> &rates() const { return rates_; }

    // ---- state ----
    const std::vector<uint32_t> &regions() const { return regions_; }      // [N]
    const std::vector<uint8_t> &lattice() const { return lattice_; }       // [N*S]
    std::vector<uint8_t> &lattice_mut() { return lattice_; }
    const std::vector<uint32_t> &reservoir() const { return reservoir_; }  // [N]
    bool has_reservoir() const { return reservoir_count_ > 0; }

private:
    uint32_t region_at(uint64_t v) const { return regions_[v]; }
    void check_regions_collision(uint64_t v);
    void set_region(uint32_t value, uint64_t v) { regions_[v] = value; }
    void update_matrices_region(float diffusion_rate, float transition_rate);
    void join_regions(const std::string &region_delete, const std::string &to_region);
    uint32_t new_species(const std::string &name, bool logging);
    float &dm(uint32_t src, uint32_t dst, uint32_t sp)
    {
        const uint32_t nsp = static_cast<uint32_t>(particle_names_.size());
        const uint32_t nreg = num_regions();
        return diffusion_[size_t(sp) + size_t(nsp) * dst + size_t(nsp) * nreg * src];
    }
    uint64_t site_index(uint64_t x, uint64_t y, uint64_t z) const
    {
        return z + y * params_.res[2] + x * uint64_t(params_.res[2]) * params_.res[1];
    }
    bool add_particle_site(uint64_t v, uint8_t sp);  // lattice.rs:280-290
    bool remove_particle_site(uint64_t v);           // lattice.rs:302-312
    void generate_point(uint32_t region_idx, float out[3]);  // region.rs:204-248
    void generate_boundary_aware(uint32_t region_idx, const uint64_t radius_voxels[3], uint64_t out[3]);
    // lattice.rs:172-278; false on a site outside the start's region or a full voxel (added sites stay listed)
    bool add_particles_cylinder(const uint64_t site[3], const float dir[3], float length, float radius, uint8_t sp,
                                uint64_t end[3], std::vector<uint64_t> &added);
    bool add_reservoir_site(uint64_t v, uint32_t sp);  // lattice.rs:292-300
    const std::vector<uint32_t> &index_list(int region_idx) const;

    LatticeParams params_;
    std::vector<uint32_t> regions_;
    std::vector<RegionDesc> region_types_;
    std::vector<uint32_t> volumes_;
    std::map<uint32_t, std::vector<uint32_t>> index_buffer_;
    bool regions_prepared_ = false;

    std::vector<std::string> particle_names_;
    std::vector<uint32_t> logging_;

    std::vector<float> diffusion_;      // [nreg][nreg][nsp], nsp = particle_names_.size()
    std::vector<int32_t> stoich_;       // [nr+1][nsp]
    std::vector<int32_t> reactants_;    // [nr+1][3]
    std::vector<float> rates_;          // [nr+1]

    std::vector<std::string> cme_names_;
    std::vector<int32_t> cme_init_;
    std::vector<int32_t> cme_reactants_;
    std::vector<int32_t> cme_delta_;
    std::vector<uint8_t> cme_emit_;
    std::vector<int32_t> cme_emit_region_;
    std::vector<float> cme_rates_;
    uint32_t cme_seed_ = uint32_t(SplitMix64(0x434D4553ull).next());

    std::vector<uint8_t> lattice_;
    std::vector<uint8_t> occupancy_;
    std::vector<uint32_t> reservoir_;
    uint64_t reservoir_count_ = 0;
    SplitMix64 rng_;
};

// Simulation::from_file (simulation.rs:210-357 + lattice_params.rs:143-166) on JSON text.
Model model_from_json(const std::string &text, uint64_t seed, uint32_t max_particles_site);

}  // namespace cs
