// capi.cpp -- extern "C" entry points declared in include/cell_sim.h.
//
// A cs_sim wraps the host model (model.hpp) and, after cs_sim_prepare_device, the device state
// (device.hpp). Exceptions never cross the ABI: they become status codes plus a thread-local message.
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cell_sim.h"
#include "device.hpp"
#include "ensemble.hpp"
#include "model.hpp"

struct cs_sim {
    std::unique_ptr<cs::Model> model;
    std::unique_ptr<cs::Device> device;
    std::deque<cs_stat_sample> stats;  // src/simulation.rs:36 (VecDeque<SolverStatisticSample<i32>>)
};

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

template <class F>
int guard(F &&f)
{
    try {
        f();
        return CS_OK;
    } catch (const cs::Error &e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(CS_ERR_ARG, "out of host memory");
    } catch (const std::exception &e) {
        return fail(CS_ERR_ARG, e.what());
    }
}

void need(const void *p, const char *what)
{
    if (!p)
        throw cs::Error(CS_ERR_ARG, std::string(what) + " is null");
}

cs::Model &model_of(cs_sim *s)
{
    need(s, "sim");
    return *s->model;
}

const cs::Model &model_of(const cs_sim *s)
{
    need(s, "sim");
    return *s->model;
}

// Builder calls change the model, which the device copy would no longer match.
cs::Model &builder(cs_sim *s)
{
    cs::Model &m = model_of(s);
    if (s->device)
        throw cs::Error(CS_ERR_STATE, "the model cannot change after cs_sim_prepare_device");
    return m;
}

std::string str(const char *p, const char *what)
{
    need(p, what);
    return std::string(p);
}

std::vector<int32_t> species_totals(cs_sim *s)
{
    const cs::Model &m = *s->model;
    const uint32_t n1 = m.num_species() + 1u, nreg = m.num_regions();
    std::vector<int32_t> out(n1, 0);
    if (s->device) {
        std::vector<int32_t> h(size_t(nreg) * n1);
        s->device->region_counts(h.data());
        for (uint32_t r = 0; r < nreg; ++r)
            for (uint32_t k = 1; k < n1; ++k)
                out[k] += h[size_t(r) * n1 + k];
    } else {
        for (uint8_t sp : m.lattice())
            if (sp)
                out[sp] += 1;
        const auto &res = m.reservoir();
        for (uint32_t r : res)
            if (r)
                out[r] += 1;
    }
    out[0] = 0;  // concentrations_stat[0] starts at 0 and is never updated (cme.wgsl:79 starts at 1)
    return out;
}

}  // namespace

extern "C" {

const char *cs_last_error(void) { return g_last_error.c_str(); }

int cs_sim_create(const cs_lattice_params *params, cs_sim **out)
{
    return guard([&] {
        need(params, "params");
        need(out, "out");
        *out = nullptr;
        cs::LatticeParams lp;
        for (int i = 0; i < 3; ++i) {
            lp.dims[i] = params->dims[i];
            lp.res[i] = params->res[i];
        }
        lp.tau = params->tau;
        lp.lambda = params->lambda;
        lp.max_particles_site = params->max_particles_site;
        auto s = std::make_unique<cs_sim>();
        s->model = std::make_unique<cs::Model>(lp);
        *out = s.release();
    });
}

int cs_sim_from_json_string(const char *json, uint64_t seed, uint32_t max_particles_site, cs_sim **out)
{
    return guard([&] {
        need(json, "json");
        need(out, "out");
        *out = nullptr;
        auto s = std::make_unique<cs_sim>();
        s->model = std::make_unique<cs::Model>(cs::model_from_json(json, seed, max_particles_site));
        *out = s.release();
    });
}

int cs_sim_from_json_file(const char *path, uint64_t seed, uint32_t max_particles_site, cs_sim **out)
{
    return guard([&] {
        need(path, "path");
        std::ifstream f(path, std::ios::binary);
        if (!f)
            throw cs::Error(CS_ERR_ARG, std::string("cannot open ") + path);
        std::stringstream ss;
        ss << f.rdbuf();
        const std::string text = ss.str();
        int rc = cs_sim_from_json_string(text.c_str(), seed, max_particles_site, out);
        if (rc != CS_OK)
            throw cs::Error(rc, g_last_error);
    });
}

void cs_sim_destroy(cs_sim *sim) { delete sim; }

int cs_sim_set_seed(cs_sim *sim, uint64_t seed)
{
    return guard([&] { builder(sim).set_seed(seed); });
}

int cs_sim_add_region_cube(cs_sim *sim, const char *name, const float p0[3], const float pf[3], float rate)
{
    return guard([&] {
        need(p0, "p0");
        need(pf, "pf");
        builder(sim).add_region_cube(str(name, "name"), p0, pf, rate);
    });
}

int cs_sim_add_region_sphere(cs_sim *sim, const char *name, const float center[3], float radius, float rate)
{
    return guard([&] {
        need(center, "center");
        builder(sim).add_region_sphere(str(name, "name"), center, radius, rate);
    });
}

int cs_sim_add_region_cylinder(cs_sim *sim, const char *name, const float p0[3], const float pf[3], float radius,
                               float rate)
{
    return guard([&] {
        need(p0, "p0");
        need(pf, "pf");
        builder(sim).add_region_cylinder(str(name, "name"), p0, pf, radius, rate);
    });
}

int cs_sim_add_region_semisphere(cs_sim *sim, const char *name, const float center[3], float radius,
                                 const float direction[3], float rate)
{
    return guard([&] {
        need(center, "center");
        need(direction, "direction");
        builder(sim).add_region_semisphere(str(name, "name"), center, radius, direction, rate);
    });
}

int cs_sim_add_region_spherical_shell(cs_sim *sim, const char *shell_name, const char *interior_name,
                                      const float center[3], float internal_radius, float external_radius,
                                      float rate)
{
    return guard([&] {
        need(center, "center");
        builder(sim).add_region_spherical_shell(str(shell_name, "shell_name"), str(interior_name, "interior_name"),
                                                center, internal_radius, external_radius, rate);
    });
}

int cs_sim_add_region_cylindrical_shell(cs_sim *sim, const char *shell_name, const char *interior_name,
                                        const float p0[3], const float pf[3], float internal_radius,
                                        float external_radius, float rate)
{
    return guard([&] {
        need(p0, "p0");
        need(pf, "pf");
        builder(sim).add_region_cylindrical_shell(str(shell_name, "shell_name"), str(interior_name, "interior_name"),
                                                  p0, pf, internal_radius, external_radius, rate);
    });
}

int cs_sim_add_region_capsid(cs_sim *sim, const char *shell_name, const char *interior_name, const float center[3],
                             const float dir[3], float internal_radius, float external_radius, float total_length,
                             float rate)
{
    return guard([&] {
        need(center, "center");
        need(dir, "dir");
        builder(sim).add_region_capsid(str(shell_name, "shell_name"), str(interior_name, "interior_name"), center,
                                       dir, internal_radius, external_radius, total_length, rate);
    });
}

int cs_sim_prepare_regions(cs_sim *sim)
{
    return guard([&] { builder(sim).prepare_regions(); });
}

int cs_sim_region_index(const cs_sim *sim, const char *name)
{
    if (!sim || !name)
        return -1;
    return sim->model->find_region_index(name);
}

int cs_sim_add_particle_count(cs_sim *sim, const char *name, const char *to_region, uint32_t count, int logging,
                              int is_reservoir)
{
    return guard([&] {
        builder(sim).add_particle_count(str(name, "name"), str(to_region, "to_region"), count, logging != 0,
                                        is_reservoir != 0);
    });
}

int cs_sim_add_particle_concentration(cs_sim *sim, const char *name, const char *to_region, float concentration,
                                      int logging, int is_reservoir)
{
    return guard([&] {
        builder(sim).add_particle_concentration(str(name, "name"), str(to_region, "to_region"), concentration,
                                                logging != 0, is_reservoir != 0);
    });
}

int cs_sim_fill_region(cs_sim *sim, const char *name, const char *to_region, int logging)
{
    return guard([&] { builder(sim).fill_region(str(name, "name"), str(to_region, "to_region"), logging != 0); });
}

int cs_sim_particle_random_walk(cs_sim *sim, const char *name, const char *to_region, float total_length,
                                float block_length, float radius, int logging)
{
    return guard([&] {
        builder(sim).particle_random_walk(str(name, "name"), str(to_region, "to_region"), total_length, block_length,
                                          radius, logging != 0);
    });
}

int cs_sim_add_sparse_region(cs_sim *sim, const char *name, float radius, const char *to_region,
                             uint32_t max_volume, float diffusion_rate)
{
    return guard([&] {
        builder(sim).add_sparse_region(str(name, "name"), radius, str(to_region, "to_region"), max_volume,
                                       diffusion_rate);
    });
}

int cs_sim_particle_index(const cs_sim *sim, const char *name)
{
    if (!sim || !name)
        return -1;
    return sim->model->find_particle(name);
}

int cs_sim_particle_name(const cs_sim *sim, uint32_t index, char *out, size_t cap)
{
    return guard([&] {
        need(out, "out");
        const cs::Model &m = model_of(sim);
        if (index > m.num_species())
            throw cs::Error(CS_ERR_ARG, "species index out of range");
        const std::string &n = m.particle_name(index);
        if (cap == 0)
            throw cs::Error(CS_ERR_ARG, "zero capacity");
        std::snprintf(out, cap, "%s", n.c_str());
    });
}

int cs_sim_set_diffusion_rate(cs_sim *sim, const char *region, float rate)
{
    return guard([&] { builder(sim).set_diffusion_rate(str(region, "region"), rate); });
}

int cs_sim_set_diffusion_rate_particle(cs_sim *sim, const char *particle, const char *region, float rate)
{
    return guard([&] { builder(sim).set_diffusion_rate_particle(str(particle, "particle"), str(region, "region"), rate); });
}

int cs_sim_set_transition_rate(cs_sim *sim, const char *from_region, const char *to_region, float rate)
{
    return guard([&] {
        builder(sim).set_transition_rate(str(from_region, "from_region"), str(to_region, "to_region"), rate);
    });
}

int cs_sim_set_transition_rate_particle(cs_sim *sim, const char *particle, const char *from_region,
                                        const char *to_region, float rate)
{
    return guard([&] {
        builder(sim).set_transition_rate_particle(str(particle, "particle"), str(from_region, "from_region"),
                                                  str(to_region, "to_region"), rate);
    });
}

int cs_sim_add_reaction(cs_sim *sim, const char *const *reactants, uint32_t n_reactants, const char *const *products,
                        uint32_t n_products, float k)
{
    return guard([&] {
        std::vector<std::string> r, p;
        if (n_reactants)
            need(reactants, "reactants");
        if (n_products)
            need(products, "products");
        for (uint32_t i = 0; i < n_reactants; ++i)
            r.push_back(str(reactants[i], "reactant"));
        for (uint32_t i = 0; i < n_products; ++i)
            p.push_back(str(products[i], "product"));
        builder(sim).add_reaction(r, p, k);
    });
}

int cs_sim_add_cme_species(cs_sim *sim, const char *name, int32_t count)
{
    return guard([&] { builder(sim).add_cme_species(str(name, "name"), count); });
}

int cs_sim_add_cme_reaction(cs_sim *sim, const char *const *reactants, uint32_t n_reactants,
                            const char *const *products, uint32_t n_products, float k)
{
    return guard([&] {
        std::vector<std::string> r, p;
        if (n_reactants)
            need(reactants, "reactants");
        if (n_products)
            need(products, "products");
        for (uint32_t i = 0; i < n_reactants; ++i)
            r.push_back(str(reactants[i], "reactant"));
        for (uint32_t i = 0; i < n_products; ++i)
            p.push_back(str(products[i], "product"));
        builder(sim).add_cme_reaction(r, p, k);
    });
}

int cs_sim_set_cme_seed(cs_sim *sim, uint32_t seed)
{
    return guard([&] {
        model_of(sim).set_cme_seed(seed);
        if (sim->device)
            sim->device->set_cme_seed(seed);
    });
}

int cs_sim_read_cme_counts(cs_sim *sim, int32_t *out, size_t n)
{
    return guard([&] {
        const cs::Model &m = model_of(sim);
        if (n != m.num_cme_species())
            throw cs::Error(CS_ERR_ARG, "buffer must hold num_cme_species entries");
        if (n == 0)
            return;
        need(out, "out");
        if (sim->device)
            sim->device->read_cme_counts(out);
        else
            std::memcpy(out, m.cme_init().data(), n * sizeof(int32_t));
    });
}

int cs_sim_cme_species_name(const cs_sim *sim, uint32_t index, char *out, size_t cap)
{
    return guard([&] {
        need(out, "out");
        const cs::Model &m = model_of(sim);
        if (index >= m.num_cme_species())
            throw cs::Error(CS_ERR_ARG, "well-mixed species index out of range");
        if (cap == 0)
            throw cs::Error(CS_ERR_ARG, "zero capacity");
        std::snprintf(out, cap, "%s", m.cme_name(index).c_str());
    });
}

int cs_sim_read_cme_tables(const cs_sim *sim, int32_t *reactants, size_t n_reactants, int32_t *delta,
                           size_t n_delta, uint8_t *emit, size_t n_emit, int32_t *emit_region, size_t n_region,
                           float *rates, size_t n_rates, uint32_t *seed)
{
    return guard([&] {
        const cs::Model &m = model_of(sim);
        auto copy = [](auto *dst, size_t n, const auto &src, const char *what) {
            if (n != src.size())
                throw cs::Error(CS_ERR_ARG, std::string(what) + ": size mismatch");
            if (n) {
                need(dst, what);
                std::memcpy(dst, src.data(), n * sizeof(src[0]));
            }
        };
        copy(reactants, n_reactants, m.cme_reactants(), "reactants");
        copy(delta, n_delta, m.cme_delta(), "delta");
        copy(emit, n_emit, m.cme_emit(), "emit");
        copy(emit_region, n_region, m.cme_emit_region(), "emit_region");
        copy(rates, n_rates, m.cme_rates(), "rates");
        need(seed, "seed");
        *seed = m.cme_seed();
    });
}

int cs_sim_set_cme_audit(cs_sim *sim, int on)
{
    return guard([&] {
        if (!sim || !sim->device)
            throw cs::Error(CS_ERR_STATE, "cs_sim_prepare_device first");
        sim->device->set_cme_audit(on != 0);
    });
}

int cs_sim_read_cme_audit(cs_sim *sim, double *max_rel_error, uint64_t *divergences, uint64_t *events)
{
    return guard([&] {
        if (!sim || !sim->device)
            throw cs::Error(CS_ERR_STATE, "cs_sim_prepare_device first");
        sim->device->read_cme_audit(max_rel_error, divergences, events);
    });
}

int cs_sim_prepare_device(cs_sim *sim, int hip_device)
{
    return guard([&] {
        cs::Model &m = model_of(sim);
        if (sim->device)
            throw cs::Error(CS_ERR_STATE, "device already prepared");
        sim->device = std::make_unique<cs::Device>(m, hip_device);
    });
}

int cs_sim_step(cs_sim *sim, uint32_t frame_num, uint32_t itime, uint32_t write_freq)
{
    return guard([&] {
        need(sim, "sim");
        if (!sim->device)
            throw cs::Error(CS_ERR_STATE, "cs_sim_prepare_device must be called before stepping");
        // simulation.rs:146-148: the readback sees the state before this step's passes.
        if (write_freq != 0 && frame_num % write_freq == 0 && frame_num > 0 &&
            !sim->model->logging_particles().empty()) {
            const std::vector<int32_t> tot = species_totals(sim);
            for (uint32_t sp : sim->model->logging_particles()) {
                cs_stat_sample smp;
                std::memset(&smp, 0, sizeof(smp));
                std::snprintf(smp.name, sizeof(smp.name), "%s", sim->model->particle_name(sp).c_str());
                smp.value = tot[sp];
                smp.iteration_count = frame_num;
                sim->stats.push_back(smp);
            }
        }
        sim->device->step(itime);
    });
}

int cs_sim_run(cs_sim *sim, uint32_t first_frame, const uint32_t *itime, uint32_t nsteps, uint32_t write_freq)
{
    return guard([&] {
        if (nsteps)
            need(itime, "itime");
        for (uint32_t t = 0; t < nsteps; ++t) {
            int rc = cs_sim_step(sim, first_frame + t, itime[t], write_freq);
            if (rc != CS_OK)
                throw cs::Error(rc, g_last_error);
        }
    });
}

int cs_sim_run_timed(cs_sim *sim, const uint32_t *itime, uint32_t nsteps, float *kernel_ms)
{
    return guard([&] {
        need(sim, "sim");
        need(kernel_ms, "kernel_ms");
        if (nsteps)
            need(itime, "itime");
        if (!sim->device)
            throw cs::Error(CS_ERR_STATE, "cs_sim_prepare_device must be called before stepping");
        *kernel_ms = sim->device->run_timed(itime, nsteps);
    });
}

int cs_sim_synchronize(cs_sim *sim)
{
    return guard([&] {
        need(sim, "sim");
        if (sim->device)
            sim->device->synchronize();
    });
}

int cs_sim_get_info(const cs_sim *sim, cs_sim_info *out)
{
    return guard([&] {
        need(out, "out");
        const cs::Model &m = model_of(sim);
        const cs::LatticeParams &p = m.params();
        for (int i = 0; i < 3; ++i) {
            out->res[i] = p.res[i];
            out->dims[i] = p.dims[i];
        }
        out->n_voxels = m.n_voxels();
        out->max_particles_site = m.slots();
        out->num_species = m.num_species();
        out->num_reactions = m.num_reactions();
        out->num_regions = m.num_regions();
        out->tau = p.tau;
        out->lambda = p.lambda;
        out->device_ready = sim->device ? 1 : 0;
        out->num_cme_species = m.num_cme_species();
        out->num_cme_reactions = m.num_cme_reactions();
    });
}

int cs_sim_read_lattice(cs_sim *sim, uint8_t *out, size_t nbytes)
{
    return guard([&] {
        need(out, "out");
        cs::Model &m = model_of(sim);
        const size_t n = m.n_voxels() * m.slots();
        if (nbytes != n)
            throw cs::Error(CS_ERR_ARG, "lattice buffer must hold N*S bytes");
        if (sim->device)
            sim->device->read_lattice(out);
        else
            std::memcpy(out, m.lattice().data(), n);
    });
}

int cs_sim_write_lattice(cs_sim *sim, const uint8_t *in, size_t nbytes)
{
    return guard([&] {
        need(in, "in");
        cs::Model &m = model_of(sim);
        const size_t n = m.n_voxels() * m.slots();
        if (nbytes != n)
            throw cs::Error(CS_ERR_ARG, "lattice buffer must hold N*S bytes");
        const uint32_t ns = m.num_species(), S = m.slots();
        for (size_t i = 0; i < n; ++i)
            if (in[i] > ns)
                throw cs::Error(CS_ERR_ARG, "lattice holds an unknown species id");
        // Occupancy is the number of occupied slots, which must come first in each voxel (the state
        // final_stage.wgsl:25-33 leaves behind).
        for (size_t v = 0; v < n / S; ++v)
            for (uint32_t k = 1; k < S; ++k)
                if (in[v * S + k] != 0 && in[v * S + k - 1] == 0)
                    throw cs::Error(CS_ERR_ARG, "lattice slots must be compacted (no empty slot before an occupied one)");
        if (sim->device)
            sim->device->write_lattice(in);
        else
            std::memcpy(m.lattice_mut().data(), in, n);
    });
}

int cs_sim_read_regions(const cs_sim *sim, uint8_t *out, size_t nbytes)
{
    return guard([&] {
        need(out, "out");
        const cs::Model &m = model_of(sim);
        if (nbytes != m.n_voxels())
            throw cs::Error(CS_ERR_ARG, "region buffer must hold N bytes");
        const auto &r = m.regions();
        for (size_t i = 0; i < r.size(); ++i)
            out[i] = uint8_t(r[i]);
    });
}

int cs_sim_read_reservoir(const cs_sim *sim, uint32_t *out, size_t n)
{
    return guard([&] {
        need(out, "out");
        const cs::Model &m = model_of(sim);
        if (n != m.n_voxels())
            throw cs::Error(CS_ERR_ARG, "reservoir buffer must hold N entries");
        std::memcpy(out, m.reservoir().data(), n * sizeof(uint32_t));
    });
}

int cs_sim_read_concentrations(cs_sim *sim, int32_t *out, size_t n)
{
    return guard([&] {
        need(out, "out");
        cs::Model &m = model_of(sim);
        const uint32_t n1 = m.num_species() + 1u, S = m.slots();
        const uint64_t N = m.n_voxels();
        if (n != N * n1)
            throw cs::Error(CS_ERR_ARG, "concentration buffer must hold N*(ns+1) entries");
        std::vector<uint8_t> lat(N * S);
        std::vector<cs::SkewVoxel> sk;
        int32_t c0 = 0;
        if (sim->device) {
            sim->device->read_lattice(lat.data());
            sim->device->read_skew(sk);
            c0 = sim->device->steps() > 0u ? 1 : 0;  // every step sets C[v][0] = 1 (cme.wgsl:31)
        } else {
            std::memcpy(lat.data(), m.lattice().data(), lat.size());
        }
        std::memset(out, 0, n * sizeof(int32_t));
        for (uint64_t v = 0; v < N; ++v) {
            int32_t *c = out + v * n1;
            c[0] = c0;
            for (uint32_t k = 0; k < S; ++k)
                if (lat[v * S + k])
                    c[lat[v * S + k]] += 1;
        }
        if (sim->device) {  // what the skewed voxels' counts hold beyond their slots
            for (const cs::SkewVoxel &e : sk)
                for (int q = 0; q < 4; ++q)
                    if (e.sp[q] != 0u && e.sp[q] < n1)
                        out[uint64_t(e.v) * n1 + e.sp[q]] += e.dc[q];
        } else {
            const auto &res = m.reservoir();
            for (uint64_t v = 0; v < N; ++v)
                if (res[v])
                    out[v * n1 + res[v]] += 1;  // lattice.rs:297
        }
    });
}

int cs_sim_read_occupancy(cs_sim *sim, uint32_t *out, size_t n)
{
    return guard([&] {
        need(out, "out");
        cs::Model &m = model_of(sim);
        const uint32_t S = m.slots();
        const uint64_t N = m.n_voxels();
        if (n != N)
            throw cs::Error(CS_ERR_ARG, "occupancy buffer must hold N entries");
        std::vector<uint8_t> lat(N * S);
        std::vector<cs::SkewVoxel> sk;
        if (sim->device) {
            sim->device->read_lattice(lat.data());
            sim->device->read_skew(sk);
        } else {
            std::memcpy(lat.data(), m.lattice().data(), lat.size());
        }
        for (uint64_t v = 0; v < N; ++v) {
            uint32_t c = 0;
            for (uint32_t k = 0; k < S; ++k)
                c += lat[v * S + k] != 0u;
            out[v] = c;
        }
        for (const cs::SkewVoxel &e : sk)
            out[e.v] += uint32_t(e.off);  // u32 wrap-around, as the reference's atomics
    });
}

int cs_sim_render_view(cs_sim *sim, uint32_t view, uint32_t slice_axis, uint32_t slice, float *out, size_t n)
{
    return guard([&] {
        need(out, "out");
        cs::Model &m = model_of(sim);
        if (n != m.n_voxels())
            throw cs::Error(CS_ERR_ARG, "view buffer must hold N floats");
        if (!sim->device)
            throw cs::Error(CS_ERR_STATE, "the view is computed on the device: call cs_sim_prepare_device first");
        sim->device->render_view(view, slice_axis, slice, out);
    });
}

int cs_sim_read_species_counts(cs_sim *sim, int32_t *out, size_t n)
{
    return guard([&] {
        need(out, "out");
        cs::Model &m = model_of(sim);
        if (n != m.num_species() + 1u)
            throw cs::Error(CS_ERR_ARG, "species count buffer must hold ns+1 entries");
        const std::vector<int32_t> t = species_totals(sim);
        std::memcpy(out, t.data(), n * sizeof(int32_t));
    });
}

int cs_sim_read_region_counts(cs_sim *sim, int32_t *out, size_t n)
{
    return guard([&] {
        need(out, "out");
        cs::Model &m = model_of(sim);
        const uint32_t n1 = m.num_species() + 1u, nreg = m.num_regions();
        if (n != size_t(nreg) * n1)
            throw cs::Error(CS_ERR_ARG, "region count buffer must hold nreg*(ns+1) entries");
        if (sim->device) {
            sim->device->region_counts(out);
        } else {
            std::memset(out, 0, n * sizeof(int32_t));
            const uint32_t S = m.slots();
            for (uint64_t v = 0; v < m.n_voxels(); ++v) {
                for (uint32_t k = 0; k < S; ++k) {
                    const uint8_t sp = m.lattice()[v * S + k];
                    if (sp)
                        out[size_t(m.regions()[v]) * n1 + sp] += 1;
                }
                if (m.reservoir()[v])
                    out[size_t(m.regions()[v]) * n1 + m.reservoir()[v]] += 1;  // lattice.rs:297
            }
        }
    });
}

int cs_sim_read_region_counts_noalloc(cs_sim *sim, int32_t *out, size_t n)
{
    return guard([&] {
        need(out, "out");
        cs::Model &m = model_of(sim);
        const uint32_t n1 = m.num_species() + 1u, nreg = m.num_regions();
        if (n != size_t(nreg) * n1)
            throw cs::Error(CS_ERR_ARG, "region count buffer must hold nreg*(ns+1) entries");
        if (!sim->device)
            throw cs::Error(CS_ERR_STATE, "no device state: call cs_sim_prepare_device first");
        sim->device->region_counts_noalloc(out);
    });
}

int cs_sim_read_tables(const cs_sim *sim, float *diffusion, size_t n_diffusion, int32_t *stoich, size_t n_stoich,
                       int32_t *reactants, size_t n_reactants, float *rates, size_t n_rates)
{
    return guard([&] {
        const cs::Model &m = model_of(sim);
        if (diffusion) {
            if (n_diffusion != m.diffusion().size())
                throw cs::Error(CS_ERR_ARG, "diffusion buffer size mismatch");
            std::memcpy(diffusion, m.diffusion().data(), n_diffusion * sizeof(float));
        }
        if (stoich) {
            if (n_stoich != m.stoich().size())
                throw cs::Error(CS_ERR_ARG, "stoichiometry buffer size mismatch");
            std::memcpy(stoich, m.stoich().data(), n_stoich * sizeof(int32_t));
        }
        if (reactants) {
            if (n_reactants != m.reactants().size())
                throw cs::Error(CS_ERR_ARG, "reactant buffer size mismatch");
            std::memcpy(reactants, m.reactants().data(), n_reactants * sizeof(int32_t));
        }
        if (rates) {
            if (n_rates != m.rates().size())
                throw cs::Error(CS_ERR_ARG, "rate buffer size mismatch");
            std::memcpy(rates, m.rates().data(), n_rates * sizeof(float));
        }
    });
}

int cs_sim_pop_stat(cs_sim *sim, cs_stat_sample *out)
{
    if (!sim || !out)
        return fail(CS_ERR_ARG, "null argument");
    if (sim->stats.empty())
        return CS_ERR_NOT_FOUND;
    *out = sim->stats.front();
    sim->stats.pop_front();
    return CS_OK;
}

int cs_sim_read_counters(cs_sim *sim, uint64_t *out, size_t n)
{
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        if (n != CS_NUM_COUNTERS)
            throw cs::Error(CS_ERR_ARG, "counter buffer must hold CS_NUM_COUNTERS entries");
        if (!sim->device)
            throw cs::Error(CS_ERR_STATE, "no device state");
        sim->device->read_counters(out);
    });
}

int cs_sim_reset_counters(cs_sim *sim)
{
    return guard([&] {
        need(sim, "sim");
        if (!sim->device)
            throw cs::Error(CS_ERR_STATE, "no device state");
        sim->device->reset_counters();
    });
}

}  // extern "C"

// ---- batched ensembles ---------------------------------------------------------------------------------
struct cs_ens {
    std::unique_ptr<cs::Ensemble> ens;
};

namespace {
cs::Ensemble &ens_of(cs_ens *e)
{
    need(e, "ensemble");
    return *e->ens;
}
const cs::Ensemble &ens_of(const cs_ens *e)
{
    need(e, "ensemble");
    return *e->ens;
}
}  // namespace

int cs_ens_create(cs_sim *const *cells, uint32_t n_cells, const uint32_t *itime_seeds, int hip_device, cs_ens **out)
{
    return guard([&] {
        need(out, "out");
        *out = nullptr;
        if (n_cells == 0)
            throw cs::Error(CS_ERR_ARG, "an ensemble needs at least one cell");
        need(cells, "cells");
        need(itime_seeds, "itime_seeds");
        std::vector<const cs::Model *> models;
        for (uint32_t i = 0; i < n_cells; ++i)
            models.push_back(&model_of(cells[i]));
        std::vector<uint32_t> seeds(itime_seeds, itime_seeds + n_cells);
        auto e = std::make_unique<cs_ens>();
        e->ens = std::make_unique<cs::Ensemble>(models, seeds, hip_device);
        *out = e.release();
    });
}

void cs_ens_destroy(cs_ens *ens) { delete ens; }

int cs_ens_run(cs_ens *ens, uint32_t nsteps)
{
    return guard([&] { ens_of(ens).run(nsteps); });
}

int cs_ens_run_graph(cs_ens *ens, uint32_t nsteps, uint32_t block)
{
    return guard([&] { ens_of(ens).run_graph(nsteps, block); });
}

int cs_ens_run_timed(cs_ens *ens, uint32_t nsteps, uint32_t block, float *ms)
{
    return guard([&] {
        need(ms, "ms");
        *ms = ens_of(ens).run_timed(nsteps, block);
    });
}

int cs_ens_synchronize(cs_ens *ens)
{
    return guard([&] { ens_of(ens).synchronize(); });
}

int cs_ens_set_tiling(cs_ens *ens, const cs_ens_tiling *t)
{
    return guard([&] {
        need(t, "tiling");
        ens_of(ens).set_tiling(cs::EnsembleTiling{t->xc, t->blocks_per_cu, t->batch});
    });
}

int cs_ens_get_tiling(const cs_ens *ens, cs_ens_tiling *t)
{
    return guard([&] {
        need(t, "tiling");
        const cs::EnsembleTiling &x = ens_of(ens).tiling();
        t->xc = x.xc;
        t->blocks_per_cu = x.blocks_per_cu;
        t->batch = x.batch;
    });
}

int cs_ens_autotune(cs_ens *ens, uint32_t trial_steps, float *ms_per_step)
{
    return guard([&] {
        if (trial_steps == 0)
            throw cs::Error(CS_ERR_ARG, "trial_steps must be positive");
        const float ms = ens_of(ens).autotune(trial_steps);
        if (ms_per_step)
            *ms_per_step = ms;
    });
}

int cs_ens_get_info(const cs_ens *ens, cs_ens_info *out)
{
    return guard([&] {
        need(out, "out");
        const cs::Ensemble &e = ens_of(ens);
        out->n_cells = e.n_cells();
        out->max_particles_site = e.slots();
        out->total_voxels = e.total_voxels();
        out->steps_done = e.steps_done();
        out->items = e.items();
    });
}

int cs_ens_cell_res(const cs_ens *ens, uint32_t cell, uint32_t res[3])
{
    return guard([&] {
        need(res, "res");
        ens_of(ens).cell_res(cell, res);
    });
}

int cs_ens_read_lattice(cs_ens *ens, uint32_t cell, uint8_t *out, size_t nbytes)
{
    return guard([&] {
        need(out, "out");
        cs::Ensemble &e = ens_of(ens);
        if (nbytes != e.cell_voxels(cell) * e.slots())
            throw cs::Error(CS_ERR_ARG, "buffer must hold the cell's N*S bytes");
        e.read_lattice(cell, out);
    });
}

int cs_ens_read_cme_counts(cs_ens *ens, uint32_t cell, int32_t *out, size_t n)
{
    return guard([&] {
        cs::Ensemble &e = ens_of(ens);
        if (cell >= e.n_cells())
            throw cs::Error(CS_ERR_ARG, "cell index out of range");
        if (n != e.cell_cme_species(cell))
            throw cs::Error(CS_ERR_ARG, "buffer must hold the cell's well-mixed species");
        if (n)
            need(out, "out");
        e.read_cme_counts(cell, out);
    });
}

int cs_ens_read_concentrations(cs_ens *ens, uint32_t cell, int32_t *out, size_t n)
{
    return guard([&] {
        need(out, "out");
        cs::Ensemble &e = ens_of(ens);
        const uint64_t N = e.cell_voxels(cell);
        const uint32_t S = e.slots(), n1 = e.cell_species(cell) + 1u;
        if (n != N * n1)
            throw cs::Error(CS_ERR_ARG, "concentration buffer must hold the cell's N*(ns+1) entries");
        std::vector<uint8_t> lat(N * S);
        std::vector<cs::SkewVoxel> sk;
        e.read_lattice(cell, lat.data());
        e.read_skew(cell, sk);
        const int32_t c0 = e.steps_done() > 0u ? 1 : 0;  // every step sets C[v][0] = 1 (cme.wgsl:31)
        std::memset(out, 0, n * sizeof(int32_t));
        for (uint64_t v = 0; v < N; ++v) {
            int32_t *c = out + v * n1;
            c[0] = c0;
            for (uint32_t k = 0; k < S; ++k)
                if (lat[v * S + k])
                    c[lat[v * S + k]] += 1;
        }
        for (const cs::SkewVoxel &x : sk)
            for (int q = 0; q < 4; ++q)
                if (x.sp[q] != 0u && x.sp[q] < n1)
                    out[uint64_t(x.v) * n1 + x.sp[q]] += x.dc[q];
    });
}

int cs_ens_read_occupancy(cs_ens *ens, uint32_t cell, uint32_t *out, size_t n)
{
    return guard([&] {
        need(out, "out");
        cs::Ensemble &e = ens_of(ens);
        const uint64_t N = e.cell_voxels(cell);
        const uint32_t S = e.slots();
        if (n != N)
            throw cs::Error(CS_ERR_ARG, "occupancy buffer must hold the cell's N entries");
        std::vector<uint8_t> lat(N * S);
        std::vector<cs::SkewVoxel> sk;
        e.read_lattice(cell, lat.data());
        e.read_skew(cell, sk);
        for (uint64_t v = 0; v < N; ++v) {
            uint32_t c = 0;
            for (uint32_t k = 0; k < S; ++k)
                c += lat[v * S + k] != 0u;
            out[v] = c;
        }
        for (const cs::SkewVoxel &x : sk)
            out[x.v] += uint32_t(x.off);  // u32 wrap-around, as the reference's atomics
    });
}

int cs_ens_read_counters(cs_ens *ens, uint64_t *out, size_t n)
{
    return guard([&] {
        need(out, "out");
        if (n != CS_NUM_COUNTERS)
            throw cs::Error(CS_ERR_ARG, "counter buffer must hold CS_NUM_COUNTERS entries");
        ens_of(ens).read_counters(out);
    });
}

int cs_ens_reset_counters(cs_ens *ens)
{
    return guard([&] { ens_of(ens).reset_counters(); });
}
