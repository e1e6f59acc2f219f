// model.cpp -- host model builder. Each routine restates the reference routine it cites, keeping the
// f32 operation order (this file is compiled with -ffp-contract=off, so nothing is fused).
#include "model.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>

#include "../../include/cell_sim.h"
#include "json.hpp"

namespace cs {

namespace {

// Rust `f32 as usize`: truncation toward zero, saturating, NaN -> 0.
uint64_t f32_to_index(float f)
{
    if (!(f > 0.f))
        return 0;
    if (f >= 1.8446744e19f)
        return UINT64_MAX;
    return static_cast<uint64_t>(f);
}

// Rust `f32 as u32`.
uint32_t f32_to_u32(float f)
{
    if (!(f > 0.f))
        return 0;
    if (f >= 4294967296.f)
        return UINT32_MAX;
    return static_cast<uint32_t>(f);
}

}  // namespace

Model::Model(const LatticeParams &p) : params_(p), rng_(0)
{
    if (p.res[0] == 0 || p.res[1] == 0 || p.res[2] == 0)
        throw Error(CS_ERR_ARG, "lattice resolution must be positive");
    const uint32_t S = p.max_particles_site;
    if (!(S == 4 || S == 8 || S == 16))
        throw Error(CS_ERR_ARG, "max_particles_site must be 4, 8 or 16");
    if (p.n_voxels() * S >= (uint64_t(1) << 32))
        throw Error(CS_ERR_ARG, "lattice too large: N*S must stay below 2^32 (u32 slot index, functions.wgsl:42-46)");
    params_.n_regions = 1;
    const uint64_t N = p.n_voxels();
    // simulation.rs:59-66: region 0 is the background cube covering the box.
    regions_.assign(N, 0u);
    region_types_.push_back({RegionKind::Cube, "background"});
    for (int i = 0; i < 3; ++i)
        region_types_.back().b[i] = p.dims[i];
    volumes_.push_back(static_cast<uint32_t>(N));
    // lattice.rs:33: species 0 is "void".
    particle_names_.push_back("void");
    // simulation.rs:71-76
    diffusion_.assign(1, 8.15E-14f / 6.f);
    stoich_.assign(1, 0);
    reactants_.assign(3, 0);
    rates_.assign(1, 0.f);
    lattice_.assign(N * S, 0);
    occupancy_.assign(N, 0);
    reservoir_.assign(N, 0u);
}

// region.rs:288-294
void Model::check_regions_collision(uint64_t v)
{
    uint32_t old = region_at(v);
    if (old != 0)
        volumes_[old] -= 1;
}

// simulation.rs:537-548 with tensor-wgpu enlarge_dimension(0|1, transition_rate): the new source row and
// destination column are filled with transition_rate, then the new diagonal gets diffusion_rate for every
// species column that exists.
void Model::update_matrices_region(float diffusion_rate, float transition_rate)
{
    const uint32_t nsp = static_cast<uint32_t>(particle_names_.size());
    // The region descriptor was pushed already, so the matrix still has one region fewer.
    const uint32_t nreg = num_regions();
    const uint32_t nreg_old = nreg - 1u;
    std::vector<float> d(size_t(nreg) * nreg * nsp, transition_rate);
    for (uint32_t s = 0; s < nreg_old; ++s)
        for (uint32_t t = 0; t < nreg_old; ++t)
            for (uint32_t k = 0; k < nsp; ++k)
                d[k + size_t(nsp) * t + size_t(nsp) * nreg * s] = diffusion_[k + size_t(nsp) * t + size_t(nsp) * nreg_old * s];
    diffusion_.swap(d);
    const uint32_t r = static_cast<uint32_t>(region_types_.size()) - 1u;
    for (uint32_t k = 0; k < nsp; ++k)
        dm(r, r, k) = diffusion_rate;
    params_.n_regions += 1;
}

// simulation.rs:550-625
void Model::add_region_cylinder(const std::string &name, const float p0[3], const float pf[3], float radius,
                                float diffusion_rate)
{
    region_types_.push_back({RegionKind::Cylinder, name});
    for (int i = 0; i < 3; ++i) {
        region_types_.back().a[i] = p0[i];
        region_types_.back().b[i] = pf[i];
    }
    region_types_.back().r = radius;
    const uint32_t new_idx = static_cast<uint32_t>(region_types_.size()) - 1u;
    const float res[3] = {float(params_.res[0]), float(params_.res[1]), float(params_.res[2])};
    float vs[3];
    params_.voxel_size(vs);
    const float rl[3] = {radius / vs[0], radius / vs[1], radius / vs[2]};
    const float ip0[3] = {p0[0] / vs[0], p0[1] / vs[1], p0[2] / vs[2]};
    const float ipf[3] = {pf[0] / vs[0], pf[1] / vs[1], pf[2] / vs[2]};
    const float v[3] = {ipf[0] - ip0[0], ipf[1] - ip0[1], ipf[2] - ip0[2]};
    float v2 = v[0] * v[0];
    v2 = v2 + v[1] * v[1];
    v2 = v2 + v[2] * v[2];
    const float rr = radius * radius;
    const float r_const = rr * v2;
    const float vs2[3] = {vs[0] * vs[0], vs[1] * vs[1], vs[2] * vs[2]};
    uint32_t volume = 0;

    const float x0 = std::fmax(ip0[0] - rl[0], 0.f), y0 = std::fmax(ip0[1] - rl[1], 0.f),
                z0 = std::fmax(ip0[2] - rl[2], 0.f);
    const float xf = std::fmin(ipf[0] + rl[0], res[0]), yf = std::fmin(ipf[1] + rl[1], res[1]),
                zf = std::fmin(ipf[2] + rl[2], res[2]);
    float x = x0, y = y0, z = z0;
    while (x < xf) {
        while (y < yf) {
            while (z < zf) {
                float w0 = x - ipf[0], w1 = y - ipf[1], w2 = z - ipf[2];
                float proj = w0 * v[0];
                proj = proj + w1 * v[1];
                proj = proj + w2 * v[2];
                if (proj > 0.f) {
                    z += 1.f;
                    continue;
                }
                w0 = x - ip0[0];
                w1 = y - ip0[1];
                w2 = z - ip0[2];
                proj = w0 * v[0];
                proj = proj + w1 * v[1];
                proj = proj + w2 * v[2];
                if (proj < 0.f) {
                    z += 1.f;
                    continue;
                }
                const float c0 = w1 * v[2] - w2 * v[1];
                const float c1 = w2 * v[0] - w0 * v[2];
                const float c2 = w0 * v[1] - w1 * v[0];
                float cp_norm = (c0 * c0) * vs2[0];
                cp_norm = cp_norm + (c1 * c1) * vs2[1];
                cp_norm = cp_norm + (c2 * c2) * vs2[2];
                if (cp_norm <= r_const) {
                    const uint64_t vi = site_index(f32_to_index(x), f32_to_index(y), f32_to_index(z));
                    if (vi >= n_voxels())
                        throw Error(CS_ERR_ARG, "cylinder voxel outside the lattice");
                    check_regions_collision(vi);
                    set_region(new_idx, vi);
                    volume += 1;
                }
                z += 1.f;
            }
            z = z0;
            y += 1.f;
        }
        y = y0;
        x += 1.f;
    }
    update_matrices_region(diffusion_rate, 0.f);
    volumes_.push_back(volume);
}

// simulation.rs:627-658
void Model::add_region_cube(const std::string &name, const float p0[3], const float pf[3], float diffusion_rate)
{
    if (!(p0[0] <= pf[0] && p0[1] <= pf[1] && p0[2] <= pf[2]))
        throw Error(CS_ERR_ARG, "cube: p0 must not exceed pf");
    float vs[3];
    params_.voxel_size(vs);
    uint64_t start[3], end[3];
    for (int i = 0; i < 3; ++i) {
        start[i] = f32_to_index(p0[i] / vs[i]);
        end[i] = f32_to_index(pf[i] / vs[i]);
        if (end[i] > params_.res[i])
            throw Error(CS_ERR_ARG, "cube extends past the lattice");
    }
    region_types_.push_back({RegionKind::Cube, name});
    for (int i = 0; i < 3; ++i) {
        region_types_.back().a[i] = p0[i];
        region_types_.back().b[i] = pf[i];
    }
    const uint32_t new_idx = static_cast<uint32_t>(region_types_.size()) - 1u;
    const uint64_t volume = (end[0] - start[0]) * (end[1] - start[1]) * (end[2] - start[2]);
    for (uint64_t x = start[0]; x < end[0]; ++x)
        for (uint64_t y = start[1]; y < end[1]; ++y)
            for (uint64_t z = start[2]; z < end[2]; ++z) {
                const uint64_t vi = site_index(x, y, z);
                check_regions_collision(vi);
                set_region(new_idx, vi);
            }
    update_matrices_region(diffusion_rate, 0.f);
    volumes_.push_back(static_cast<uint32_t>(volume));
}

// simulation.rs:660-699
void Model::add_region_sphere(const std::string &name, const float center[3], float radius, float diffusion_rate)
{
    region_types_.push_back({RegionKind::Sphere, name});
    for (int i = 0; i < 3; ++i)
        region_types_.back().a[i] = center[i];
    region_types_.back().r = radius;
    const float res[3] = {float(params_.res[0]), float(params_.res[1]), float(params_.res[2])};
    const uint32_t new_idx = static_cast<uint32_t>(region_types_.size()) - 1u;
    float vs[3];
    params_.voxel_size(vs);
    const float c[3] = {center[0] / vs[0], center[1] / vs[1], center[2] / vs[2]};
    const float r2 = radius * radius;
    const float vs2[3] = {vs[0] * vs[0], vs[1] * vs[1], vs[2] * vs[2]};
    uint32_t volume = 0;
    for (float x = 0.f; x < res[0]; x += 1.f)
        for (float y = 0.f; y < res[1]; y += 1.f)
            for (float z = 0.f; z < res[2]; z += 1.f) {
                const float a = x - c[0], b = y - c[1], e = z - c[2];
                float dist = (a * a) * vs2[0];
                dist = dist + (b * b) * vs2[1];
                dist = dist + (e * e) * vs2[2];
                if (dist < r2) {
                    const uint64_t vi = site_index(uint64_t(x), uint64_t(y), uint64_t(z));
                    check_regions_collision(vi);
                    set_region(new_idx, vi);
                    volume += 1;
                }
            }
    update_matrices_region(diffusion_rate, 0.f);
    volumes_.push_back(volume);
}

// simulation.rs:701-743
void Model::add_region_semisphere(const std::string &name, const float center[3], float radius,
                                  const float direction[3], float diffusion_rate)
{
    region_types_.push_back({RegionKind::SemiSphere, name});
    for (int i = 0; i < 3; ++i)
        region_types_.back().a[i] = center[i];
    region_types_.back().r = radius;
    const float res[3] = {float(params_.res[0]), float(params_.res[1]), float(params_.res[2])};
    float vs[3];
    params_.voxel_size(vs);
    const float c[3] = {center[0] / vs[0], center[1] / vs[1], center[2] / vs[2]};
    const float r2 = radius * radius;
    const float vs2[3] = {vs[0] * vs[0], vs[1] * vs[1], vs[2] * vs[2]};
    const uint32_t new_idx = static_cast<uint32_t>(region_types_.size()) - 1u;
    uint32_t volume = 0;
    for (float x = 0.f; x < res[0]; x += 1.f)
        for (float y = 0.f; y < res[1]; y += 1.f)
            for (float z = 0.f; z < res[2]; z += 1.f) {
                const float a = x - c[0], b = y - c[1], e = z - c[2];
                float side = a * direction[0];
                side = side + b * direction[1];
                side = side + e * direction[2];
                if (side < 0.f)
                    continue;
                float dist = (a * a) * vs2[0];
                dist = dist + (b * b) * vs2[1];
                dist = dist + (e * e) * vs2[2];
                if (dist < r2) {
                    const uint64_t vi = site_index(uint64_t(x), uint64_t(y), uint64_t(z));
                    check_regions_collision(vi);
                    set_region(new_idx, vi);
                    volume += 1;
                }
            }
    update_matrices_region(diffusion_rate, 0.f);
    volumes_.push_back(volume);
}

// simulation.rs:439-443
void Model::add_region_spherical_shell(const std::string &shell, const std::string &interior, const float center[3],
                                       float internal_radius, float external_radius, float diffusion_rate)
{
    add_region_sphere(shell, center, external_radius, diffusion_rate);
    add_region_sphere(interior, center, internal_radius, diffusion_rate);
}

// simulation.rs:444-448
void Model::add_region_cylindrical_shell(const std::string &shell, const std::string &interior, const float p0[3],
                                         const float pf[3], float internal_radius, float external_radius,
                                         float diffusion_rate)
{
    add_region_cylinder(shell, p0, pf, external_radius, diffusion_rate);
    add_region_cylinder(interior, p0, pf, internal_radius, diffusion_rate);
}

// simulation.rs:449-476
void Model::add_region_capsid(const std::string &shell, const std::string &interior, const float center[3],
                              const float dir[3], float internal_radius, float external_radius, float total_length,
                              float diffusion_rate)
{
    const float cyl = total_length - 2.f * external_radius;
    float p0[3], pf[3], neg[3];
    for (int i = 0; i < 3; ++i) {
        p0[i] = center[i] - dir[i] * cyl / 2.f;
        pf[i] = center[i] + dir[i] * cyl / 2.f;
        neg[i] = -dir[i];
    }
    add_region_cylinder(shell, p0, pf, external_radius, diffusion_rate);
    add_region_cylinder(interior, p0, pf, internal_radius, diffusion_rate);
    add_region_semisphere("cap1", p0, external_radius, neg, diffusion_rate);
    add_region_semisphere("cap2", pf, external_radius, dir, diffusion_rate);
    add_region_semisphere("inside1", p0, internal_radius, neg, diffusion_rate);
    add_region_semisphere("inside2", pf, internal_radius, dir, diffusion_rate);
    join_regions("cap1", shell);
    join_regions("cap2", shell);
    join_regions("inside1", interior);
    join_regions("inside2", interior);
}

// simulation.rs:745-756: first primitive region with that name.
int Model::find_region_index(const std::string &name) const
{
    for (size_t i = 0; i < region_types_.size(); ++i)
        if (region_types_[i].name == name)
            return static_cast<int>(i);
    return -1;
}

// simulation.rs:758-790, with tensor-wgpu remove_element_at(0|1, idx) on the diffusion matrix.
void Model::join_regions(const std::string &region_delete, const std::string &to_region)
{
    const int di = find_region_index(region_delete), ti = find_region_index(to_region);
    if (di < 0 || ti < 0)
        throw Error(CS_ERR_NOT_FOUND, "join_regions: region not found");
    const uint32_t d = static_cast<uint32_t>(di), t = static_cast<uint32_t>(ti);
    const uint64_t N = n_voxels();
    for (uint64_t v = 0; v < N; ++v) {
        if (regions_[v] == d)
            regions_[v] = t;
        if (regions_[v] > d)
            regions_[v] -= 1;
    }
    const uint32_t nsp = static_cast<uint32_t>(particle_names_.size());
    const uint32_t nreg = num_regions();
    std::vector<float> nd;
    nd.reserve(size_t(nreg - 1) * (nreg - 1) * nsp);
    for (uint32_t s = 0; s < nreg; ++s) {
        if (s == d)
            continue;
        for (uint32_t q = 0; q < nreg; ++q) {
            if (q == d)
                continue;
            for (uint32_t k = 0; k < nsp; ++k)
                nd.push_back(diffusion_[k + size_t(nsp) * q + size_t(nsp) * nreg * s]);
        }
    }
    diffusion_.swap(nd);
    params_.n_regions -= 1;
    volumes_[t] += volumes_[d];
    region_types_.erase(region_types_.begin() + d);
    volumes_.erase(volumes_.begin() + d);
    regions_prepared_ = false;
}

// region.rs:129-148: per-region lists of voxel indices in ascending order.
void Model::prepare_regions()
{
    index_buffer_.clear();
    const uint64_t N = n_voxels();
    for (uint64_t v = 0; v < N; ++v)
        index_buffer_[regions_[v]].push_back(static_cast<uint32_t>(v));
    regions_prepared_ = true;
}

const std::vector<uint32_t> &Model::index_list(int region_idx) const
{
    if (!regions_prepared_)
        throw Error(CS_ERR_STATE, "prepare_regions must be called before adding particles");
    auto it = index_buffer_.find(static_cast<uint32_t>(region_idx));
    if (it == index_buffer_.end())
        throw Error(CS_ERR_FULL, "region has no voxels");
    return it->second;
}

// region.rs:204-248: a random point (real units) of a region's primitive; region kinds without a rule give
// the origin, as the reference's catch-all arm.
void Model::generate_point(uint32_t region_idx, float out[3])
{
    const RegionDesc &d = region_types_.at(region_idx);
    const float two_pi = 6.2831853071795864f, pi = 3.1415926535897932f;
    switch (d.kind) {
    case RegionKind::Cube:
        for (int i = 0; i < 3; ++i)
            out[i] = d.a[i] + (d.b[i] - d.a[i]) * rng_.unit();
        break;
    case RegionKind::Sphere:
    case RegionKind::SemiSphere: {  // the semisphere draws from its whole sphere (region.rs:225-234)
        const float theta = two_pi * rng_.unit(), phi = pi * rng_.unit(), r = d.r * rng_.unit();
        out[0] = r * std::cos(theta) * std::sin(phi) + d.a[0];
        out[1] = r * std::sin(theta) * std::sin(phi) + d.a[1];
        out[2] = r * std::cos(phi) + d.a[2];
        break;
    }
    case RegionKind::Cylinder: {  // a disc around p0 in x-y, z between the end centers (region.rs:235-244)
        const float theta = two_pi * rng_.unit(), r = d.r * rng_.unit();
        out[0] = r * std::cos(theta) + d.a[0];
        out[1] = r * std::sin(theta) + d.a[1];
        out[2] = d.a[2] + (d.b[2] - d.a[2]) * rng_.unit();
        break;
    }
    default:
        out[0] = out[1] = out[2] = 0.f;
    }
}

// region.rs:150-194: a voxel of the region whose box of +-radius_voxels lies inside the lattice and holds
// only that region's voxels; ten draws, then an error (the reference panics).
void Model::generate_boundary_aware(uint32_t region_idx, const uint64_t rv[3], uint64_t out[3])
{
    float vs[3];
    params_.voxel_size(vs);
    for (int attempt = 0; attempt < 10; ++attempt) {
        float pt[3];
        generate_point(region_idx, pt);
        uint64_t p[3];
        bool ok = true;
        for (int i = 0; i < 3; ++i) {
            const float f = pt[i] / vs[i];  // usize cast: saturates at 0, truncates
            p[i] = f > 0.f ? uint64_t(f) : 0u;
            if (p[i] >= params_.res[i])
                ok = false;
        }
        if (!ok || region_at(site_index(p[0], p[1], p[2])) != region_idx)
            continue;
        for (int i = 0; i < 3; ++i)
            if (p[i] < rv[i] || p[i] + rv[i] >= params_.res[i])
                ok = false;
        if (!ok)
            continue;
        // every voxel of [p - rv, p + rv) in the region (the reference compares the label sum)
        for (uint64_t x = p[0] - rv[0]; ok && x < p[0] + rv[0]; ++x)
            for (uint64_t y = p[1] - rv[1]; ok && y < p[1] + rv[1]; ++y)
                for (uint64_t z = p[2] - rv[2]; ok && z < p[2] + rv[2]; ++z)
                    if (region_at(site_index(x, y, z)) != region_idx)
                        ok = false;
        if (!ok)
            continue;
        for (int i = 0; i < 3; ++i)
            out[i] = p[i];
        return;
    }
    throw Error(CS_ERR_FULL, "no point of region " + region_types_.at(region_idx).name + " leaves room for the box");
}

// simulation.rs:483-535
void Model::add_sparse_region(const std::string &name, float radius, const std::string &to_region,
                              uint32_t max_volume, float diffusion_rate)
{
    const int ti = find_region_index(to_region);
    if (ti < 0)
        throw Error(CS_ERR_NOT_FOUND, "region not found: " + to_region);
    if (!(volumes_[ti] > max_volume))
        throw Error(CS_ERR_ARG, "sparse region: max_volume must be below the volume of " + to_region);
    float vs[3];
    params_.voxel_size(vs);
    const float vs2[3] = {vs[0] * vs[0], vs[1] * vs[1], vs[2] * vs[2]};
    const float r2 = radius * radius;
    uint64_t rv[3];
    for (int i = 0; i < 3; ++i) {
        const float f = radius / vs[i];
        rv[i] = f > 0.f ? uint64_t(f) : 0u;
    }
    if (rv[0] == 0 && rv[1] == 0 && rv[2] == 0)
        throw Error(CS_ERR_ARG, "sparse region: the sphere is smaller than a voxel");
    const uint32_t new_idx = static_cast<uint32_t>(region_types_.size());
    uint32_t volume = 0;
    while (volume < max_volume) {
        uint64_t p[3];
        generate_boundary_aware(uint32_t(ti), rv, p);
        for (uint64_t x = p[0] - rv[0]; x < p[0] + rv[0]; ++x)
            for (uint64_t y = p[1] - rv[1]; y < p[1] + rv[1]; ++y)
                for (uint64_t z = p[2] - rv[2]; z < p[2] + rv[2]; ++z) {
                    const float dx = float(x > p[0] ? x - p[0] : p[0] - x);
                    const float dy = float(y > p[1] ? y - p[1] : p[1] - y);
                    const float dz = float(z > p[2] ? z - p[2] : p[2] - z);
                    float dist = (dx * dx) * vs2[0];
                    dist = dist + (dy * dy) * vs2[1];
                    dist = dist + (dz * dz) * vs2[2];
                    if (dist < r2) {
                        const uint64_t vi = site_index(x, y, z);
                        check_regions_collision(vi);
                        set_region(new_idx, vi);
                        volume += 1;
                    }
                }
    }
    RegionDesc d{RegionKind::Sparse, name};
    d.r = radius;  // region.rs:82-85: a sphere at the origin stands for the whole set
    region_types_.push_back(d);
    update_matrices_region(diffusion_rate, 0.f);
    volumes_.push_back(volume);
    if (regions_prepared_)
        prepare_regions();
}

// lattice.rs:302-312: empties the voxel's last occupied slot.
bool Model::remove_particle_site(uint64_t v)
{
    if (occupancy_[v] == 0)
        return false;
    occupancy_[v] -= 1;
    lattice_[v * slots() + occupancy_[v]] = 0;
    return true;
}

// f32 -> u32 as Rust's `as` casts: truncation toward zero, saturating at 0 and 2^32 - 1 (NaN gives 0).
static uint32_t f32_to_u32_saturating(float f)
{
    if (!(f > 0.f))
        return 0u;
    if (f >= 4294967296.f)
        return 0xFFFFFFFFu;
    return uint32_t(f);
}

// lattice.rs:172-278. The box is the reference's: from min(ip0 - r, ipf + r) to max(ipf + r, ip0 - r) per
// axis, clipped to the lattice, walked in whole-voxel steps from that corner. A voxel belongs to the cylinder
// when it lies between the two end planes and within the radius (the radius test is in real units, scaled by
// the voxel size). The first voxel of the box outside the start voxel's region ends the block (false).
bool Model::add_particles_cylinder(const uint64_t site[3], const float dir[3], float length, float radius,
                                   uint8_t sp, uint64_t end[3], std::vector<uint64_t> &added)
{
    const uint32_t start_region = region_at(site_index(site[0], site[1], site[2]));
    float vs[3];
    params_.voxel_size(vs);
    const float res[3] = {float(params_.res[0]), float(params_.res[1]), float(params_.res[2])};
    float ip0[3], ipf[3], v[3], lo[3], hi[3];
    for (int i = 0; i < 3; ++i) {
        const float len = length / vs[i], rad = radius / vs[i];
        ip0[i] = float(site[i]);
        ipf[i] = ip0[i] + dir[i] * len;
        v[i] = ipf[i] - ip0[i];
        lo[i] = std::fmax(std::fmin(ip0[i] - rad, ipf[i] + rad), 0.f);
        hi[i] = std::fmin(std::fmax(ipf[i] + rad, ip0[i] - rad), res[i]);
    }
    const float v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    const float r_const = radius * radius * v2;
    const float vs2[3] = {vs[0] * vs[0], vs[1] * vs[1], vs[2] * vs[2]};
    for (float x = lo[0]; x < hi[0]; x += 1.f)
        for (float y = lo[1]; y < hi[1]; y += 1.f)
            for (float z = lo[2]; z < hi[2]; z += 1.f) {
                const uint64_t vi = site_index(uint64_t(x), uint64_t(y), uint64_t(z));
                if (region_at(vi) != start_region)
                    return false;
                const float w1[3] = {x - ipf[0], y - ipf[1], z - ipf[2]};
                if (w1[0] * v[0] + w1[1] * v[1] + w1[2] * v[2] > 0.f)
                    continue;  // past the far end plane
                const float w[3] = {x - ip0[0], y - ip0[1], z - ip0[2]};
                if (w[0] * v[0] + w[1] * v[1] + w[2] * v[2] < 0.f)
                    continue;  // before the near end plane
                const float c0 = w[1] * v[2] - w[2] * v[1], c1 = w[2] * v[0] - w[0] * v[2],
                            c2 = w[0] * v[1] - w[1] * v[0];
                const float cn = c0 * c0 * vs2[0] + c1 * c1 * vs2[1] + c2 * c2 * vs2[2];
                if (cn <= r_const) {
                    if (!add_particle_site(vi, sp))
                        return false;
                    added.push_back(vi);
                }
            }
    for (int i = 0; i < 3; ++i)
        end[i] = f32_to_u32_saturating(ipf[i]);
    return true;
}

// simulation.rs:842-864 and lattice.rs:113-170.
uint32_t Model::particle_random_walk(const std::string &name, const std::string &to_region, float total_length,
                                     float block_length, float radius, bool logging)
{
    const int ri = find_region_index(to_region);
    if (ri < 0)
        throw Error(CS_ERR_NOT_FOUND, "region not found: " + to_region);
    if (particle_names_.size() >= 255)
        throw Error(CS_ERR_UNSUPPORTED, "at most 254 species fit the one-byte slot encoding");
    if (!(block_length > 0.f) || !(total_length >= 0.f) || !(radius >= 0.f))
        throw Error(CS_ERR_ARG, "random walk: lengths and radius must be positive");
    const uint8_t sp = uint8_t(particle_names_.size());
    const int back = 25;  // blocks stepped back when a block fails
    float vs[3];
    params_.voxel_size(vs);
    const float safety = block_length * float(back) * 0.4f;
    uint64_t rv[3];
    for (int i = 0; i < 3; ++i) {
        const float f = safety / vs[i];
        rv[i] = f > 0.f ? uint64_t(f) : 0u;
    }
    uint64_t site[3];
    generate_boundary_aware(uint32_t(ri), rv, site);
    auto direction = [&](float d[3]) {  // utils.rs:14-21, uniform on the sphere
        const float theta = 6.2831853071795864f * rng_.unit();
        const float phi = std::acos(1.f - 2.f * rng_.unit());
        d[0] = std::sin(phi) * std::cos(theta);
        d[1] = std::sin(phi) * std::sin(theta);
        d[2] = std::cos(phi);
    };
    float dir[3];
    direction(dir);
    std::deque<std::array<uint64_t, 3>> sites{{site[0], site[1], site[2]}};
    std::deque<std::vector<uint64_t>> blocks;
    float length = 0.f;
    // The reference retries without bound; a block that keeps failing (radius wider than the region, say)
    // ends the walk with an error here instead.
    uint64_t failures = 0;
    while (length < total_length) {
        std::vector<uint64_t> added;
        uint64_t end[3];
        if (!add_particles_cylinder(site, dir, block_length, radius, sp, end, added)) {
            // the failed block's own particles stay, as in the reference; the last `back` blocks go
            if (++failures > 100000)
                throw Error(CS_ERR_FULL, "random walk of " + name + " keeps leaving region " + to_region);
            const size_t steps = std::min<size_t>(sites.size() - 1, size_t(back));
            for (size_t k = 0; k < steps; ++k) {
                sites.pop_back();
                if (!blocks.empty()) {
                    for (uint64_t vi : blocks.back())
                        if (!remove_particle_site(vi))
                            throw Error(CS_ERR_STATE, "random walk: stepping back met an empty voxel");
                    blocks.pop_back();
                }
                length -= block_length;
            }
            for (int i = 0; i < 3; ++i)
                site[i] = sites.front()[i];
        } else {
            length += block_length;
            sites.push_back({end[0], end[1], end[2]});
            if (sites.size() > size_t(back) + 1u)
                sites.pop_front();
            blocks.push_back(std::move(added));
            if (blocks.size() > size_t(back) + 1u)
                blocks.pop_front();
            for (int i = 0; i < 3; ++i)
                site[i] = end[i];
        }
        direction(dir);
    }
    new_species(name, logging);
    return sp;
}

int Model::find_particle(const std::string &name) const
{
    for (size_t i = 1; i < particle_names_.size(); ++i)
        if (particle_names_[i] == name)
            return static_cast<int>(i);
    return -1;
}

// lattice.rs:280-290
bool Model::add_particle_site(uint64_t v, uint8_t sp)
{
    const uint32_t S = slots();
    if (occupancy_[v] == S)
        return false;
    lattice_[v * S + occupancy_[v]] = sp;
    occupancy_[v] += 1;
    return true;
}

// lattice.rs:292-300
bool Model::add_reservoir_site(uint64_t v, uint32_t sp)
{
    if (reservoir_[v] > 0u)
        return false;
    reservoir_[v] = sp;
    reservoir_count_ += 1;
    return true;
}

// Bookkeeping shared by add_particle_count / fill_region (simulation.rs:796-813, 827-839): a new species
// column in the diffusion matrix copies the last column (tensor-wgpu copy_dimension(2)) and the
// stoichiometry matrix gains a zero column.
uint32_t Model::new_species(const std::string &name, bool logging)
{
    if (find_cme(name) >= 0)
        throw Error(CS_ERR_ARG, "name already used by a well-mixed species: " + name);
    const uint32_t nsp_old = static_cast<uint32_t>(particle_names_.size());
    if (nsp_old >= 255)
        throw Error(CS_ERR_UNSUPPORTED, "at most 254 species fit the one-byte slot encoding");
    const uint32_t idx = nsp_old;
    particle_names_.push_back(name);
    const uint32_t nreg = num_regions();
    std::vector<float> d(size_t(nreg) * nreg * (nsp_old + 1));
    for (uint32_t s = 0; s < nreg; ++s)
        for (uint32_t q = 0; q < nreg; ++q) {
            for (uint32_t k = 0; k < nsp_old; ++k)
                d[k + size_t(nsp_old + 1) * q + size_t(nsp_old + 1) * nreg * s] =
                    diffusion_[k + size_t(nsp_old) * q + size_t(nsp_old) * nreg * s];
            d[nsp_old + size_t(nsp_old + 1) * q + size_t(nsp_old + 1) * nreg * s] =
                diffusion_[(nsp_old - 1) + size_t(nsp_old) * q + size_t(nsp_old) * nreg * s];
        }
    diffusion_.swap(d);
    const uint32_t nrow = num_reactions() + 1u;
    std::vector<int32_t> st(size_t(nrow) * (nsp_old + 1), 0);
    for (uint32_t r = 0; r < nrow; ++r)
        for (uint32_t k = 0; k < nsp_old; ++k)
            st[size_t(r) * (nsp_old + 1) + k] = stoich_[size_t(r) * nsp_old + k];
    stoich_.swap(st);
    if (logging)
        logging_.push_back(idx);
    return idx;
}

// simulation.rs:792-814 + lattice.rs:63-97. Up to 100 uniform draws from the region's voxel list per
// particle; the draws come from the build's seeded SplitMix64 instead of rand::thread_rng.
void Model::add_particle_count(const std::string &name, const std::string &to_region, uint32_t count, bool logging,
                               bool is_reservoir)
{
    const int ri = find_region_index(to_region);
    if (ri < 0)
        throw Error(CS_ERR_NOT_FOUND, "region not found: " + to_region);
    if (particle_names_.size() >= 255)
        throw Error(CS_ERR_UNSUPPORTED, "at most 254 species fit the one-byte slot encoding");
    const uint32_t sp = static_cast<uint32_t>(particle_names_.size());
    const std::vector<uint32_t> *list = nullptr;
    if (count > 0)
        list = &index_list(ri);
    // Placement is undone if any particle cannot be placed (the reference panics, lattice.rs:93-95).
    std::vector<uint64_t> placed;
    placed.reserve(count);
    for (uint32_t n = 0; n < count; ++n) {
        int tries = 0;
        while (tries < 100) {
            const uint64_t pos = (*list)[rng_.below(list->size())];
            const bool ok = is_reservoir ? add_reservoir_site(pos, sp) : add_particle_site(pos, uint8_t(sp));
            if (ok) {
                placed.push_back(pos);
                break;
            }
            ++tries;
        }
        if (tries == 100) {
            for (auto it = placed.rbegin(); it != placed.rend(); ++it) {
                if (is_reservoir) {
                    reservoir_[*it] = 0u;
                    reservoir_count_ -= 1;
                } else {
                    occupancy_[*it] -= 1;
                    lattice_[*it * slots() + occupancy_[*it]] = 0;
                }
            }
            throw Error(CS_ERR_FULL, "could not add particle " + name + " to region " + to_region);
        }
    }
    new_species(name, logging);
}

// simulation.rs:816-823
void Model::add_particle_concentration(const std::string &name, const std::string &to_region, float concentration,
                                       bool logging, bool is_reservoir)
{
    const int ri = find_region_index(to_region);
    if (ri < 0)
        throw Error(CS_ERR_NOT_FOUND, "region not found: " + to_region);
    const float volume = static_cast<float>(volumes_[ri]);
    const uint32_t count = f32_to_u32(concentration * volume);
    add_particle_count(name, to_region, count, logging, is_reservoir);
}

// simulation.rs:825-840 + lattice.rs:99-111
void Model::fill_region(const std::string &name, const std::string &to_region, bool logging)
{
    const int ri = find_region_index(to_region);
    if (ri < 0)
        throw Error(CS_ERR_NOT_FOUND, "region not found: " + to_region);
    if (particle_names_.size() >= 255)
        throw Error(CS_ERR_UNSUPPORTED, "at most 254 species fit the one-byte slot encoding");
    const uint32_t sp = static_cast<uint32_t>(particle_names_.size());
    const auto &list = index_list(ri);
    for (uint32_t pos : list)
        if (occupancy_[pos] == slots())
            throw Error(CS_ERR_FULL, "could not add particle to region " + to_region);
    for (uint32_t pos : list)
        add_particle_site(pos, uint8_t(sp));
    new_species(name, logging);
}

namespace {
int position_of(const std::vector<std::string> &names, const std::string &n)
{
    for (size_t i = 0; i < names.size(); ++i)
        if (names[i] == n)
            return static_cast<int>(i);
    return -1;
}
}  // namespace

// simulation.rs:867-875
void Model::set_diffusion_rate(const std::string &region, float rate)
{
    const int r = find_region_index(region);
    if (r < 0)
        throw Error(CS_ERR_NOT_FOUND, "region not found: " + region);
    for (uint32_t k = 0; k < particle_names_.size(); ++k)
        dm(r, r, k) = rate;
}

// simulation.rs:877-883 (the particle is looked up among all names, void included)
void Model::set_diffusion_rate_particle(const std::string &particle, const std::string &region, float rate)
{
    const int r = find_region_index(region);
    if (r < 0)
        throw Error(CS_ERR_NOT_FOUND, "region not found: " + region);
    const int p = position_of(particle_names_, particle);
    if (p < 0)
        throw Error(CS_ERR_NOT_FOUND, "particle not found: " + particle);
    dm(r, r, p) = rate;
}

// simulation.rs:885-894
void Model::set_transition_rate(const std::string &from, const std::string &to, float rate)
{
    const int f = find_region_index(from), t = find_region_index(to);
    if (f < 0 || t < 0)
        throw Error(CS_ERR_NOT_FOUND, "region not found");
    for (uint32_t k = 0; k < particle_names_.size(); ++k)
        dm(f, t, k) = rate;
}

// simulation.rs:896-902
void Model::set_transition_rate_particle(const std::string &particle, const std::string &from, const std::string &to,
                                         float rate)
{
    const int f = find_region_index(from), t = find_region_index(to);
    if (f < 0 || t < 0)
        throw Error(CS_ERR_NOT_FOUND, "region not found");
    const int p = position_of(particle_names_, particle);
    if (p < 0)
        throw Error(CS_ERR_NOT_FOUND, "particle not found: " + particle);
    dm(f, t, p) = rate;
}

// simulation.rs:904-947
void Model::add_reaction(const std::vector<std::string> &reactants, const std::vector<std::string> &products, float k)
{
    std::vector<int32_t> ri, pi;
    for (const auto &r : reactants) {
        int i = find_particle(r);
        if (i < 0)
            throw Error(CS_ERR_NOT_FOUND, "reactant not found: " + r);
        ri.push_back(i);
    }
    for (const auto &p : products) {
        int i = find_particle(p);
        if (i < 0)
            throw Error(CS_ERR_NOT_FOUND, "product not found: " + p);
        pi.push_back(i);
    }
    if (ri.size() > 3)
        throw Error(CS_ERR_ARG, "at most three reactants (simulation.rs:937 pads the list to three)");
    const uint32_t nsp = static_cast<uint32_t>(particle_names_.size());
    std::vector<int32_t> row(nsp, 0);
    for (int32_t r : ri)
        row[r] -= 1;
    for (int32_t p : pi)
        row[p] += 1;
    stoich_.insert(stoich_.end(), row.begin(), row.end());
    while (ri.size() < 3)
        ri.push_back(0);
    reactants_.insert(reactants_.end(), ri.begin(), ri.end());
    rates_.push_back(k);
}

// ---------------------------------------------------------------------------------------------------
// Well-mixed (CME) species and reactions. No counterpart in the reference: the semantics are the
// build's own (see the CME phase in device.hip).

int Model::find_cme(const std::string &name) const
{
    for (size_t i = 0; i < cme_names_.size(); ++i)
        if (cme_names_[i] == name)
            return static_cast<int>(i);
    return -1;
}

void Model::add_cme_species(const std::string &name, int32_t count)
{
    if (name.empty() || name.find('@') != std::string::npos)
        throw Error(CS_ERR_ARG, "invalid well-mixed species name: " + name);
    if (count < 0)
        throw Error(CS_ERR_ARG, "well-mixed species counts are non-negative");
    if (find_particle(name) >= 0)
        throw Error(CS_ERR_ARG, "name already used by a lattice species: " + name);
    const int i = find_cme(name);
    if (i >= 0) {  // adding again raises the initial count, like add_particle_count on a known species
        cme_init_[size_t(i)] += count;
        return;
    }
    if (cme_names_.size() >= 255)
        throw Error(CS_ERR_UNSUPPORTED, "at most 255 well-mixed species");
    const size_t nc = cme_names_.size(), ncr = cme_rates_.size();
    std::vector<int32_t> d(ncr * (nc + 1), 0);
    for (size_t r = 0; r < ncr; ++r)
        for (size_t k = 0; k < nc; ++k)
            d[r * (nc + 1) + k] = cme_delta_[r * nc + k];
    cme_delta_.swap(d);
    cme_names_.push_back(name);
    cme_init_.push_back(count);
}

void Model::add_cme_reaction(const std::vector<std::string> &reactants, const std::vector<std::string> &products,
                             float k)
{
    if (!(k >= 0.0f) || !std::isfinite(k))
        throw Error(CS_ERR_ARG, "well-mixed reaction rates must be finite and non-negative");
    if (reactants.size() > 3)
        throw Error(CS_ERR_ARG, "at most three reactants");
    if (cme_rates_.size() >= 255)
        throw Error(CS_ERR_UNSUPPORTED, "at most 255 well-mixed reactions");
    const size_t nc = cme_names_.size();
    std::vector<int32_t> ri;
    std::vector<int32_t> delta(nc, 0);
    for (const auto &r : reactants) {
        const int i = find_cme(r);
        if (i < 0)
            throw Error(CS_ERR_NOT_FOUND, "well-mixed reactant not found: " + r);
        ri.push_back(i);
        delta[size_t(i)] -= 1;
    }
    uint8_t emit[4] = {0, 0, 0, 0};
    int n_emit = 0, region = -1;
    for (const auto &p : products) {
        const size_t at = p.find('@');
        if (at == std::string::npos) {
            const int i = find_cme(p);
            if (i < 0)
                throw Error(CS_ERR_NOT_FOUND, "well-mixed product not found (lattice products are name@region): " + p);
            delta[size_t(i)] += 1;
            continue;
        }
        const int sp = find_particle(p.substr(0, at));
        if (sp < 0)
            throw Error(CS_ERR_NOT_FOUND, "lattice species not found: " + p.substr(0, at));
        const int rg = find_region_index(p.substr(at + 1));
        if (rg < 0)
            throw Error(CS_ERR_NOT_FOUND, "region not found: " + p.substr(at + 1));
        if (region >= 0 && rg != region)
            throw Error(CS_ERR_UNSUPPORTED, "the lattice products of one reaction go to one region");
        if (n_emit == 4)
            throw Error(CS_ERR_UNSUPPORTED, "at most four lattice products per reaction");
        region = rg;
        emit[n_emit++] = uint8_t(sp);
    }
    while (ri.size() < 3)
        ri.push_back(-1);
    cme_reactants_.insert(cme_reactants_.end(), ri.begin(), ri.end());
    cme_delta_.insert(cme_delta_.end(), delta.begin(), delta.end());
    cme_emit_.insert(cme_emit_.end(), emit, emit + 4);
    cme_emit_region_.push_back(region);
    cme_rates_.push_back(k);
}

// ---------------------------------------------------------------------------------------------------
// JSON loader: simulation.rs:210-357, lattice_params.rs:143-166, utils.rs:103-125.

namespace {

const json::Value &need(const json::Value &obj, const char *key)
{
    const json::Value *v = obj.get(key);
    if (!v)
        throw Error(CS_ERR_PARSE, std::string("missing key \"") + key + "\"");
    return *v;
}

float as_f32(const json::Value &v, const char *what)
{
    if (v.kind != json::Value::Number)
        throw Error(CS_ERR_PARSE, std::string(what) + " must be a number");
    return static_cast<float>(v.num);
}

std::string as_str(const json::Value &v, const char *what)
{
    if (v.kind != json::Value::String)
        throw Error(CS_ERR_PARSE, std::string(what) + " must be a string");
    return v.str;
}

bool as_bool(const json::Value &v, const char *what)
{
    if (v.kind != json::Value::Bool)
        throw Error(CS_ERR_PARSE, std::string(what) + " must be a boolean");
    return v.b;
}

// utils.rs:103-125 for f32
void array3_f32(const json::Value &v, float out[3], const char *what)
{
    if (v.kind != json::Value::Array || v.arr.size() != 3)
        throw Error(CS_ERR_PARSE, std::string(what) + " must be an array of three numbers");
    for (int i = 0; i < 3; ++i)
        out[i] = as_f32(v.arr[i], what);
}

// utils.rs:139-146: only non-negative integral values that fit u32
void array3_u32(const json::Value &v, uint32_t out[3], const char *what)
{
    if (v.kind != json::Value::Array || v.arr.size() != 3)
        throw Error(CS_ERR_PARSE, std::string(what) + " must be an array of three numbers");
    for (int i = 0; i < 3; ++i) {
        if (v.arr[i].kind != json::Value::Number)
            throw Error(CS_ERR_PARSE, std::string(what) + " must be numeric");
        const double n = v.arr[i].num;
        if (!(n >= 0.0 && n <= 4294967295.0 && n == std::floor(n)))
            throw Error(CS_ERR_PARSE, std::string(what) + ": conversion error");
        out[i] = static_cast<uint32_t>(n);
    }
}

// Rust's str::parse::<f32>: decimal or inf/infinity/nan with an optional sign; no hex, no spaces.
bool parse_rust_f32(const std::string &tok, float &out)
{
    if (tok.empty())
        return false;
    size_t i = 0;
    if (tok[i] == '+' || tok[i] == '-')
        ++i;
    std::string rest = tok.substr(i);
    std::string low;
    for (char c : rest)
        low += char(std::tolower(static_cast<unsigned char>(c)));
    if (low == "inf" || low == "infinity" || low == "nan") {
        out = std::strtof(tok.c_str(), nullptr);
        return true;
    }
    bool digits = false, dot = false;
    size_t j = 0;
    for (; j < rest.size(); ++j) {
        char c = rest[j];
        if (c >= '0' && c <= '9')
            digits = true;
        else if (c == '.' && !dot)
            dot = true;
        else
            break;
    }
    if (!digits)
        return false;
    if (j < rest.size()) {
        if (rest[j] != 'e' && rest[j] != 'E')
            return false;
        ++j;
        if (j < rest.size() && (rest[j] == '+' || rest[j] == '-'))
            ++j;
        bool ed = false;
        for (; j < rest.size(); ++j) {
            if (rest[j] >= '0' && rest[j] <= '9')
                ed = true;
            else
                return false;
        }
        if (!ed)
            return false;
    }
    out = std::strtof(tok.c_str(), nullptr);
    return true;
}

std::vector<std::string> split_ws(const std::string &s)
{
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i])))
            ++i;
        size_t j = i;
        while (j < s.size() && !std::isspace(static_cast<unsigned char>(s[j])))
            ++j;
        if (j > i)
            out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

}  // namespace

Model model_from_json(const std::string &text, uint64_t seed, uint32_t max_particles_site)
{
    const json::Value root = json::parse(text);
    if (root.kind != json::Value::Object)
        throw Error(CS_ERR_PARSE, "model must be a JSON object");
    const json::Value &par = need(root, "parameters");
    LatticeParams lp;
    array3_f32(need(par, "dimensions"), lp.dims, "dimensions");
    array3_u32(need(par, "lattice_resolution"), lp.res, "lattice_resolution");
    lp.lambda = as_f32(need(par, "lambda"), "lambda");
    lp.tau = as_f32(need(par, "tau"), "tau");
    lp.max_particles_site = max_particles_site ? max_particles_site : 8u;
    Model m(lp);
    m.set_seed(seed);

    // simulation.rs:238-281. "cube" and "capsid" are the reference's types; "sphere", "semisphere",
    // "cylinder", "spherical_shell" and "cylindrical_shell" expose the other RegionType variants.
    const json::Value &regs = need(root, "regions");
    if (regs.kind != json::Value::Array)
        throw Error(CS_ERR_PARSE, "regions must be an array");
    for (const auto &r : regs.arr) {
        const std::string type = as_str(need(r, "type"), "type");
        const float rate = as_f32(need(r, "base_diffusion_rate"), "base_diffusion_rate");
        if (type == "cube") {
            float p0[3], pf[3];
            array3_f32(need(r, "p0"), p0, "p0");
            array3_f32(need(r, "pf"), pf, "pf");
            m.add_region_cube(as_str(need(r, "name"), "name"), p0, pf, rate);
        } else if (type == "capsid") {
            float c[3], d[3];
            array3_f32(need(r, "center"), c, "center");
            array3_f32(need(r, "dir"), d, "dir");
            m.add_region_capsid(as_str(need(r, "shell_name"), "shell_name"),
                                as_str(need(r, "interior_name"), "interior_name"), c, d,
                                as_f32(need(r, "internal_radius"), "internal_radius"),
                                as_f32(need(r, "external_radius"), "external_radius"),
                                as_f32(need(r, "total_length"), "total_length"), rate);
        } else if (type == "sphere") {
            float c[3];
            array3_f32(need(r, "center"), c, "center");
            m.add_region_sphere(as_str(need(r, "name"), "name"), c, as_f32(need(r, "radius"), "radius"), rate);
        } else if (type == "semisphere") {
            float c[3], d[3];
            array3_f32(need(r, "center"), c, "center");
            array3_f32(need(r, "direction"), d, "direction");
            m.add_region_semisphere(as_str(need(r, "name"), "name"), c, as_f32(need(r, "radius"), "radius"), d, rate);
        } else if (type == "cylinder") {
            float p0[3], pf[3];
            array3_f32(need(r, "p0"), p0, "p0");
            array3_f32(need(r, "pf"), pf, "pf");
            m.add_region_cylinder(as_str(need(r, "name"), "name"), p0, pf, as_f32(need(r, "radius"), "radius"), rate);
        } else if (type == "spherical_shell") {
            float c[3];
            array3_f32(need(r, "center"), c, "center");
            m.add_region_spherical_shell(as_str(need(r, "shell_name"), "shell_name"),
                                         as_str(need(r, "interior_name"), "interior_name"), c,
                                         as_f32(need(r, "internal_radius"), "internal_radius"),
                                         as_f32(need(r, "external_radius"), "external_radius"), rate);
        } else if (type == "cylindrical_shell") {
            float p0[3], pf[3];
            array3_f32(need(r, "p0"), p0, "p0");
            array3_f32(need(r, "pf"), pf, "pf");
            m.add_region_cylindrical_shell(as_str(need(r, "shell_name"), "shell_name"),
                                           as_str(need(r, "interior_name"), "interior_name"), p0, pf,
                                           as_f32(need(r, "internal_radius"), "internal_radius"),
                                           as_f32(need(r, "external_radius"), "external_radius"), rate);
        } else if (type == "sparse") {
            throw Error(CS_ERR_UNSUPPORTED, "sparse regions are not implemented (the reference panics too)");
        } else {
            throw Error(CS_ERR_PARSE, "unknown region type: " + type);
        }
    }
    m.prepare_regions();

    // simulation.rs:283-313
    const json::Value &parts = need(root, "particles");
    if (parts.kind != json::Value::Array)
        throw Error(CS_ERR_PARSE, "particles must be an array");
    for (const auto &p : parts.arr) {
        const std::string name = as_str(need(p, "name"), "name");
        const std::string to = as_str(need(p, "to_region"), "to_region");
        const bool logging = as_bool(need(p, "logging"), "logging");
        const bool res = as_bool(need(p, "is_reservoir"), "is_reservoir");
        if (const json::Value *c = p.get("count")) {
            if (!c->is_uint)
                throw Error(CS_ERR_PARSE, "count must be a non-negative integer");
            m.add_particle_count(name, to, static_cast<uint32_t>(c->uint_val), logging, res);
        } else if (const json::Value *c = p.get("concentration")) {
            m.add_particle_concentration(name, to, as_f32(*c, "concentration"), logging, res);
        } else {
            throw Error(CS_ERR_PARSE, "particle must have count or concentration");
        }
        if (const json::Value *d = p.get("diffusion_rate")) {
            if (d->kind != json::Value::Object)
                throw Error(CS_ERR_PARSE, "diffusion_rate must be an object");
            for (const auto &kv : d->obj)
                m.set_diffusion_rate_particle(name, kv.first, as_f32(kv.second, "diffusion_rate"));
        }
    }

    // simulation.rs:315-356: keys in sorted order; numeric tokens multiply k.
    const json::Value *reacts = root.get("reactions");
    if (reacts && reacts->kind != json::Value::Null) {
        if (reacts->kind != json::Value::Object)
            throw Error(CS_ERR_PARSE, "reactions must be an object");
        for (const auto &kv : reacts->obj) {
            float k = as_f32(kv.second, "reaction rate");
            std::vector<std::string> reactants, products;
            bool before = true;
            for (const auto &tok : split_ws(kv.first)) {
                float nk;
                if (parse_rust_f32(tok, nk)) {
                    k = k * nk;
                    continue;
                }
                if (tok == "->")
                    before = false;
                else if (tok == "+")
                    continue;
                else if (before)
                    reactants.push_back(tok);
                else
                    products.push_back(tok);
            }
            m.add_reaction(reactants, products, k);
        }
    }
    return m;
}

}  // namespace cs
