"""Synthetic workloads of the benchmark configurations, built through the Simulation API.

* `spherical_cell` -- 64^3 lattice, 32 nm voxels, S=8; extracellular space plus a spherical cell of
  radius 28 voxels; A + B <-> C (1e6 /M/s forward, 0.1 /s back), D = 1e-12 m^2/s; 2000 A and 2000 B
  placed uniformly in the cell from seed 42; dt = 50 us.
* `large_cell` -- 256^3 lattice, S=16, 32 nm voxels (the edge is not fixed by the workload; 32 nm keeps
  6p <= 1 for the fastest species); four regions (extracellular, a 2-voxel membrane shell, cytoplasm,
  nucleoid); 20 species with region-dependent D log-uniform in 1e-14..1e-11 m^2/s; 30 first- and
  second-order reactions with log-uniform rates (1e-3..1e1 /s, 1e5..1e8 /M/s); about 2M particles
  placed uniformly by region from seed 1234; dt = 10 us. `edge` scales the lattice down for tests.

Bimolecular rates are converted from /M/s to per-voxel /s (k / (N_A * V_voxel * 1000 L/m^3)) before
they reach the model: the reference applies k verbatim (res/shader/cme.wgsl:42-44).
"""
from __future__ import annotations

import numpy as np

from .simulation import LatticeParams, RegionType, Simulation

AVOGADRO = 6.02214076e23


def per_voxel_rate(k_molar: float, voxel_edge_m: float) -> float:
    """Second-order rate constant in /M/s to /s for one voxel of edge `voxel_edge_m`."""
    volume_l = voxel_edge_m ** 3 * 1e3
    return k_molar / (AVOGADRO * volume_l)


def spherical_cell(seed: int = 42, edge: int = 64, slots: int = 8, counts=(2000, 2000)) -> Simulation:
    lam = 32e-9
    box = edge * lam
    sim = Simulation(LatticeParams([box] * 3, [edge] * 3, 50e-6, lam, slots))
    sim.set_seed(seed)
    centre = [edge / 2 * lam] * 3
    radius = 28 * lam * edge / 64
    sim.add_region(RegionType.Sphere("cell", centre, radius), 1e-12)
    sim.prepare_regions()
    sim.add_particle_count("A", "cell", counts[0], True, False)
    sim.add_particle_count("B", "cell", counts[1], True, False)
    sim.add_particle_count("C", "cell", 0, True, False)
    for sp in ("A", "B", "C"):
        sim.set_diffusion_rate_particle(sp, "cell", 1e-12)
    sim.add_reaction(["A", "B"], ["C"], per_voxel_rate(1e6, lam))
    sim.add_reaction(["C"], ["A", "B"], 0.1)
    return sim


LARGE_CELL_REGIONS = ("background", "membrane", "cytoplasm", "nucleoid")


def large_cell(seed: int = 1234, edge: int = 256, slots: int = 16, n_particles: int = 2_000_000,
               n_species: int = 20, n_reactions: int = 30) -> Simulation:
    lam = 32e-9
    tau = 10e-6
    rng = np.random.Generator(np.random.PCG64(seed))
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, tau, lam, slots))
    sim.set_seed(seed)
    c = [edge / 2 * lam] * 3
    r_ext = 0.47 * edge * lam
    r_int = r_ext - 2 * lam
    r_nuc = 0.2 * edge * lam
    sim.add_region(RegionType.SphericalShell("membrane", "cytoplasm", c, r_int, r_ext), 1e-12)
    sim.add_region(RegionType.Sphere("nucleoid", c, r_nuc), 1e-12)
    sim.prepare_regions()

    names = [f"S{i:02d}" for i in range(n_species)]
    # Home region of each species: extracellular, membrane, cytoplasm and nucleoid in turn; the particle
    # budget is shared equally.
    per_species = n_particles // n_species
    for i, nm in enumerate(names):
        home = LARGE_CELL_REGIONS[i % 4]
        sim.add_particle_count(nm, home, per_species, i < 4, False)

    def loguniform(lo, hi):
        return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))

    # Region-dependent diffusion, plus transitions between adjacent regions.
    adjacent = [("background", "membrane"), ("membrane", "cytoplasm"), ("cytoplasm", "nucleoid")]
    for nm in names:
        for reg in LARGE_CELL_REGIONS:
            sim.set_diffusion_rate_particle(nm, reg, loguniform(1e-14, 1e-11))
        for a, b in adjacent:
            sim.set_transition_rate_particle(nm, a, b, loguniform(1e-14, 1e-11))
            sim.set_transition_rate_particle(nm, b, a, loguniform(1e-14, 1e-11))

    # Reaction network: half unimolecular (conversion or dissociation), half bimolecular association
    # between distinct species.
    for j in range(n_reactions):
        if j % 2 == 0:
            a = int(rng.integers(n_species))
            if rng.uniform() < 0.5:
                prods = [int(rng.integers(n_species))]
            else:
                prods = [int(rng.integers(n_species)), int(rng.integers(n_species))]
            sim.add_reaction([names[a]], [names[p] for p in prods], loguniform(1e-3, 1e1))
        else:
            a, b = rng.choice(n_species, size=2, replace=False)
            p = int(rng.integers(n_species))
            sim.add_reaction([names[int(a)], names[int(b)]], [names[p]],
                             per_voxel_rate(loguniform(1e5, 1e8), lam))
    return sim
