"""Synthetic workloads of the benchmark configurations, built through the Simulation API.

* `spherical_cell` -- 64^3 lattice, 32 nm voxels, S=8; extracellular space plus a spherical cell of
  radius 28 voxels; A + B <-> C (1e6 /M/s forward, 0.1 /s back), D = 1e-12 m^2/s; 2000 A and 2000 B
  placed uniformly in the cell from seed 42; dt = 50 us.
* `large_cell` -- 256^3 lattice, S=16, 32 nm voxels (the edge is not fixed by the workload; 32 nm keeps
  6p <= 1 for the fastest species); four regions (extracellular, a 2-voxel membrane shell, cytoplasm,
  nucleoid); 20 species with region-dependent D log-uniform in 1e-14..1e-11 m^2/s; 30 first- and
  second-order reactions with log-uniform rates (1e-3..1e1 /s, 1e5..1e8 /M/s); about 2M particles
  placed uniformly by region from seed 1234; dt = 10 us. `edge` scales the lattice down for tests.

* `ensemble_cells` -- replicate cells of the gene-expression network with edges drawn log-uniform in
  32..96 voxels, every rate constant (and the diffusion constants) perturbed per replicate by a
  log-normal factor (sigma 0.2), S=8, dt = 8 us (see ENSEMBLE_TAU), seed 99; per-cell itime seeds for an
  Ensemble.
* `precision_stress` -- 512^3 lattice, S=8, 32 nm voxels, dt = 5 us; the large-cell regions; 12 lattice
  species (about 10M particles in the cytoplasm and nucleoid) with region-dependent D, 40 well-mixed species
  (10..2000 copies each); 120 reactions in total: 40 per-voxel lattice reactions with rate constants
  log-uniform over 1e-6..1e1 /s, 72 well-mixed reactions and 8 well-mixed reactions that emit lattice particles,
  whose propensities at the initial counts are log-uniform over 1e-6..1e4 /s; seed 2024. `edge` scales it
  down for tests.
* `gene_expression` -- 128^3 lattice, 16 nm voxels, S=8, dt = 20 us; extracellular space, a membrane shell
  one voxel thick, cytoplasm and a nucleoid sphere of radius 20 voxels; well-mixed gene_off / gene_on
  switching (0.01 /s on, 0.05 /s off) with transcription of mRNA into the nucleoid at 0.5 /s while on;
  lattice mRNA (D = 5e-14 m^2/s, degraded at 0.01 /s, translated at 0.2 /s) and protein
  (D = 2.5e-12 m^2/s, degraded at 0.001 /s) in the cytoplasm and nucleoid. `rate_scale` multiplies every
  rate (tests use it to see events in few steps); `rate_factors` perturbs them per replicate.

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


GENE_REGIONS = ("background", "membrane", "cytoplasm", "nucleoid")


def gene_expression(seed: int = 7, edge: int = 128, slots: int = 8, tau: float = 20e-6, lam: float = 16e-9,
                    rate_scale: float = 1.0, rate_factors=None, nucleoid_radius: float = 20.0,
                    gene_copies: int = 1, populated: bool = False) -> Simulation:
    """Hybrid gene expression: well-mixed gene switching, lattice mRNA and protein.

    `rate_factors` (optional sequence of 7 positive factors) multiplies, in order: gene on, gene off,
    transcription, mRNA degradation, translation, protein degradation, and both diffusion constants.
    `populated` starts the cell near its stationary state instead of an empty lattice with the genes off: each
    gene copy on with probability p_on = k_on / (k_on + k_off) (drawn from `seed`), copies * p_on *
    transcription / mRNA decay mRNA in the nucleoid and that times translation / protein decay proteins in the
    cytoplasm (the means of the switching model)."""
    f = list(rate_factors) if rate_factors is not None else [1.0] * 7
    box = edge * lam
    sim = Simulation(LatticeParams([box] * 3, [edge] * 3, tau, lam, slots))
    sim.set_seed(seed)
    c = [box / 2] * 3
    r_ext = 0.45 * box
    r_int = r_ext - lam
    sim.add_region(RegionType.SphericalShell("membrane", "cytoplasm", c, r_int, r_ext), 0.0)
    sim.add_region(RegionType.Sphere("nucleoid", c, nucleoid_radius * lam * edge / 128), 0.0)
    sim.prepare_regions()
    n_mrna = n_prot = n_on = 0
    if populated:
        p_on = 0.01 * f[0] / (0.01 * f[0] + 0.05 * f[1])
        draw = np.random.Generator(np.random.PCG64(seed + 0x5EED))
        n_on = int(np.count_nonzero(draw.uniform(size=gene_copies) < p_on))
        m = gene_copies * p_on * (0.5 * f[2]) / (0.01 * f[3])
        n_mrna = int(round(m))
        n_prot = int(round(m * (0.2 * f[4]) / (0.001 * f[5])))
    sim.add_particle_count("mRNA", "nucleoid", n_mrna, True, False)
    sim.add_particle_count("protein", "cytoplasm", n_prot, True, False)
    d = {"mRNA": 5e-14 * f[6], "protein": 2.5e-12 * f[6]}
    for sp, dc in d.items():
        for reg in GENE_REGIONS:
            sim.set_diffusion_rate_particle(sp, reg, dc if reg in ("cytoplasm", "nucleoid") else 0.0)
        sim.set_transition_rate_particle(sp, "cytoplasm", "nucleoid", dc)
        sim.set_transition_rate_particle(sp, "nucleoid", "cytoplasm", dc)
    k = rate_scale
    sim.add_reaction(["mRNA"], [], 0.01 * k * f[3])
    sim.add_reaction(["mRNA"], ["mRNA", "protein"], 0.2 * k * f[4])
    sim.add_reaction(["protein"], [], 0.001 * k * f[5])
    sim.add_cme_species("gene_off", gene_copies - n_on)
    sim.add_cme_species("gene_on", n_on)
    sim.add_cme_reaction(["gene_off"], ["gene_on"], 0.01 * k * f[0])
    sim.add_cme_reaction(["gene_on"], ["gene_off"], 0.05 * k * f[1])
    sim.add_cme_reaction(["gene_on"], ["gene_on", "mRNA@nucleoid"], 0.5 * k * f[2])
    return sim


# Time step of the replicate cells: the gene-expression network at 8 us keeps every replicate's hop vector below 1
# (protein: 6 * D * f * tau / lambda^2 = 0.47 f, with f the replicate's diffusion factor). At the single cell's
# 20 us it is 1.17 f: the reference's cumulative vector then truncates the +z entry (rdme.wgsl:138-191), and every
# protein drifts onto the -z membrane, where the packed voxels make each step a capacity replay.
ENSEMBLE_TAU = 8e-6


def ensemble_cells(n_cells: int = 512, seed: int = 99, edge_range=(32, 96), sigma: float = 0.2,
                   rate_scale: float = 1.0, slots: int = 8, populated: bool = False, tau: float = ENSEMBLE_TAU):
    """Replicate cells for a batched ensemble: returns (cells, itime_seeds). `populated`: see gene_expression."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = edge_range
    edges = np.exp(rng.uniform(np.log(lo), np.log(hi), size=n_cells)).round().astype(int)
    cells, seeds = [], []
    for i, e in enumerate(edges):
        factors = np.exp(rng.normal(0.0, sigma, size=7))
        cells.append(gene_expression(seed=seed * 100003 + i, edge=int(e), slots=slots, tau=tau, rate_scale=rate_scale,
                                     rate_factors=factors, populated=populated))
        seeds.append(int(rng.integers(0, 2**32)))
    return cells, seeds


def precision_stress(seed: int = 2024, edge: int = 512, slots: int = 8, n_particles: int = 10_000_000,
                     n_rdme: int = 12, n_cme: int = 40, n_reactions: int = 120, tau: float = 5e-6,
                     lam: float = 32e-9) -> Simulation:
    rng = np.random.Generator(np.random.PCG64(seed))
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, tau, lam, slots))
    sim.set_seed(seed)
    c = [edge / 2 * lam] * 3
    r_ext = 0.47 * edge * lam
    sim.add_region(RegionType.SphericalShell("membrane", "cytoplasm", c, r_ext - 2 * lam, r_ext), 1e-12)
    sim.add_region(RegionType.Sphere("nucleoid", c, 0.2 * edge * lam), 1e-12)
    sim.prepare_regions()

    def loguniform(lo, hi):
        return float(np.exp(rng.uniform(np.log(lo), np.log(hi))))

    names = [f"R{i:02d}" for i in range(n_rdme)]
    per = n_particles // n_rdme
    for i, nm in enumerate(names):  # cytoplasm and nucleoid: ~0.2 particles per voxel
        sim.add_particle_count(nm, "nucleoid" if i % 3 == 2 else "cytoplasm", per, i < 4, False)
    for nm in names:
        for reg in LARGE_CELL_REGIONS:
            sim.set_diffusion_rate_particle(nm, reg, loguniform(1e-14, 1e-11))
    cme = [f"W{i:02d}" for i in range(n_cme)]
    init = {}
    for nm in cme:
        init[nm] = int(rng.integers(10, 2000))
        sim.add_cme_species(nm, init[nm])

    def rate_for(reac, a):
        """k giving propensity a at the initial counts: a = k * h, h the product of binomial(count, multiplicity)
        over distinct reactants (the well-mixed propensity's definition)."""
        h = 1.0
        for nm in set(reac):
            m, n = reac.count(nm), init[nm]
            h *= float(np.prod([n - j for j in range(m)])) / float(np.prod(range(1, m + 1)))
        return a / h
    n_lattice = n_reactions // 3
    n_emit = n_reactions // 15
    n_wm = n_reactions - n_lattice - n_emit
    for _ in range(n_lattice):
        a = names[int(rng.integers(n_rdme))]
        if rng.uniform() < 0.5:
            sim.add_reaction([a], [names[int(rng.integers(n_rdme))]], loguniform(1e-6, 1e1))
        else:  # two distinct species (k*n^2 of a homodimer row fires on a single particle)
            b = names[(names.index(a) + 1 + int(rng.integers(n_rdme - 1))) % n_rdme]
            sim.add_reaction([a, b], [names[int(rng.integers(n_rdme))]], loguniform(1e-6, 1e1))
    # well-mixed reactions: initial propensities log-uniform over 1e-6..1e4 /s
    for _ in range(n_wm):
        order = 1 if rng.uniform() < 0.7 else 2
        reac = [cme[int(i)] for i in rng.integers(0, n_cme, size=order)]
        prods = [cme[int(i)] for i in rng.integers(0, n_cme, size=int(rng.integers(0, 3)))]
        sim.add_cme_reaction(reac, prods, rate_for(reac, loguniform(1e-6, 1e4)))
    for _ in range(n_emit):
        src = cme[int(rng.integers(n_cme))]
        sim.add_cme_reaction([src], [src, f"{names[int(rng.integers(n_rdme))]}@cytoplasm"],
                             rate_for([src], loguniform(1e-6, 1e4)))
    return sim
