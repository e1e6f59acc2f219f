"""The benchmark workloads at their stated sizes and the reference's own model file, on the GPU against the oracle.

* the 64^3 spherical cell (A + B <-> C, 2000 A and 2000 B, dt = 50 us) for its full 1000 steps;
* the 128^3 hybrid gene-expression cell for 200 of its 10000 steps, four gene copies, started near the stationary
  state (genes on with probability k_on / (k_on + k_off), mRNA and protein at their stationary means), every rate
  x20000 (a common factor leaves those means as they are; at the stated rates nothing reacts in 200 steps of
  20 us); its protein hops with 6p = 1.17, so the +z entry of its cumulative vector is cut at 1 and it never stays
  (rdme.wgsl:138-191);
* the reference's saved_models/easy.json (tests/golden/easy.json: a capsid cell, A + B <-> C, a reservoir
  species in the membrane) loaded by Simulation.from_file (src/simulation.rs:210-357), placed from a seed and
  stepped for 150 of its 3 ms steps.
Each compares the oracle's whole state (lattice, per-voxel counts, occupancy words, global counts) and the
event counters.
"""
from pathlib import Path

import numpy as np
import pytest

from cell_simulation_amd import Simulation, configs, default_itime_schedule
from cell_simulation_amd._lib import CS_CNT_CME_EVENTS, CS_CNT_FIRED
from tests.conftest import announce
from tests.test_gpu_parity import run_parity

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_spherical_cell_64_full_run(oracle_mod):
    announce("spherical_cell 64^3 S=8, the full 1000 steps, seed 42")
    sim = configs.spherical_cell(seed=42)
    orc = oracle_mod.Oracle.from_simulation(sim)
    c = run_parity(sim, orc, default_itime_schedule(42, 1000), 250, "spherical_cell 1000")
    assert orc.counters.moves > 0 and c[CS_CNT_FIRED] > 100
    counts = sim.species_counts()
    assert counts[1] + counts[3] == 2000 and counts[2] + counts[3] == 2000  # A + C and B + C are conserved


def test_gene_expression_128_populated(oracle_mod):
    announce("gene_expression 128^3 S=8, 4 genes, populated near the stationary state, rates x20000, 200 steps")
    sim = configs.gene_expression(seed=7, populated=True, rate_scale=20000.0, gene_copies=4)
    orc = oracle_mod.Oracle.from_simulation(sim)
    assert sim.species_counts()[2] > 5000
    c = run_parity(sim, orc, default_itime_schedule(7, 200), 50, "gene_expression 128")
    assert c[CS_CNT_CME_EVENTS] == orc.cme_counters.events > 0
    assert np.array_equal(sim.cme_counts(), orc.cme_counts)


def test_easy_json_model(oracle_mod):
    announce("saved_models/easy.json via Simulation.from_file, seed 5, 150 steps")
    sim = Simulation.from_file(str(GOLDEN / "easy.json"), seed=5)
    inf = sim.info()
    assert tuple(inf["res"]) == (32, 32, 64) and inf["num_species"] == 4
    assert np.count_nonzero(sim.reservoir()) > 100  # the IeX reservoir sites of the membrane
    orc = oracle_mod.Oracle.from_simulation(sim)
    c = run_parity(sim, orc, default_itime_schedule(5, 150), 50, "easy.json")
    assert c[CS_CNT_FIRED] > 0 and orc.counters.moves > 0


def test_jammed_cell_s8(oracle_mod):
    """A 32^3 S=8 cell whose 16000 proteins fill a sphere of radius 14 voxels and hop with 6p = 2.1: the cumulative
    vector reaches 1 at -y (rdme.wgsl:138-191), so no particle stays or hops +y, +z or -z and the proteins pile up
    against the -y side of the sphere. Thousands of rejected claims and slot-S aliases follow, as in the smallest
    replicate cells whose perturbed diffusion constant takes 6p far past 1."""
    from cell_simulation_amd import LatticeParams, RegionType

    announce("jammed 32^3 S=8 cell, 16000 proteins with 6p = 2.1, 40 steps step by step")
    lam, tau, edge = 16e-9, 20e-6, 32
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, tau, lam, 8))
    sim.set_seed(162)
    sim.add_region(RegionType.Sphere("cytoplasm", [edge * lam / 2] * 3, 14 * lam), 4.5e-12)
    sim.prepare_regions()
    sim.add_particle_count("protein", "cytoplasm", 16000, True, False)
    sim.add_particle_count("mRNA", "cytoplasm", 60, True, False)
    sim.set_diffusion_rate_particle("protein", "background", 0.0)
    sim.set_diffusion_rate_particle("mRNA", "cytoplasm", 5e-14)
    sim.add_reaction(["mRNA"], ["mRNA", "protein"], 4000.0)
    sim.add_reaction(["protein"], [], 20.0)
    orc = oracle_mod.Oracle.from_simulation(sim)
    c = run_parity(sim, orc, default_itime_schedule(162, 40), 1, "jammed")
    assert orc.counters.failed_claims > 3000 and orc.counters.alias_writes > 200
    print(f"oracle {orc.counters.as_dict()}, tight voxel-steps {int(c[0])}")
