"""cs_sim_read_region_counts_noalloc (the entry point that allocates nothing, its bins in a fixed per-GPU array)
against cs_sim_read_region_counts, through the C ABI: two calls in a row, the second on a smaller model, so the
array holds the first call's bins when the second starts."""
import numpy as np
import pytest

from cell_simulation_amd import LatticeParams, RegionType, Simulation, configs, default_itime_schedule
from tests.conftest import announce

pytestmark = pytest.mark.gpu


def _reservoir_model():
    lam = 32e-9
    sim = Simulation(LatticeParams([20 * lam] * 3, [20] * 3, 50e-6, lam, 8))
    sim.set_seed(23)
    sim.add_region(RegionType.Cube("left", [0, 0, 0], [10 * lam, 20 * lam, 20 * lam]), 2e-12)
    sim.add_region(RegionType.Sphere("ball", [14 * lam, 10 * lam, 10 * lam], 4 * lam), 1e-12)
    sim.prepare_regions()
    sim.add_particle_count("Ares", "left", 300, True, True)
    sim.add_particle_count("A", "background", 1500, True, False)
    sim.add_particle_count("B", "left", 900, True, False)
    sim.add_particle_count("C", "ball", 100, True, False)
    sim.set_transition_rate_particle("A", "background", "left", 1e-12)
    sim.add_reaction(["Ares"], ["Ares", "B"], 4000.0)
    sim.add_reaction(["A", "B"], ["C"], 3000.0)
    sim.add_reaction(["A", "A"], ["B"], 800.0)  # counts below zero: offsets the slots do not show
    return sim


def test_region_counts_noalloc_matches_region_counts():
    announce("region counts without allocation: 20^3 reservoir model (4 regions x 5 species), then 16^3 cell")
    big = _reservoir_model()
    big.prepare_for_gpu(0)
    big.run(0, default_itime_schedule(23, 30))
    want = big.region_counts()
    got = big.region_counts_noalloc()
    assert want.shape == (3, 5) and np.array_equal(got, want), (got, want)
    assert np.array_equal(got.sum(axis=0)[1:], big.species_counts()[1:])
    # the reservoir units and negative counts are in the counts and in no slot
    lat = big.lattice().reshape(-1, 8)
    reg = big.regions().ravel()
    slots = np.zeros_like(want)
    for s in range(8):
        np.add.at(slots, (reg, lat[:, s].astype(np.int64)), 1)
    slots[:, 0] = want[:, 0]
    assert not np.array_equal(slots, want)
    # the second call: a smaller lattice and fewer bins (2 regions x 4 species)
    small = configs.spherical_cell(seed=5, edge=16, counts=(150, 150))
    small.prepare_for_gpu(0)
    small.run(0, default_itime_schedule(5, 20))
    want2 = small.region_counts()
    got2 = small.region_counts_noalloc()
    assert want2.shape == (2, 4) and np.array_equal(got2, want2), (got2, want2)
    assert got2[:, 1:].sum() > 0
    # and the first model again, after the smaller call
    assert np.array_equal(big.region_counts_noalloc(), want)
