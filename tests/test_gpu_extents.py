"""Degenerate and ragged lattice extents on the HIP path, step by step against the oracle.

Lattices one voxel wide along one or two axes, a single voxel, a single plane, extents just past a tile
(64 columns along z, 4 rows along y) and an empty lattice: the wave-march kernel's halo records, side planes and
labels outside the lattice, the block-tiled kernel and the one-lane-per-voxel kernel must all give the oracle's
lattices, counts and event counters. The y == 0 rows use X % 0 (hazard H4), which every one of these lattices has.
"""
import os

import pytest

from cell_simulation_amd import LatticeParams, Simulation, default_itime_schedule
from tests.conftest import announce
from tests.test_gpu_parity import run_parity

pytestmark = pytest.mark.gpu

# (extents, slots, A, B)
CASES = [
    ((1, 1, 200), 8, 40, 20),
    ((150, 1, 1), 4, 24, 12),
    ((1, 70, 1), 8, 40, 20),
    ((1, 1, 1), 4, 2, 1),
    ((3, 2, 129), 16, 300, 150),
    ((65, 5, 1), 8, 150, 60),
    ((1, 9, 67), 8, 200, 80),
    ((20, 20, 20), 8, 0, 0),
]


def _model(dims, slots, n_a, n_b, seed):
    lam = 32e-9
    sim = Simulation(LatticeParams([d * lam for d in dims], list(dims), 50e-6, lam, slots))
    sim.set_seed(seed)
    sim.prepare_regions()
    sim.add_particle_count("A", "background", n_a, True, False)
    sim.add_particle_count("B", "background", n_b, True, False)
    sim.set_diffusion_rate_particle("A", "background", 5e-12)
    sim.set_diffusion_rate_particle("B", "background", 2e-12)
    sim.add_reaction(["A", "B"], ["B"], 300.0)
    sim.add_reaction(["B"], ["A", "B"], 20.0)
    return sim


@pytest.mark.parametrize("kernel", ["wave", "tiled", "simple"])
@pytest.mark.parametrize("case", range(len(CASES)), ids=["x".join(map(str, c[0])) + f"-S{c[1]}" for c in CASES])
def test_extents_match_oracle(oracle_mod, case, kernel):
    dims, slots, n_a, n_b = CASES[case]
    seed = 100 + case
    announce(f"extents {dims} S={slots} ({kernel}), {n_a} A + {n_b} B, 120 steps")
    sim = _model(dims, slots, n_a, n_b, seed)
    orc = oracle_mod.Oracle.from_simulation(sim)
    saved = os.environ.get("CS_STEP_KERNEL")
    os.environ["CS_STEP_KERNEL"] = kernel
    try:
        run_parity(sim, orc, default_itime_schedule(seed, 120), 40, f"{dims} {kernel}")
    finally:
        if saved is None:
            os.environ.pop("CS_STEP_KERNEL", None)
        else:
            os.environ["CS_STEP_KERNEL"] = saved
    if n_a + n_b == 0:
        assert int(sim.lattice().sum()) == 0
