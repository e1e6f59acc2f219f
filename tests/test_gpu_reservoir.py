"""Reservoir particles on the HIP path, step by step against the oracle.

A reservoir site (src/lattice.rs:292-300) adds one unit of its species to the voxel's counts and to no
slot; the reaction (res/shader/cme.wgsl:76-110) skips the species reservoirs[v*S] at voxel v, an index in
lattice units, so it names the voxel's own entry only at v = 0 and another voxel's entry below N/S (an index
past the end reads 0). From then on the reference's per-voxel counts and occupancy word differ from the
slots at a few voxels: units that join the slots, slots that vanish from them while the counts keep them,
rollbacks that subtract the skipped species, occupancy words off by a signed amount (modulo 2^32, which
refuses every claim into an emptied voxel with a word of -1). The HIP step runs those voxels' reactions in
a pass of their own and replays the claims into voxels whose word is off the slot count; these cases check
the lattice, the per-voxel counts, the occupancy words, the global counts and the event counters after
every step, with each kernel.
"""
import os

import numpy as np
import pytest

from cell_simulation_amd import LatticeParams, RegionType, Simulation, default_itime_schedule
from cell_simulation_amd._lib import (CS_CNT_ALIAS, CS_CNT_EXCEPTION, CS_CNT_FAILED_CLAIM, CS_CNT_FIRED,
                                      CS_CNT_NEAR_THRESHOLD, CS_CNT_ROLLBACK)
from tests.conftest import announce, check_state, skewed_voxels

pytestmark = pytest.mark.gpu


def _reservoir_model(edge, slots, seed, n_res=400, density=0.3):
    lam = 32e-9
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, 50e-6, lam, slots))
    sim.set_seed(seed)
    sim.add_region(RegionType.Cube("cell", [0, 0, 0], [edge * lam] * 3), 2e-12)
    sim.prepare_regions()
    n = int(density * edge ** 3)
    sim.add_particle_count("Ares", "cell", n_res, True, True)
    sim.add_particle_count("A", "cell", n, True, False)
    sim.add_particle_count("B", "cell", n // 2, True, False)
    sim.add_particle_count("C", "cell", 0, True, False)
    sim.add_reaction(["Ares"], ["Ares", "B"], 4000.0)
    sim.add_reaction(["B"], [], 1000.0)
    sim.add_reaction(["A", "B"], ["C"], 5000.0)
    sim.add_reaction(["C"], ["A", "B"], 2000.0)
    sim.add_reaction(["Ares", "B"], ["A"], 3000.0)
    sim.add_reaction(["A", "A"], ["A", "A", "A"], 500.0)  # fills voxels: rollbacks near S
    return sim


def _check_step(sim, orc, tag):
    check_state(sim, orc, tag)


@pytest.mark.parametrize("kernel", ["wave", "wave-summary", "tiled", "simple"])
def test_reservoir_parity(oracle_mod, kernel):
    announce(f"reservoir model 20^3 S=8, 400 reservoir sites ({kernel}), 60 steps, step by step")
    os.environ["CS_STEP_KERNEL"] = kernel.split("-")[0]
    if kernel == "wave-summary":
        os.environ["CS_OCC_SUMMARY"] = "on"  # the replay's outputs must mark the occupancy summary
    try:
        sim = _reservoir_model(20, 8, seed=23)
        orc = oracle_mod.Oracle.from_simulation(sim)
        sim.prepare_for_gpu(0)
    finally:
        os.environ.pop("CS_STEP_KERNEL", None)
        os.environ.pop("CS_OCC_SUMMARY", None)
    S = 8
    res = sim.reservoir().ravel()
    N = res.size
    # the case must hold voxels that skip another voxel's reservoir species (v*S < N) and sites of their own
    k_vox = [v for v in range(N // S) if res[v * S]]
    assert len(k_vox) > 10 and np.count_nonzero(res) == 400
    _check_step(sim, orc, f"{kernel} initial state")
    # A + A -> 3A fills voxels: rejected claims, rollbacks and slot-S aliases (the first at step 41)
    skewed = 0
    for t, it in enumerate(default_itime_schedule(23, 60)):
        orc.step(int(it))
        sim.step(t, int(it), 0)
        _check_step(sim, orc, f"{kernel} step {t}")
        slots = np.count_nonzero(orc.lattice_u8().reshape(N, S), axis=1)
        skewed = max(skewed, int(np.count_nonzero(orc.occupancy.astype(np.int64) != slots)))
    c = sim.counters()
    seen = orc.counters.as_dict()
    print(f"oracle {seen}; occupancy words off the slot count: up to {skewed}")
    assert c[CS_CNT_EXCEPTION] == 0
    assert c[CS_CNT_FIRED] == seen["fired"] > 0
    assert c[CS_CNT_ROLLBACK] == seen["rollbacks"] > 0
    assert c[CS_CNT_NEAR_THRESHOLD] == seen["near_threshold"]
    assert c[CS_CNT_FAILED_CLAIM] == seen["failed_claims"] > 0
    assert c[CS_CNT_ALIAS] == seen["alias_writes"] > 0
    assert skewed > 20, "the case must leave occupancy words off the slot count"


def test_reservoir_with_well_mixed_species(oracle_mod):
    """Reservoir sites next to a well-mixed pool whose reactions emit lattice particles into the sites' region:
    the emitted particles join the slots after the voxel's particles, and the reservoir voxels keep their
    offsets (lattice.rs:292-300, cme.wgsl:80)."""
    announce("reservoir model 20^3 S=8 + well-mixed pool emitting A, 50 steps")
    sim = _reservoir_model(20, 8, seed=29, n_res=300, density=0.2)
    sim.add_cme_species("G", 5)
    sim.add_cme_species("H", 0)
    sim.add_cme_reaction(["G"], ["H", "A@cell"], 20000.0)
    sim.add_cme_reaction(["H"], ["G"], 15000.0)
    orc = oracle_mod.Oracle.from_simulation(sim)
    sim.prepare_for_gpu(0)
    for t, it in enumerate(default_itime_schedule(29, 50)):
        orc.step(int(it))
        sim.step(t, int(it), 0)
        _check_step(sim, orc, f"reservoir + pool step {t}")
        assert np.array_equal(sim.cme_counts(), orc.cme_counts)
    c = sim.counters()
    assert c[CS_CNT_EXCEPTION] == 0 and c[CS_CNT_FIRED] == orc.counters.fired > 0
    assert orc.cme_counters.events > 20 and skewed_voxels(orc) > 100


def test_view_matches_oracle(oracle_mod):
    """final_stage.wgsl:35-67 on the device against the oracle, on a reservoir model after 30 steps (so
    some occupancy words are off the slot count): every view, with and without slicing."""
    announce("view kernel: reservoir model 20^3 S=8 after 30 steps, views 0-2, slices")
    sim = _reservoir_model(20, 8, seed=23)
    orc = oracle_mod.Oracle.from_simulation(sim)
    sim.prepare_for_gpu(0)
    its = default_itime_schedule(23, 30)
    sim.run(0, its)
    orc.run(its)
    _check_step(sim, orc, "after 30 steps")
    for view, axis, sl in ((0, 0, 0), (1, 0, 0), (2, 0, 0), (0, 2, 7), (1, 1, 3), (0, 0, 25)):
        got = sim.render_view(view, axis, sl).ravel()
        exp = orc.view(view, axis, sl)
        assert np.array_equal(got, exp), (view, axis, sl, int(np.count_nonzero(got != exp)))
    assert np.count_nonzero(orc.view(0)) > 1000 and np.count_nonzero(orc.view(2)) == 400
