"""Bit-exact parity of the HIP step against the serial CPU oracle, through the C ABI.

Each case builds one model with the Simulation API, hands the same initial lattice, region labels and
tables to the oracle, runs both with the same per-step seeds and compares the reference's whole state: the
integer lattice (species per slot) byte for byte, the per-voxel counts, the occupancy words and the global
counts, plus the event counters (reactions fired, near-threshold draws, rejected claims, slot-S aliases). The
HIP path also reports how many voxel-steps were "tight" (could fill up, so the capacity hazards of
rdme.wgsl:57-87 apply); those are replayed in canonical order, and the voxels whose counters leave the slots
(slot-S aliases, rdme.wgsl:30-40; counts below zero, cme.wgsl:42-44) keep their offsets, so every case runs
to its end.
"""
import numpy as np
import pytest

from cell_simulation_amd import LatticeParams, RegionType, Simulation, configs, default_itime_schedule
from cell_simulation_amd._lib import (CS_CNT_ALIAS, CS_CNT_EXCEPTION, CS_CNT_FAILED_CLAIM, CS_CNT_FIRED,
                                      CS_CNT_NEAR_THRESHOLD, CS_CNT_ROLLBACK, CS_CNT_TIGHT)
from tests.conftest import announce, check_state, skewed_voxels

pytestmark = pytest.mark.gpu


def _compare(sim, orc, tag):
    g = sim.lattice()
    o = orc.lattice_u8()
    if not np.array_equal(g, o):
        diff = np.argwhere(np.any(g != o, axis=-1))
        v = tuple(diff[0])
        raise AssertionError(f"{tag}: {len(diff)} voxels differ; first {v}: gpu {g[v]} oracle {o[v]}")


def check_counters(sim, orc):
    c = sim.counters()
    assert c[CS_CNT_EXCEPTION] == 0
    assert c[CS_CNT_NEAR_THRESHOLD] == orc.counters.near_threshold
    assert c[CS_CNT_FIRED] == orc.counters.fired
    assert c[CS_CNT_ROLLBACK] == orc.counters.rollbacks
    assert c[CS_CNT_FAILED_CLAIM] == orc.counters.failed_claims
    assert c[CS_CNT_ALIAS] == orc.counters.alias_writes
    return c


def run_parity(sim, orc, itimes, check_every, tag):
    sim.prepare_for_gpu(0)
    done = 0
    while done < len(itimes):
        n = min(check_every, len(itimes) - done)
        sim.run(done, itimes[done:done + n])
        orc.run(itimes[done:done + n])
        done += n
        check_state(sim, orc, f"{tag} after {done} steps")
    return check_counters(sim, orc)


def test_spherical_cell_64(oracle_mod):
    announce("spherical_cell 64^3 S=8, 300 steps, seed 42")
    sim = configs.spherical_cell(seed=42)
    orc = oracle_mod.Oracle.from_simulation(sim)
    c = run_parity(sim, orc, default_itime_schedule(42, 300), 100, "spherical_cell")
    assert orc.counters.moves > 0 and c[CS_CNT_FIRED] > 0


def test_large_cell_48(oracle_mod):
    announce("large_cell 48^3 S=16 (20 species, 30 reactions, 4 regions), 120 steps, seed 1234")
    sim = configs.large_cell(seed=1234, edge=48, n_particles=13_000)
    orc = oracle_mod.Oracle.from_simulation(sim)
    run_parity(sim, orc, default_itime_schedule(1234, 120), 40, "large_cell_48")
    assert orc.counters.fired > 0


def test_large_cell_s8_fast_reactions(oracle_mod):
    """Same network with every rate scaled up so reactions (and rollbacks at S=8) are frequent."""
    announce("large_cell 32^3 S=8 fast reactions, 100 steps, seed 5")
    sim = configs.large_cell(seed=5, edge=32, slots=8, n_particles=4_000)
    orc = oracle_mod.Oracle.from_simulation(sim)
    run_parity(sim, orc, default_itime_schedule(5, 100), 25, "large_cell_s8")


def _dense_model(edge, slots, density, seed, k_fwd=50.0, k_back=5.0, zero_order=0.0):
    lam = 32e-9
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, 50e-6, lam, slots))
    sim.set_seed(seed)
    sim.add_region(RegionType.Cube("left", [0, 0, 0], [edge * lam / 2, edge * lam, edge * lam]), 2e-12)
    sim.prepare_regions()
    n = int(density * edge ** 3 / 3)
    sim.add_particle_count("A", "background", n, True, False)
    sim.add_particle_count("B", "left", n, True, False)
    sim.add_particle_count("C", "left", n // 2, False, False)
    sim.set_diffusion_rate_particle("A", "background", 3e-12)
    sim.set_transition_rate_particle("A", "background", "left", 1e-12)
    sim.set_transition_rate_particle("B", "left", "background", 2e-12)
    sim.add_reaction(["A", "B"], ["C"], k_fwd)
    sim.add_reaction(["C"], ["A", "B"], k_back)
    sim.add_reaction(["C"], ["C", "C"], 1.0)
    if zero_order:
        sim.add_reaction([], ["A"], zero_order)
    return sim


def test_dense_moderate(oracle_mod):
    announce("dense 24^3 S=8 density 0.6, two regions with transitions, 200 steps")
    sim = _dense_model(24, 8, 0.6, seed=11)
    orc = oracle_mod.Oracle.from_simulation(sim)
    run_parity(sim, orc, default_itime_schedule(11, 200), 50, "dense")


def test_zero_order_creation(oracle_mod):
    """A zero-order reaction fires in empty voxels too (all reactant slots padded with species 0)."""
    announce("zero-order creation 16^3 S=4, 150 steps")
    sim = _dense_model(16, 4, 0.05, seed=3, zero_order=2.0)
    orc = oracle_mod.Oracle.from_simulation(sim)
    run_parity(sim, orc, default_itime_schedule(3, 150), 50, "zero_order")


def test_odd_dims_and_y0_row(oracle_mod):
    """Non-cubic odd extents; the y == 0 rows use X % 0 (hazard H4)."""
    announce("odd dims 7x5x9 S=4, 200 steps")
    lam = 32e-9
    sim = Simulation(LatticeParams([7 * lam, 5 * lam, 9 * lam], [7, 5, 9], 50e-6, lam, 4))
    sim.set_seed(8)
    sim.prepare_regions()
    sim.add_particle_count("A", "background", 16, True, False)
    sim.add_particle_count("B", "background", 8, True, False)
    sim.set_diffusion_rate_particle("A", "background", 5e-12)
    sim.set_diffusion_rate_particle("B", "background", 2e-12)
    sim.add_reaction(["A", "B"], ["B"], 300.0)
    sim.add_reaction(["B"], ["A", "B"], 20.0)
    orc = oracle_mod.Oracle.from_simulation(sim)
    run_parity(sim, orc, default_itime_schedule(8, 200), 20, "odd")


def test_forced_hops(oracle_mod):
    """p[0] = 1 for species A (every A hops -x) and p = [0,0,0,0,0,1] for B (every B hops +z)."""
    announce("forced hops 6x4x5 S=4, 12 steps")
    lam = 32e-9
    tau = 50e-6
    d_one = lam * lam / tau  # D*tau/lambda^2 == 1 exactly in f32 for these values
    sim = Simulation(LatticeParams([6 * lam, 4 * lam, 5 * lam], [6, 4, 5], tau, lam, 4))
    sim.set_seed(2)
    sim.prepare_regions()
    sim.add_particle_count("A", "background", 10, True, False)
    sim.add_particle_count("B", "background", 10, True, False)
    sim.set_diffusion_rate_particle("A", "background", d_one)
    sim.set_diffusion_rate_particle("B", "background", 0.0)
    sim.set_transition_rate_particle("B", "background", "background", 0.0)
    orc = oracle_mod.Oracle.from_simulation(sim)
    run_parity(sim, orc, default_itime_schedule(2, 12), 1, "forced")


def test_stat_samples_follow_write_freq(oracle_mod):
    announce("statistics cadence on spherical_cell, write_freq 25, 101 steps")
    sim = configs.spherical_cell(seed=42)
    orc = oracle_mod.Oracle.from_simulation(sim)
    sim.prepare_for_gpu(0)
    its = default_itime_schedule(42, 101)
    expected = []
    for f in range(101):
        if f % 25 == 0 and f > 0:
            for sp, name in ((1, "A"), (2, "B"), (3, "C")):
                expected.append((name, int(orc.stat[sp]), f))
        sim.step(f, int(its[f]), 25)
        orc.step(int(its[f]))
    got = [(s.name, s.value, s.iteration_count) for s in sim.stats]
    assert got == expected


def test_region_counts_match_lattice(oracle_mod):
    announce("region counts, large_cell 40^3, 30 steps")
    sim = configs.large_cell(seed=77, edge=40, n_particles=8_000)
    sim.prepare_for_gpu(0)
    sim.run(0, default_itime_schedule(77, 30))
    lat = sim.lattice().reshape(-1, sim.info()["max_particles_site"])
    reg = sim.regions().ravel()
    n1 = sim.info()["num_species"] + 1
    exp = np.zeros((sim.info()["num_regions"], n1), dtype=np.int64)
    for s in range(lat.shape[1]):
        np.add.at(exp, (reg, lat[:, s].astype(np.int64)), 1)
    exp[:, 0] = 0
    got = sim.region_counts()
    got[:, 0] = 0
    assert np.array_equal(got, exp)
    assert np.array_equal(sim.species_counts()[1:], exp.sum(axis=0)[1:])


@pytest.mark.parametrize("kernel", ["wave", "wave-overflow", "wave-summary", "wave-bitmapsort", "wave-splittab",
                                    "wave-globaltab", "wave-replayblock", "wave-replaylane", "tiled", "simple"])
def test_dense_capacity_replay(oracle_mod, kernel):
    """S=4 at high density: claims into full voxels are rejected (rdme.wgsl:62, 73-81) or land in the next
    voxel's slot 0 (rdme.wgsl:30-40), after which the reference's occupancy word and counts leave the slots.
    Step by step to the end against the oracle's whole state and event counters."""
    import os

    for walled, dens, nsteps in ((False, 1.4, 30), (True, 0.7, 40)):
        announce(f"dense capacity replay ({kernel}), 20^3 S=4 density {dens} wall={walled}, step by step")
        os.environ["CS_STEP_KERNEL"] = kernel.split("-")[0]
        if kernel == "wave-overflow":
            os.environ["CS_WAVE_EVENT_CAP"] = "2"  # most planes overflow and take the global form
        if kernel == "wave-summary":
            os.environ["CS_OCC_SUMMARY"] = "on"  # the replay's outputs must mark the occupancy summary
        if kernel == "wave-bitmapsort":
            os.environ["CS_REPLAY_LDS_SORT"] = "1"  # every noted list sorted through the replay's bitmaps
        if kernel == "wave-splittab":
            os.environ["CS_WAVE_GLOBAL_TABLES"] = "1"  # the reaction tables read from global memory
        if kernel == "wave-globaltab":
            os.environ["CS_WAVE_GLOBAL_TABLES"] = "2"  # every model table read from global memory
        if kernel == "wave-replayblock":
            os.environ["CS_REPLAY_FORMS"] = "block"  # every cluster of up to 64 noted voxels replayed by the block
        if kernel == "wave-replaylane":
            os.environ["CS_REPLAY_FORMS"] = "lane"  # every cluster replayed by one lane, arrays in global memory
        try:
            sim = _walled_model(20, 4, dens, seed=31) if walled else _dense_model(20, 4, dens, seed=31)
            orc = oracle_mod.Oracle.from_simulation(sim)
            sim.prepare_for_gpu(0)
        finally:
            os.environ.pop("CS_STEP_KERNEL", None)
            os.environ.pop("CS_WAVE_EVENT_CAP", None)
            os.environ.pop("CS_OCC_SUMMARY", None)
            os.environ.pop("CS_REPLAY_LDS_SORT", None)
            os.environ.pop("CS_WAVE_GLOBAL_TABLES", None)
            os.environ.pop("CS_REPLAY_FORMS", None)
        skew = 0
        for t, it in enumerate(default_itime_schedule(31, nsteps)):
            sim.step(t, int(it), 0)
            orc.step(int(it))
            check_state(sim, orc, f"dense replay step {t}")
            skew = max(skew, skewed_voxels(orc))
        c = check_counters(sim, orc)
        assert c[CS_CNT_TIGHT] > 0 and orc.counters.failed_claims > 0, "the case must exercise the replay"
        if not walled:
            assert orc.counters.alias_writes > 0 and skew > 0, "the case must leave slot-S aliases"
        print(f"oracle {orc.counters.as_dict()}; tight {int(c[CS_CNT_TIGHT])}; skewed voxels up to {skew}")


def _big_network(edge, seed, n_species=160, n_reactions=250):
    """Two regions, n_species lattice species and n_reactions first- and second-order reactions: the tables
    (stoichiometry alone (n_reactions + 1) x (n_species + 1) bytes) exceed what the wave kernel stages in LDS."""
    rng = np.random.default_rng(seed)
    lam = 32e-9
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, 50e-6, lam, 8))
    sim.set_seed(seed)
    sim.add_region(RegionType.Sphere("core", [edge * lam / 2] * 3, edge * lam / 3), 1e-12)
    sim.prepare_regions()
    names = [f"P{i:03d}" for i in range(n_species)]
    for i, nm in enumerate(names):
        sim.add_particle_count(nm, "core" if i % 2 else "background", 40, True, False)
        sim.set_diffusion_rate_particle(nm, "background", float(rng.uniform(1e-13, 4e-12)))
        sim.set_diffusion_rate_particle(nm, "core", float(rng.uniform(1e-13, 4e-12)))
        sim.set_transition_rate_particle(nm, "background", "core", 1e-12)
        sim.set_transition_rate_particle(nm, "core", "background", 1e-12)
    for j in range(n_reactions):
        a = names[int(rng.integers(n_species))]
        if j % 3:
            sim.add_reaction([a], [names[int(rng.integers(n_species))]], float(rng.uniform(50.0, 2000.0)))
        else:
            b = names[int(rng.integers(n_species))]
            sim.add_reaction([a, b], [names[int(rng.integers(n_species))]], float(rng.uniform(50.0, 2000.0)))
    return sim


def test_tables_past_lds_read_from_global_memory(oracle_mod):
    """A 160-species, 250-reaction network: the tables (about 48 KB) do not fit the wave kernel's LDS beside the
    plane records, so it stages the hop tables and reads the reaction tables from global memory. Against the
    oracle's whole state."""
    announce("160 species x 250 reactions (tables past LDS), 24^3 S=8, 60 steps")
    sim = _big_network(24, seed=8)
    orc = oracle_mod.Oracle.from_simulation(sim)
    sim.prepare_for_gpu(0)
    for t, it in enumerate(default_itime_schedule(8, 60)):
        sim.step(t, int(it), 0)
        orc.step(int(it))
    check_state(sim, orc, "tables past LDS after 60 steps")
    c = check_counters(sim, orc)
    assert orc.counters.fired > 20, "the case must fire reactions"
    print(f"oracle {orc.counters.as_dict()}; device counters {c.tolist()}")


def test_dense_s16_full_heads(oracle_mod):
    """S=16 at about 7 particles per voxel: many voxels hold 8 or more, past the 8-slot head the wave kernel
    stages, so their hops and gathers take the whole record from memory. Step by step to the end against the
    oracle's whole state (slot-S aliases included)."""
    announce("dense S=16 16^3 density 7 (full 8-slot heads), step by step")
    sim = _dense_model(16, 16, 7.0, seed=17)
    orc = oracle_mod.Oracle.from_simulation(sim)
    sim.prepare_for_gpu(0)
    lat = sim.lattice()
    assert (np.count_nonzero(lat, axis=-1) >= 8).sum() > 100, "the case must hold voxels past the head"
    for t, it in enumerate(default_itime_schedule(17, 12)):
        sim.step(t, int(it), 0)
        orc.step(int(it))
        check_state(sim, orc, f"dense S=16 step {t}")
    check_counters(sim, orc)
    print(f"oracle {orc.counters.as_dict()}")


def test_more_than_1024_tight_voxels_in_one_step(oracle_mod):
    """A 32^3 S=4 lattice at 2.2 particles per voxel: thousands of tight voxels per step (the replay's list
    and its clusters), slot-S aliases and rejected claims, against the oracle's whole state step by step."""
    announce("32^3 S=4 density 2.2: >1024 tight voxels per step, 8 steps")
    sim = _dense_model(32, 4, 2.2, seed=5)
    orc = oracle_mod.Oracle.from_simulation(sim)
    sim.prepare_for_gpu(0)
    most = 0
    for t, it in enumerate(default_itime_schedule(5, 8)):
        before = int(sim.counters()[CS_CNT_TIGHT])
        sim.step(t, int(it), 0)
        orc.step(int(it))
        most = max(most, int(sim.counters()[CS_CNT_TIGHT]) - before)
        check_state(sim, orc, f"32^3 dense step {t}")
    check_counters(sim, orc)
    assert most > 1024, most
    assert orc.counters.alias_writes > 0
    print(f"most tight voxels in a step: {most}; oracle {orc.counters.as_dict()}")


def test_homodimer_negative_counts(oracle_mod):
    """k*n^2 of a homodimer row (cme.wgsl:42-44) fires on a single particle and takes its count to -1
    (cme.wgsl:85-91); the negative count stays in C[v][sp], enters later propensities (a negative product) and
    the occupancy word drops below the slot count. The HIP path keeps both as offsets of the voxel."""
    announce("homodimer A + A -> B on single particles, 24^3 S=8, 80 steps")
    lam = 32e-9
    edge = 24
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, 50e-6, lam, 8))
    sim.set_seed(44)
    sim.prepare_regions()
    sim.add_particle_count("A", "background", 3000, True, False)
    sim.add_particle_count("B", "background", 500, True, False)
    sim.add_particle_count("C", "background", 200, True, False)
    sim.set_diffusion_rate_particle("A", "background", 2e-12)
    sim.set_diffusion_rate_particle("B", "background", 1e-12)
    sim.set_diffusion_rate_particle("C", "background", 3e-12)
    sim.add_reaction(["A", "A"], ["B"], 900.0)
    sim.add_reaction(["B"], ["A", "C"], 300.0)
    sim.add_reaction(["A", "C"], ["C", "C"], 400.0)
    sim.add_reaction(["C", "C"], ["A"], 200.0)
    orc = oracle_mod.Oracle.from_simulation(sim)
    sim.prepare_for_gpu(0)
    neg = 0
    for t, it in enumerate(default_itime_schedule(44, 80)):
        sim.step(t, int(it), 0)
        orc.step(int(it))
        check_state(sim, orc, f"homodimer step {t}")
        neg = max(neg, int(np.count_nonzero(orc.counts_array()[:, 1:] < 0)))
    c = check_counters(sim, orc)
    assert neg > 10, "the case must take counts below zero"
    assert c[CS_CNT_FIRED] > 0
    print(f"negative per-voxel counts: up to {neg}; oracle {orc.counters.as_dict()}")


def _walled_model(edge, slots, density, seed):
    """Every voxel's slot 0 holds an immobile, unreactive species, so a claim into a full voxel finds the
    next voxel's slot 0 taken and is rejected (never the slot-S alias)."""
    lam = 32e-9
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, 50e-6, lam, slots))
    sim.set_seed(seed)
    sim.add_region(RegionType.Cube("left", [0, 0, 0], [edge * lam / 2, edge * lam, edge * lam]), 4e-12)
    sim.prepare_regions()
    sim.fill_region("W", "background", False)
    sim.fill_region("W2", "left", False)  # "left" voxels: fill_region targets a region, so fill it too
    n = int(density * edge ** 3 / 2)
    sim.add_particle_count("A", "background", n, True, False)
    sim.add_particle_count("B", "left", n // 2, True, False)
    sim.set_diffusion_rate_particle("W", "background", 0.0)
    sim.set_diffusion_rate_particle("W", "left", 0.0)
    sim.set_diffusion_rate_particle("W2", "left", 0.0)
    sim.set_diffusion_rate_particle("W2", "background", 0.0)
    sim.set_diffusion_rate_particle("A", "background", 6e-12)
    sim.set_transition_rate_particle("A", "background", "left", 3e-12)
    sim.set_transition_rate_particle("B", "left", "background", 3e-12)
    sim.add_reaction(["A", "B"], ["B", "B"], 40.0)
    sim.add_reaction(["B"], ["A"], 8.0)
    return sim


def _run_with_env(env, build, seed, nsteps):
    import os

    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        sim = build()
        sim.prepare_for_gpu(0)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    sim.run(0, default_itime_schedule(seed, nsteps))
    return sim.lattice(), sim.counters()


@pytest.mark.parametrize("kernel", ["wave", "wave-summary", "tiled"])
@pytest.mark.parametrize("planes", ["1", "3", "16", "1000"])
def test_tiled_kernel_matches_simple_kernel(planes, kernel):
    """The wave-march and block-tiled kernels give the same lattices as the one-lane-per-voxel kernel for
    any chunk length along x, on a lattice whose y and z extents are not multiples of the tiles (4 x 64).
    "wave-summary" keeps the occupancy summaries (planes skipped when their tiles are empty) although this
    lattice would not."""
    announce(f"{kernel} (planes={planes}) vs simple, 37x30x70 S=8, 40 steps")

    def build():
        lam = 32e-9
        sim = Simulation(LatticeParams([37 * lam, 30 * lam, 70 * lam], [37, 30, 70], 50e-6, lam, 8))
        sim.set_seed(21)
        sim.add_region(RegionType.Sphere("ball", [18 * lam, 15 * lam, 35 * lam], 12 * lam), 2e-12)
        sim.prepare_regions()
        sim.add_particle_count("A", "background", 9000, True, False)
        sim.add_particle_count("B", "ball", 4000, True, False)
        sim.add_particle_count("C", "ball", 1000, False, False)
        sim.set_diffusion_rate_particle("A", "background", 4e-12)
        sim.set_transition_rate_particle("A", "background", "ball", 1e-12)
        sim.set_transition_rate_particle("B", "ball", "background", 2e-12)
        sim.add_reaction(["A", "B"], ["C"], 80.0)
        sim.add_reaction(["C"], ["A", "B"], 9.0)
        return sim

    env = {"CS_STEP_KERNEL": kernel.split("-")[0], "CS_TILE_PLANES": planes}
    env["CS_OCC_SUMMARY"] = "on" if kernel == "wave-summary" else "off"
    a, ca = _run_with_env(env, build, 21, 40)
    b, cb = _run_with_env({"CS_STEP_KERNEL": "simple"}, build, 21, 40)
    assert np.array_equal(a, b)
    assert np.array_equal(ca[:5], cb[:5])


def test_large_cell_256_first_steps(oracle_mod):
    """Full 256^3 S=16 workload: two steps against the oracle, then size-independent properties."""
    announce("large_cell 256^3 S=16, 2 steps vs oracle + 20 steps properties")
    sim = configs.large_cell(seed=1234)
    orc = oracle_mod.Oracle.from_simulation(sim)
    run_parity(sim, orc, default_itime_schedule(1234, 2), 1, "large_cell_256")
    del orc
    its = default_itime_schedule(1234, 20, first=2)
    start = sim.lattice()
    sim.run(2, its)
    a = sim.lattice()
    # determinism: the same source state and seeds give the same lattice
    sim.write_lattice(start)
    sim.run(2, its)
    b = sim.lattice()
    assert np.array_equal(a, b)
    flat = a.reshape(-1, a.shape[-1])
    nz = flat != 0
    # compacted: no empty slot before an occupied one
    assert not np.any(~nz[:, :-1] & nz[:, 1:])
    assert flat.max() <= sim.info()["num_species"]
    assert sim.counters()[CS_CNT_TIGHT] == 0


def test_sparse_region_and_random_walk_model(oracle_mod):
    """A model built with add_sparse_region and particle_random_walk (simulation.rs:483-535, 842-864) steps
    bit-exactly: the labels of the sparse balls and the chain of walk particles are ordinary model state."""
    announce("sparse region + random walk 64^3 S=16, 120 steps, seed 8")
    lam = 32e-9
    edge = 64
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, 50e-6, lam, 16))
    sim.set_seed(8)
    sim.add_region(RegionType.Sphere("cell", [edge * lam / 2] * 3, 0.45 * edge * lam), 1e-12)
    sim.prepare_regions()
    sim.particle_random_walk("W", "cell", 30 * lam, 1.2 * lam, lam, True)
    sim.add_sparse_region("granules", RegionType.Sphere("g", [0, 0, 0], 2 * lam), "cell", 300, 1e-13)
    sim.add_particle_count("E", "granules", 200, True, False)
    sim.add_particle_count("F", "cell", 1500, True, False)
    sim.set_transition_rate_particle("F", "cell", "granules", 5e-13)
    sim.add_reaction(["E", "F"], ["W"], 3000.0)
    sim.add_reaction(["W"], ["E", "F"], 50.0)
    orc = oracle_mod.Oracle.from_simulation(sim)
    run_parity(sim, orc, default_itime_schedule(8, 120), 40, "sparse_walk")
    assert orc.counters.fired > 0 and orc.counters.moves > 0
