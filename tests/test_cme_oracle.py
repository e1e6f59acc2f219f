"""The well-mixed (CME) phase of the oracle against closed forms (no reference counterpart exists).

The CME phase is the build's own extension (cs_oracle.h documents its arithmetic); what pins it here is
mathematics, not a reference run:
* combinatorial factors h against binomial coefficients;
* a single two-state gene: P(on at T) = k_on / (k_on + k_off) * (1 - exp(-(k_on + k_off) T)), checked over
  independent replicates within 3 sigma of the binomial spread;
* a birth process into a region: the number of lattice products after T is Poisson(k T) (mean and
  variance within 3 sigma), every product lands in the target region and the lattice stays compacted;
* the compensated f32 total and choice against plain f64 on networks whose rates span ten decades.
A planted mistake in the oracle (a dropped reactant, a wrong waiting-time scale, a skipped emission)
moves one of these statistics by many sigma.
"""
import ctypes as C
import math

import numpy as np
import pytest

from cell_simulation_amd import LatticeParams, RegionType, Simulation
from cell_simulation_amd._lib import CellSimError


def _box(tau=1e-3, edge=6, slots=8):
    lam = 32e-9
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, tau, lam, slots))
    sim.set_seed(5)
    sim.add_region(RegionType.Sphere("cell", [edge * lam / 2] * 3, 2.2 * lam), 0.0)
    sim.prepare_regions()
    return sim


def test_cme_h_matches_binomials(oracle_mod):
    lib = oracle_mod.lib()
    counts = np.array([7, 5, 0, 1], dtype=np.int32)

    def h(row):
        r = np.array(row, dtype=np.int32)
        return lib.cso_cme_h(r.ctypes.data, counts.ctypes.data)

    assert h([0, -1, -1]) == 7.0
    assert h([0, 0, -1]) == math.comb(7, 2)
    assert h([0, 0, 0]) == math.comb(7, 3)
    assert h([0, 1, -1]) == 35.0
    assert h([1, 0, 1]) == math.comb(5, 2) * 7
    assert h([2, -1, -1]) == 0.0
    assert h([3, 3, -1]) == 0.0  # one molecule cannot pair with itself
    assert h([-1, -1, -1]) == 1.0  # zero order


def test_two_state_gene_probability(oracle_mod):
    sim = _box()
    sim.add_cme_species("off", 1)
    sim.add_cme_species("on", 0)
    k_on, k_off = 50.0, 30.0
    sim.add_cme_reaction(["off"], ["on"], k_on)
    sim.add_cme_reaction(["on"], ["off"], k_off)
    orc = oracle_mod.Oracle.from_simulation(sim)
    n_rep, n_steps, tau = 3000, 40, 1e-3
    on = 0
    for rep in range(n_rep):
        orc.cme_counts[:] = (1, 0)
        orc.cme.seed = (0x1234567 + 7919 * rep) & 0xFFFFFFFF
        for t in range(n_steps):
            orc.cme_phase(0x9E3779B1 * (t + 1) & 0xFFFFFFFF)
        assert orc.cme_counts.sum() == 1
        on += int(orc.cme_counts[1])
    T = n_steps * tau
    p = k_on / (k_on + k_off) * (1.0 - math.exp(-(k_on + k_off) * T))
    sigma = math.sqrt(p * (1 - p) / n_rep)
    assert abs(on / n_rep - p) < 3 * sigma, (on / n_rep, p, sigma)
    assert orc.cme_counters.near == 0 and orc.cme_counters.exceptions == 0


def test_transcription_is_poisson_into_region(oracle_mod):
    sim = _box()
    sim.add_particle_count("mRNA", "cell", 0, True, False)
    sim.add_cme_species("gene_on", 1)
    k = 200.0
    sim.add_cme_reaction(["gene_on"], ["gene_on", "mRNA@cell"], k)
    cell = sim.region_index("cell")
    n_rep, n_steps, tau = 600, 40, 1e-3
    totals = []
    for rep in range(n_rep):
        orc = oracle_mod.Oracle.from_simulation(sim)
        orc.cme.seed = (0xBEEF + 104729 * rep) & 0xFFFFFFFF
        for t in range(n_steps):
            orc.cme_phase((2654435761 * (t + 3)) & 0xFFFFFFFF)
        lat = orc.lattice.reshape(-1, orc.S)
        occ = np.count_nonzero(lat, axis=1)
        # compacted: no empty slot before an occupied one
        assert np.all(np.cumsum(lat == 0, axis=1)[lat != 0] == 0)
        assert np.array_equal(occ, orc.occupancy)
        where = np.nonzero(occ)[0]
        assert np.all(orc.regions[where] == cell)
        assert int(orc.stat[1]) == int(occ.sum())
        totals.append(int(occ.sum()))
    lam = k * n_steps * tau
    x = np.array(totals, dtype=np.float64)
    assert abs(x.mean() - lam) < 3 * math.sqrt(lam / n_rep), (x.mean(), lam)
    var_sigma = math.sqrt(lam / n_rep + 2 * lam * lam / (n_rep - 1))
    assert abs(x.var(ddof=1) - lam) < 3 * var_sigma, (x.var(ddof=1), lam)


def test_emission_into_full_region_is_dropped(oracle_mod):
    sim = _box(edge=4, slots=4)
    sim.add_particle_count("X", "cell", 0, True, False)
    for k in range(4):  # fill_region places one particle per voxel
        sim.fill_region(f"F{k}", "cell", False)
    sim.add_cme_species("src", 1)
    sim.add_cme_reaction(["src"], ["src", "X@cell"], 1e4)
    orc = oracle_mod.Oracle.from_simulation(sim)
    before = orc.lattice.copy()
    orc.cme_phase(12345)
    assert orc.cme_counters.events > 0
    assert orc.cme_counters.emit_fail == orc.cme_counters.events
    assert np.array_equal(orc.lattice, before)


def _random_network(rng, nc, ncr):
    reac = np.full((ncr, 3), -1, dtype=np.int32)
    for r in range(ncr):
        order = rng.integers(0, 3)
        reac[r, :order] = rng.integers(0, nc, size=order)
    rates = np.exp(rng.uniform(np.log(1e-6), np.log(1e4), size=ncr)).astype(np.float32)
    return reac, rates


def test_compensated_f32_against_f64(oracle_mod):
    lib = oracle_mod.lib()
    rng = np.random.Generator(np.random.PCG64(2024))
    nc, ncr = 40, 120
    reac, rates = _random_network(rng, nc, ncr)
    cm = oracle_mod.CmeModel()
    zeros = np.zeros(ncr * max(nc, 4), dtype=np.int32)
    cm.nc, cm.ncr = nc, ncr
    cm.reactants, cm.rates = reac.ctypes.data, rates.ctypes.data
    cm.delta = cm.emit = cm.emit_region = zeros.ctypes.data
    worst, diverged, trials = 0.0, 0, 0
    for _ in range(200):
        counts = rng.integers(0, 2000, size=nc).astype(np.int32)
        t32 = lib.cso_cme_total32(C.byref(cm), counts.ctypes.data)
        t64 = lib.cso_cme_total64(C.byref(cm), counts.ctypes.data)
        worst = max(worst, abs(t32 - t64) / t64)
        for u2 in rng.uniform(0, 1, size=50):
            trials += 1
            diverged += lib.cso_cme_choice32(C.byref(cm), counts.ctypes.data, u2) != \
                lib.cso_cme_choice64(C.byref(cm), counts.ctypes.data, u2)
    # every term is one f32 product of an f32 rate and an f32-rounded h: relative error <= ~2^-23
    assert worst < 2.5e-7, worst
    assert diverged / trials < 1e-3, (diverged, trials)


def test_cme_model_validation():
    sim = _box()
    sim.add_particle_count("P", "cell", 0, True, False)
    sim.add_cme_species("G", 1)
    with pytest.raises(CellSimError):
        sim.add_cme_species("P", 1)  # lattice name
    with pytest.raises(CellSimError):
        sim.add_cme_reaction(["nope"], [], 1.0)
    with pytest.raises(CellSimError):
        sim.add_cme_reaction(["G"], ["P"], 1.0)  # lattice product without a region
    with pytest.raises(CellSimError):
        sim.add_cme_reaction(["G"], ["P@nowhere"], 1.0)
    with pytest.raises(CellSimError):
        sim.add_cme_reaction(["G"], ["G"], -1.0)
    sim.add_cme_reaction(["G"], ["G", "P@cell"], 1.0)
    t = sim.cme_tables()
    assert t["emit"][0, 0] == sim.particle_index("P") and t["emit_region"][0] == sim.region_index("cell")
    assert list(t["delta"][0]) == [0]
