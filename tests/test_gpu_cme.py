"""Bit-exact parity of the HIP well-mixed (CME) phase and the hybrid step against the oracle.

The CME phase has no reference counterpart; its definition is the oracle's (cs_oracle.h). These cases run
the hybrid gene-expression model (well-mixed gene switching, transcription of lattice mRNA into the
nucleoid, lattice translation and decay) and a 40-species / 120-reaction well-mixed network whose rates span
ten decades, and require the lattice, the well-mixed counts and the event counters to match the oracle
exactly. The second case also reads the device's f64 audit of the compensated f32 propensities.
"""
import numpy as np
import pytest

from cell_simulation_amd import LatticeParams, RegionType, Simulation, configs, default_itime_schedule
from cell_simulation_amd._lib import CS_CNT_CME_EVENTS, CS_CNT_CME_NEAR, CS_CNT_EMIT_FAIL, CS_CNT_EXCEPTION
from tests.conftest import announce, check_state

pytestmark = pytest.mark.gpu


def _compare(sim, orc, tag):
    g = sim.lattice()
    o = orc.lattice_u8()
    if not np.array_equal(g, o):
        diff = np.argwhere(np.any(g != o, axis=-1))
        v = tuple(diff[0])
        raise AssertionError(f"{tag}: {len(diff)} voxels differ; first {v}: gpu {g[v]} oracle {o[v]}")
    assert np.array_equal(sim.cme_counts(), orc.cme_counts), (tag, sim.cme_counts(), orc.cme_counts)


def test_gene_expression_hybrid_parity(oracle_mod):
    announce("gene_expression 32^3 S=8, rates x20000, 4 genes, 300 steps, seed 7")
    sim = configs.gene_expression(seed=7, edge=32, rate_scale=20000.0, gene_copies=4)
    orc = oracle_mod.Oracle.from_simulation(sim)
    sim.prepare_for_gpu(0)
    its = default_itime_schedule(7, 300)
    for k in range(0, 300, 100):
        sim.run(k, its[k:k + 100])
        orc.run(its[k:k + 100])
        _compare(sim, orc, f"after {k + 100} steps")
    c = sim.counters()
    assert c[CS_CNT_CME_EVENTS] == orc.cme_counters.events > 10
    assert c[CS_CNT_EMIT_FAIL] == orc.cme_counters.emit_fail
    assert c[CS_CNT_CME_NEAR] == orc.cme_counters.near
    assert c[CS_CNT_EXCEPTION] == 0
    counts = sim.species_counts()
    assert counts[1] > 0 and np.array_equal(counts[1:], orc.stat[1:])


def _precision_network(seed=2024, nc=40, ncr=120, edge=8):
    rng = np.random.Generator(np.random.PCG64(seed))
    lam = 32e-9
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, 5e-6, lam, 8))
    sim.set_seed(seed)
    sim.add_region(RegionType.Sphere("cell", [edge * lam / 2] * 3, 3.5 * lam), 0.0)
    sim.prepare_regions()
    sim.add_particle_count("R", "cell", 0, True, False)
    names = [f"c{i:02d}" for i in range(nc)]
    for nm in names:
        sim.add_cme_species(nm, int(rng.integers(50, 5000)))
    for r in range(ncr):
        order = int(rng.integers(1, 3))
        reac = [names[int(i)] for i in rng.integers(0, nc, size=order)]
        prods = [names[int(i)] for i in rng.integers(0, nc, size=int(rng.integers(0, 3)))]
        if r % 17 == 0:
            prods.append("R@cell")
        k = float(np.exp(rng.uniform(np.log(1e-6), np.log(1e4))))
        # keep the event rate per step modest: bimolecular rows see count products
        if order == 2:
            k *= 1e-3
        sim.add_cme_reaction(reac, prods, k)
    return sim


def test_precision_network_parity_and_audit(oracle_mod):
    announce("well-mixed network 40 species x 120 reactions, rates 1e-6..1e4, 60 steps")
    sim = _precision_network()
    orc = oracle_mod.Oracle.from_simulation(sim)
    sim.prepare_for_gpu(0)
    sim.set_cme_audit(True)
    its = default_itime_schedule(2024, 60)
    sim.run(0, its)
    orc.run(its)
    _compare(sim, orc, "precision network after 60 steps")
    c = sim.counters()
    assert c[CS_CNT_CME_EVENTS] == orc.cme_counters.events > 100
    a = sim.cme_audit()
    assert a["events"] == c[CS_CNT_CME_EVENTS]
    assert a["max_rel_error"] < 2.5e-7, a
    assert a["divergences"] <= max(1, a["events"] // 1000), a


def test_precision_stress_small_parity(oracle_mod):
    announce("precision_stress 48^3 S=8 (12 lattice + 40 well-mixed species, 120 reactions), 6000 particles, 40 steps")
    sim = configs.precision_stress(edge=48, n_particles=6000)
    orc = oracle_mod.Oracle.from_simulation(sim)
    sim.prepare_for_gpu(0)
    sim.set_cme_audit(True)
    its = default_itime_schedule(2024, 40)
    sim.run(0, its)
    orc.run(its)
    _compare(sim, orc, "precision_stress 48^3 after 40 steps")
    check_state(sim, orc, "precision_stress 48^3 after 40 steps")
    c = sim.counters()
    assert c[CS_CNT_CME_EVENTS] == orc.cme_counters.events > 0
    a = sim.cme_audit()
    assert a["events"] == c[CS_CNT_CME_EVENTS] and a["max_rel_error"] < 2.5e-7, a


def test_ensemble_gene_switching_matches_closed_form():
    """400 replicate cells of a two-state gene: the fraction switched on after T is
    k_on / (k_on + k_off) * (1 - exp(-(k_on + k_off) T)) within 3 binomial sigma (an f64 closed form)."""
    import math

    from cell_simulation_amd import Ensemble

    announce("ensemble of 400 two-state gene cells, 40 steps of 1 ms, against the closed form")
    lam = 32e-9
    cells, seeds = [], []
    k_on, k_off, n_steps, tau = 50.0, 30.0, 40, 1e-3
    for i in range(400):
        sim = Simulation(LatticeParams([4 * lam] * 3, [4] * 3, tau, lam, 4))
        sim.set_seed(1000 + i)
        sim.add_region(RegionType.Sphere("cell", [2 * lam] * 3, 1.5 * lam), 0.0)
        sim.prepare_regions()
        sim.add_cme_species("off", 1)
        sim.add_cme_species("on", 0)
        sim.add_cme_reaction(["off"], ["on"], k_on)
        sim.add_cme_reaction(["on"], ["off"], k_off)
        cells.append(sim)
        seeds.append(77 + 13 * i)
    ens = Ensemble(cells, seeds)
    ens.run_graph(n_steps, block=20)
    on = sum(int(ens.cme_counts(i)[1]) for i in range(len(cells)))
    p = k_on / (k_on + k_off) * (1.0 - math.exp(-(k_on + k_off) * n_steps * tau))
    sigma = math.sqrt(p * (1 - p) / len(cells))
    assert abs(on / len(cells) - p) < 3 * sigma, (on / len(cells), p, sigma)
