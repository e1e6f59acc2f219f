"""Batched ensemble against the serial oracle and against plain per-cell simulations, bit for bit.

Each ensemble cell at step t uses itime = default_itime_schedule(itime_seed)[t], so it must evolve exactly as
the oracle of that cell run with that schedule, and as the same cell built as a plain Simulation and stepped
alone: lattices, per-voxel counts, occupancy words and well-mixed counts are compared for every cell after
eager steps, after graph-replayed blocks, under different tilings (queue batch sizes, the queue-free
round-robin, planes per item) and across an autotune, which must leave the state as it found it. Dense cells
exercise the per-cell replay of tight voxels (rejected claims, slot-S aliases) and reservoir cells the skewed
voxels' offsets; the 512-cell gene-expression workload is checked for conservation, compaction, region
confinement and against the oracle on its smallest cells.
"""
import numpy as np
import pytest

from cell_simulation_amd import Ensemble, LatticeParams, RegionType, Simulation, configs, default_itime_schedule
from cell_simulation_amd._lib import CS_CNT_ALIAS, CS_CNT_EXCEPTION, CS_CNT_FAILED_CLAIM, CS_CNT_TIGHT
from tests.conftest import announce

pytestmark = pytest.mark.gpu

N_CELLS = 7


def _cells():
    return configs.ensemble_cells(N_CELLS, seed=11, edge_range=(16, 72), rate_scale=20000.0, tau=20e-6)


def _plain(cells, seeds, nsteps):
    out = []
    for c, s in zip(cells, seeds):
        c.prepare_for_gpu(0)
        c.run(0, default_itime_schedule(s, nsteps))
        out.append((c.lattice(), c.cme_counts()))
    return out


def _check(ens, ref, tag):
    for i, (lat, cme) in enumerate(ref):
        g = ens.lattice(i)
        if not np.array_equal(g, lat):
            diff = np.argwhere(np.any(g != lat, axis=-1))
            raise AssertionError(f"{tag}: cell {i} {len(diff)} voxels differ, first {tuple(diff[0])}")
        assert np.array_equal(ens.cme_counts(i), cme), (tag, i, ens.cme_counts(i), cme)


def test_ensemble_matches_plain_cells():
    announce(f"ensemble of {N_CELLS} gene-expression cells (edges 16..72), 60 eager + 40 graph steps")
    cells, seeds = _cells()
    ens = Ensemble(cells, seeds)
    ref_cells, _ = _cells()
    ens.run(60)
    _check(ens, _plain(ref_cells, seeds, 60), "eager 60")
    ens.run_graph(40, block=20)
    ref_cells, _ = _cells()
    _check(ens, _plain(ref_cells, seeds, 100), "graph 100")
    assert ens.info()["steps_done"] == 100
    assert ens.counters()[CS_CNT_EXCEPTION] == 0


@pytest.mark.parametrize("tiling", [dict(xc=8, blocks_per_cu=2, batch=0), dict(xc=32, blocks_per_cu=4, batch=4),
                                    dict(xc=5, blocks_per_cu=1, batch=1), dict(xc=60, blocks_per_cu=2, batch=0),
                                    dict(xc=61, blocks_per_cu=3, batch=1)])
def test_ensemble_tilings_agree(tiling):
    announce(f"ensemble tiling {tiling}, 30 steps")
    cells, seeds = _cells()
    ens = Ensemble(cells, seeds)
    ens.set_tiling(**tiling)
    ens.run_graph(30, block=10)
    ref_cells, _ = _cells()
    _check(ens, _plain(ref_cells, seeds, 30), f"tiling {tiling}")


def test_autotune_keeps_state():
    announce("ensemble autotune (2 trial steps per tiling) then 20 steps")
    cells, seeds = _cells()
    ens = Ensemble(cells, seeds)
    ens.run(10)
    ms = ens.autotune(2)
    assert ms > 0 and ens.info()["steps_done"] == 10
    ens.run(20)
    ref_cells, _ = _cells()
    _check(ens, _plain(ref_cells, seeds, 30), "after autotune")


def _check_oracle(ens, orcs, tag, cells=None):
    for i in (range(len(orcs)) if cells is None else cells):
        o = orcs[i]
        g = ens.lattice(i)
        ol = o.lattice_u8()
        if not np.array_equal(g, ol):
            diff = np.argwhere(np.any(g != ol, axis=-1))
            raise AssertionError(f"{tag}: cell {i} {len(diff)} voxels differ from the oracle, first {tuple(diff[0])}")
        gc, oc = ens.concentrations(i), o.counts_array()
        if not np.array_equal(gc, oc):
            bad = np.argwhere(np.any(gc != oc, axis=1))[:, 0]
            raise AssertionError(f"{tag}: cell {i} counts of {len(bad)} voxels differ, first {bad[0]}: "
                                 f"{gc[bad[0]]} vs {oc[bad[0]]}")
        assert np.array_equal(ens.occupancy(i), o.occupancy), (tag, i)
        if o.cme is not None:
            assert np.array_equal(ens.cme_counts(i), o.cme_counts), (tag, i, ens.cme_counts(i), o.cme_counts)


def test_ensemble_matches_oracle(oracle_mod):
    announce(f"ensemble of {N_CELLS} gene-expression cells (edges 16..72) vs the oracle, 30 eager + 40 graph steps")
    cells, seeds = _cells()
    orcs = [oracle_mod.Oracle.from_simulation(c) for c in cells]
    ens = Ensemble(cells, seeds)
    ens.run(30)
    for o, sd in zip(orcs, seeds):
        o.run(default_itime_schedule(sd, 30))
    _check_oracle(ens, orcs, "eager 30")
    ens.run_graph(40, block=20)
    for o, sd in zip(orcs, seeds):
        o.run(default_itime_schedule(sd, 40, first=30))
    _check_oracle(ens, orcs, "graph 70")
    assert ens.counters()[CS_CNT_EXCEPTION] == 0
    assert sum(o.cme_counters.events for o in orcs) > 0


def _dense_cell(edge, density, seed):
    lam = 32e-9
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, 50e-6, lam, 4))
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
    sim.add_reaction(["A", "B"], ["C"], 50.0)
    sim.add_reaction(["C"], ["A", "B"], 5.0)
    sim.add_reaction(["A", "A"], ["B"], 40.0)  # k*n^2 fires on single particles: counts below zero
    return sim


def _reservoir_cell(edge, seed):
    lam = 32e-9
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, 50e-6, lam, 4))
    sim.set_seed(seed)
    sim.add_region(RegionType.Cube("cell", [0, 0, 0], [edge * lam] * 3), 2e-12)
    sim.prepare_regions()
    sim.add_particle_count("Ares", "cell", edge ** 3 // 20, True, True)
    sim.add_particle_count("A", "cell", edge ** 3 // 4, True, False)
    sim.add_particle_count("B", "cell", edge ** 3 // 8, True, False)
    sim.add_reaction(["Ares"], ["Ares", "B"], 4000.0)
    sim.add_reaction(["B"], [], 1000.0)
    sim.add_reaction(["A", "B"], ["A", "A"], 3000.0)
    return sim


def test_dense_and_reservoir_cells_replayed_per_cell(oracle_mod):
    """Dense S=4 cells (tight voxels, rejected claims, slot-S aliases, negative counts) and reservoir cells in one
    ensemble: every cell's replay follows the oracle step by step, eager and through graph blocks."""
    announce("ensemble of 4 dense + 2 reservoir S=4 cells vs the oracle, 12 eager + 12 graph steps")
    def build():
        cells = [_dense_cell(12, 1.5, 3), _dense_cell(17, 1.3, 4), _reservoir_cell(14, 5), _dense_cell(9, 1.6, 6),
                 _reservoir_cell(11, 7), _dense_cell(20, 1.4, 8)]
        return cells, [101 + 7 * i for i in range(len(cells))]
    cells, seeds = build()
    orcs = [oracle_mod.Oracle.from_simulation(c) for c in cells]
    ens = Ensemble(cells, seeds)
    for t in range(12):
        ens.run(1)
        for o, sd in zip(orcs, seeds):
            o.run(default_itime_schedule(sd, 1, first=t))
        _check_oracle(ens, orcs, f"eager step {t}")
    ens.run_graph(12, block=4)
    for o, sd in zip(orcs, seeds):
        o.run(default_itime_schedule(sd, 12, first=12))
    _check_oracle(ens, orcs, "graph 24")
    c = ens.counters()
    assert c[CS_CNT_EXCEPTION] == 0 and c[CS_CNT_TIGHT] > 0
    assert c[CS_CNT_FAILED_CLAIM] == sum(o.counters.failed_claims for o in orcs) > 0
    assert c[CS_CNT_ALIAS] == sum(o.counters.alias_writes for o in orcs) > 0
    assert any(np.any(o.counts_array()[:, 1:] < 0) for o in orcs), "the case must take counts below zero"


@pytest.mark.parametrize("populated,tau", [(False, configs.ENSEMBLE_TAU), (True, configs.ENSEMBLE_TAU),
                                           (True, 20e-6)])
def test_gene_expression_512_cells(oracle_mod, populated, tau):
    """The 512-cell workload (edges 32..96, S=8, rates x LogNormal(0.2), seed 99) for two 100-step graph blocks:
    the three smallest cells and the two most crowded against the oracle's whole state, every cell's slots
    compacted, particles outside the cytoplasm and nucleoid only where a slot-S alias put them (a claim into a full
    voxel lands in slot 0 of the next voxel, rdme.wgsl:30-40), no exception. `populated` starts each cell near its
    stationary state (see configs.gene_expression). At 20 us the protein's 6p passes 1, the cut cumulative vector
    drives the proteins against the -z side of the cell, and the packed voxels are replayed every step."""
    announce(f"512 gene-expression cells (populated={populated}, dt={tau}), 200 steps in graph blocks of 100")
    cells, seeds = configs.ensemble_cells(512, seed=99, populated=populated, tau=tau)
    sizes = [int(np.prod(c.info()["res"])) for c in cells]
    crowd = [int(np.count_nonzero(c.lattice())) / sizes[i] for i, c in enumerate(cells)]
    small = sorted(range(len(cells)), key=lambda i: sizes[i])[:3]
    small += [i for i in sorted(range(len(cells)), key=lambda i: -crowd[i]) if i not in small][:2]
    orcs = {i: oracle_mod.Oracle.from_simulation(cells[i]) for i in small}
    regions = [c.regions() for c in cells]
    start = [int(np.count_nonzero(c.lattice())) for c in cells]
    ens = Ensemble(cells, seeds)
    ens.run_graph(200, block=100)
    for i, o in orcs.items():
        o.run(default_itime_schedule(seeds[i], 200))
    _check_oracle(ens, orcs, "512 cells after 200 steps", cells=small)
    moved = outside = 0
    for i in range(len(cells)):
        lat = ens.lattice(i)
        flat = lat.reshape(-1, lat.shape[-1])
        nz = flat != 0
        assert not np.any(~nz[:, :-1] & nz[:, 1:]), f"cell {i} not compacted"
        occ = np.count_nonzero(flat, axis=1)
        reg = regions[i].ravel()
        outside += int(occ[(reg == 0) | (reg == 1)].sum())
        assert flat.max() <= 2
        moved += abs(int(nz.sum()) - start[i])
    c = ens.counters()
    assert c[CS_CNT_EXCEPTION] == 0
    assert outside <= c[CS_CNT_ALIAS], (outside, c[CS_CNT_ALIAS])
    if populated:
        assert sum(start) > 500_000
    print(f"particles at start {sum(start)}, net change summed over cells {moved}, outside the cell {outside}, "
          f"counters {c.tolist()}")
