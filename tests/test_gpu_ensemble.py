"""Batched ensemble vs plain per-cell simulations, bit for bit.

Each ensemble cell at step t uses itime = default_itime_schedule(itime_seed)[t], so it must evolve exactly
as the same cell built as a plain Simulation and stepped alone with that schedule: lattices and well-mixed
counts are compared for every cell after eager steps, after graph-replayed blocks, under different tilings
(queue batch sizes, the queue-free round-robin, planes per item) and across an autotune, which must leave
the state as it found it.
"""
import numpy as np
import pytest

from cell_simulation_amd import Ensemble, configs, default_itime_schedule
from cell_simulation_amd._lib import CS_CNT_EXCEPTION
from tests.conftest import announce

pytestmark = pytest.mark.gpu

N_CELLS = 7


def _cells():
    return configs.ensemble_cells(N_CELLS, seed=11, edge_range=(16, 72), rate_scale=20000.0)


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
