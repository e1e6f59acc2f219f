"""Ensemble statistics of lattice species against an f64 closed form, on the GPU: 400 replicate cells of the
isomerisation model of tests/test_oracle_stats.py (32^3, 40 particles diffusing, A <-> B) in one batched
ensemble; the replicates' sample mean and sample variance of the A count must lie within 3 sigma of
Binomial(40, p), every cell must keep its 40 particles, and the replicates must differ from each other
(independent seeds and placements)."""
import numpy as np
import pytest

from cell_simulation_amd import Ensemble
from tests.conftest import announce
from tests.test_oracle_stats import (EDGE, N_STEPS, cell_seed, check_a_counts, isomerisation_cell, itime_seed,
                                     species_counts)

pytestmark = pytest.mark.gpu

N_CELLS = 400


def test_ensemble_isomerisation_mean_and_variance():
    announce(f"ensemble of {N_CELLS} cells {EDGE}^3, A <-> B, {N_STEPS} steps: mean and variance vs f64 closed form")
    cells = [isomerisation_cell(cell_seed(i)) for i in range(N_CELLS)]
    ens = Ensemble(cells, [itime_seed(i) for i in range(N_CELLS)])
    ens.run_graph(N_STEPS, block=50)
    n_a = np.array([species_counts(ens.lattice(i))[0] for i in range(N_CELLS)], np.int64)
    print(check_a_counts(n_a))
    assert len({ens.lattice(i).tobytes() for i in range(8)}) == 8
