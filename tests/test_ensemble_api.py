"""Host-side checks of the batched-ensemble entry points (no GPU needed: they fail before device setup)."""
import pytest

from cell_simulation_amd import Ensemble, configs
from cell_simulation_amd._lib import CS_ERR_ARG, CellSimError


def test_ensemble_rejects_mixed_slot_counts():
    a = configs.gene_expression(edge=16, slots=8)
    b = configs.gene_expression(edge=16, slots=4)
    with pytest.raises(CellSimError) as e:
        Ensemble([a, b], [1, 2])
    assert e.value.code == CS_ERR_ARG


def test_ensemble_needs_cells_and_seeds():
    with pytest.raises(CellSimError) as e:
        Ensemble([], [])
    assert e.value.code == CS_ERR_ARG
    a = configs.gene_expression(edge=16)
    with pytest.raises(ValueError):
        Ensemble([a], [1, 2])


def test_ensemble_cells_are_reproducible():
    c1, s1 = configs.ensemble_cells(6, seed=3, edge_range=(12, 30))
    c2, s2 = configs.ensemble_cells(6, seed=3, edge_range=(12, 30))
    assert s1 == s2
    for x, y in zip(c1, c2):
        assert x.info()["res"] == y.info()["res"]
        assert (x.lattice() == y.lattice()).all()
        assert (x.cme_tables()["rates"] == y.cme_tables()["rates"]).all()
    rates = [c.cme_tables()["rates"][0] for c in c1]
    assert len(set(float(r) for r in rates)) == len(rates)  # perturbed per replicate
