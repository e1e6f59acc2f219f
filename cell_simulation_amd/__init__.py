"""MI355X (gfx950) implementation of cell_simulation's reaction-diffusion master equation step.

The hot path is libcellsim.so (HIP kernels + C ABI, include/cell_sim.h); this package is its Python
mirror of the reference's host API.
"""
from ._lib import CellSimError, lib  # noqa: F401
from .ensemble import Ensemble  # noqa: F401
from .stats import StatisticContainer, collect_stats, make_all_stats  # noqa: F401
from .simulation import (  # noqa: F401
    LatticeParams,
    RegionType,
    Simulation,
    SolverStatisticSample,
    default_itime_schedule,
)
