"""Batched ensembles of independent cells on one device (no reference counterpart).

An `Ensemble` copies the host models of many `Simulation`s (any lattice sizes, own tables, regions and
well-mixed pools; one slot count) into one device state and steps them together: a prologue launch per
step derives each cell's itime from a device-side step counter and runs the well-mixed phases, a
persistent launch takes (cell, tile, x-chunk) items from a device-side queue, and an epilogue replays
tight voxels and re-arms the queue. `run_graph` replays a captured hipGraph of `block` steps. Cell c at
step t uses itime = low 32 bits of SplitMix64((itime_seeds[c] << 32) + t), i.e.
default_itime_schedule(itime_seeds[c]), so each cell evolves exactly as its Simulation would alone.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import check, lib


class Ensemble:
    def __init__(self, cells: Sequence, itime_seeds: Sequence[int], device: int = 0):
        n = len(cells)
        if len(itime_seeds) != n:
            raise ValueError("one itime seed per cell")
        handles = (C.c_void_p * max(n, 1))(*[c._h.value for c in cells])
        seeds = (C.c_uint32 * max(n, 1))(*[int(s) & 0xFFFFFFFF for s in itime_seeds])
        h = C.c_void_p()
        check(lib().cs_ens_create(handles, n, seeds, int(device), C.byref(h)))
        self._h = h
        self._nc = [c.info()["num_cme_species"] for c in cells]
        self._ns = [c.info()["num_species"] for c in cells]

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().cs_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps: int) -> None:
        check(lib().cs_ens_run(self._h, int(nsteps)))

    def run_graph(self, nsteps: int, block: int = 100) -> None:
        check(lib().cs_ens_run_graph(self._h, int(nsteps), int(block)))

    def run_timed(self, nsteps: int, block: int = 100) -> float:
        ms = C.c_float(0.0)
        check(lib().cs_ens_run_timed(self._h, int(nsteps), int(block), C.byref(ms)))
        return float(ms.value)

    def synchronize(self) -> None:
        check(lib().cs_ens_synchronize(self._h))

    @property
    def tiling(self) -> dict:
        t = _lib.EnsTilingC()
        check(lib().cs_ens_get_tiling(self._h, C.byref(t)))
        return {"xc": int(t.xc), "blocks_per_cu": int(t.blocks_per_cu), "batch": int(t.batch)}

    def set_tiling(self, xc: int = 16, blocks_per_cu: int = 3, batch: int = 1) -> None:
        t = _lib.EnsTilingC(int(xc), int(blocks_per_cu), int(batch))
        check(lib().cs_ens_set_tiling(self._h, C.byref(t)))

    def autotune(self, trial_steps: int = 10) -> float:
        ms = C.c_float(0.0)
        check(lib().cs_ens_autotune(self._h, int(trial_steps), C.byref(ms)))
        return float(ms.value)

    def info(self) -> dict:
        i = _lib.EnsInfoC()
        check(lib().cs_ens_get_info(self._h, C.byref(i)))
        return {"n_cells": int(i.n_cells), "max_particles_site": int(i.max_particles_site),
                "total_voxels": int(i.total_voxels), "steps_done": int(i.steps_done), "items": int(i.items)}

    def cell_res(self, cell: int) -> tuple:
        r = (C.c_uint32 * 3)()
        check(lib().cs_ens_cell_res(self._h, int(cell), r))
        return tuple(int(x) for x in r)

    def lattice(self, cell: int) -> np.ndarray:
        res = self.cell_res(cell)
        S = self.info()["max_particles_site"]
        out = np.empty(res[0] * res[1] * res[2] * S, dtype=np.uint8)
        check(lib().cs_ens_read_lattice(self._h, int(cell), out.ctypes.data, out.nbytes))
        return out.reshape(*res, S)

    def cme_counts(self, cell: int) -> np.ndarray:
        n = self._nc[cell]
        out = np.zeros(max(n, 1), dtype=np.int32)
        check(lib().cs_ens_read_cme_counts(self._h, int(cell), out.ctypes.data, n))
        return out[:n]

    def concentrations(self, cell: int) -> np.ndarray:
        """The cell's per-voxel counts C[v][sp], int32[N, ns+1] (see Simulation.concentrations)."""
        res = self.cell_res(cell)
        n = res[0] * res[1] * res[2]
        n1 = self._ns[cell] + 1
        out = np.empty(n * n1, dtype=np.int32)
        check(lib().cs_ens_read_concentrations(self._h, int(cell), out.ctypes.data, out.size))
        return out.reshape(n, n1)

    def occupancy(self, cell: int) -> np.ndarray:
        """The cell's occupancy words, uint32[N] (see Simulation.occupancy)."""
        res = self.cell_res(cell)
        out = np.empty(res[0] * res[1] * res[2], dtype=np.uint32)
        check(lib().cs_ens_read_occupancy(self._h, int(cell), out.ctypes.data, out.size))
        return out

    def counters(self) -> np.ndarray:
        out = np.zeros(_lib.CS_NUM_COUNTERS, dtype=np.uint64)
        check(lib().cs_ens_read_counters(self._h, out.ctypes.data, out.size))
        return out

    def reset_counters(self) -> None:
        check(lib().cs_ens_reset_counters(self._h))
