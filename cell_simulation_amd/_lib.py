"""ctypes binding of libcellsim.so (declarations in include/cell_sim.h).

The library is loaded from this package directory only; there is no fallback implementation, so a
missing or stale build fails here, loudly.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libcellsim.so"

CS_OK = 0
CS_ERR_ARG = 1
CS_ERR_NOT_FOUND = 2
CS_ERR_STATE = 3
CS_ERR_HIP = 4
CS_ERR_UNSUPPORTED = 5
CS_ERR_FULL = 6
CS_ERR_PARSE = 7

CS_CNT_TIGHT = 0
CS_CNT_EXCEPTION = 1
CS_CNT_NEAR_THRESHOLD = 2
CS_CNT_FIRED = 3
CS_CNT_ROLLBACK = 4
CS_CNT_CME_EVENTS = 5
CS_CNT_CME_NEAR = 6
CS_CNT_EMIT_FAIL = 7
CS_CNT_FAILED_CLAIM = 8
CS_CNT_ALIAS = 9
CS_NUM_COUNTERS = 10


class LatticeParamsC(C.Structure):
    _fields_ = [
        ("dims", C.c_float * 3),
        ("res", C.c_uint32 * 3),
        ("tau", C.c_float),
        ("lambda_", C.c_float),
        ("max_particles_site", C.c_uint32),
    ]


class SimInfoC(C.Structure):
    _fields_ = [
        ("res", C.c_uint32 * 3),
        ("n_voxels", C.c_uint64),
        ("max_particles_site", C.c_uint32),
        ("num_species", C.c_uint32),
        ("num_reactions", C.c_uint32),
        ("num_regions", C.c_uint32),
        ("tau", C.c_float),
        ("lambda_", C.c_float),
        ("dims", C.c_float * 3),
        ("device_ready", C.c_int),
        ("num_cme_species", C.c_uint32),
        ("num_cme_reactions", C.c_uint32),
    ]


class EnsTilingC(C.Structure):
    _fields_ = [("xc", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("batch", C.c_uint32)]


class EnsInfoC(C.Structure):
    _fields_ = [("n_cells", C.c_uint32), ("max_particles_site", C.c_uint32), ("total_voxels", C.c_uint64),
                ("steps_done", C.c_uint64), ("items", C.c_uint32)]


class StatSampleC(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("value", C.c_int32), ("iteration_count", C.c_uint32)]


class CellSimError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_F3 = C.POINTER(C.c_float)
_SIM = C.c_void_p

# name: (restype, argtypes)
SIGNATURES = {
    "cs_last_error": (C.c_char_p, []),
    "cs_sim_create": (C.c_int, [C.POINTER(LatticeParamsC), C.POINTER(_SIM)]),
    "cs_sim_from_json_file": (C.c_int, [C.c_char_p, C.c_uint64, C.c_uint32, C.POINTER(_SIM)]),
    "cs_sim_from_json_string": (C.c_int, [C.c_char_p, C.c_uint64, C.c_uint32, C.POINTER(_SIM)]),
    "cs_sim_destroy": (None, [_SIM]),
    "cs_sim_set_seed": (C.c_int, [_SIM, C.c_uint64]),
    "cs_sim_add_region_cube": (C.c_int, [_SIM, C.c_char_p, _F3, _F3, C.c_float]),
    "cs_sim_add_region_sphere": (C.c_int, [_SIM, C.c_char_p, _F3, C.c_float, C.c_float]),
    "cs_sim_add_region_cylinder": (C.c_int, [_SIM, C.c_char_p, _F3, _F3, C.c_float, C.c_float]),
    "cs_sim_add_region_semisphere": (C.c_int, [_SIM, C.c_char_p, _F3, C.c_float, _F3, C.c_float]),
    "cs_sim_add_region_spherical_shell": (C.c_int, [_SIM, C.c_char_p, C.c_char_p, _F3, C.c_float, C.c_float,
                                                    C.c_float]),
    "cs_sim_add_region_cylindrical_shell": (C.c_int, [_SIM, C.c_char_p, C.c_char_p, _F3, _F3, C.c_float,
                                                      C.c_float, C.c_float]),
    "cs_sim_add_region_capsid": (C.c_int, [_SIM, C.c_char_p, C.c_char_p, _F3, _F3, C.c_float, C.c_float,
                                           C.c_float, C.c_float]),
    "cs_sim_prepare_regions": (C.c_int, [_SIM]),
    "cs_sim_region_index": (C.c_int, [_SIM, C.c_char_p]),
    "cs_sim_add_particle_count": (C.c_int, [_SIM, C.c_char_p, C.c_char_p, C.c_uint32, C.c_int, C.c_int]),
    "cs_sim_add_particle_concentration": (C.c_int, [_SIM, C.c_char_p, C.c_char_p, C.c_float, C.c_int, C.c_int]),
    "cs_sim_fill_region": (C.c_int, [_SIM, C.c_char_p, C.c_char_p, C.c_int]),
    "cs_sim_particle_random_walk": (C.c_int, [_SIM, C.c_char_p, C.c_char_p, C.c_float, C.c_float, C.c_float,
                                              C.c_int]),
    "cs_sim_add_sparse_region": (C.c_int, [_SIM, C.c_char_p, C.c_float, C.c_char_p, C.c_uint32, C.c_float]),
    "cs_sim_particle_index": (C.c_int, [_SIM, C.c_char_p]),
    "cs_sim_particle_name": (C.c_int, [_SIM, C.c_uint32, C.c_char_p, C.c_size_t]),
    "cs_sim_set_diffusion_rate": (C.c_int, [_SIM, C.c_char_p, C.c_float]),
    "cs_sim_set_diffusion_rate_particle": (C.c_int, [_SIM, C.c_char_p, C.c_char_p, C.c_float]),
    "cs_sim_set_transition_rate": (C.c_int, [_SIM, C.c_char_p, C.c_char_p, C.c_float]),
    "cs_sim_set_transition_rate_particle": (C.c_int, [_SIM, C.c_char_p, C.c_char_p, C.c_char_p, C.c_float]),
    "cs_sim_add_reaction": (C.c_int, [_SIM, C.POINTER(C.c_char_p), C.c_uint32, C.POINTER(C.c_char_p),
                                      C.c_uint32, C.c_float]),
    "cs_sim_prepare_device": (C.c_int, [_SIM, C.c_int]),
    "cs_sim_step": (C.c_int, [_SIM, C.c_uint32, C.c_uint32, C.c_uint32]),
    "cs_sim_run": (C.c_int, [_SIM, C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32]),
    "cs_sim_synchronize": (C.c_int, [_SIM]),
    "cs_sim_run_timed": (C.c_int, [_SIM, C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(C.c_float)]),
    "cs_sim_get_info": (C.c_int, [_SIM, C.POINTER(SimInfoC)]),
    "cs_sim_read_lattice": (C.c_int, [_SIM, C.c_void_p, C.c_size_t]),
    "cs_sim_write_lattice": (C.c_int, [_SIM, C.c_void_p, C.c_size_t]),
    "cs_sim_read_regions": (C.c_int, [_SIM, C.c_void_p, C.c_size_t]),
    "cs_sim_read_reservoir": (C.c_int, [_SIM, C.c_void_p, C.c_size_t]),
    "cs_sim_read_species_counts": (C.c_int, [_SIM, C.c_void_p, C.c_size_t]),
    "cs_sim_read_concentrations": (C.c_int, [_SIM, C.c_void_p, C.c_size_t]),
    "cs_sim_read_occupancy": (C.c_int, [_SIM, C.c_void_p, C.c_size_t]),
    "cs_sim_render_view": (C.c_int, [_SIM, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_size_t]),
    "cs_sim_read_region_counts": (C.c_int, [_SIM, C.c_void_p, C.c_size_t]),
    "cs_sim_read_region_counts_noalloc": (C.c_int, [_SIM, C.c_void_p, C.c_size_t]),
    "cs_sim_read_tables": (C.c_int, [_SIM, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p,
                                     C.c_size_t, C.c_void_p, C.c_size_t]),
    "cs_sim_pop_stat": (C.c_int, [_SIM, C.POINTER(StatSampleC)]),
    "cs_sim_read_counters": (C.c_int, [_SIM, C.c_void_p, C.c_size_t]),
    "cs_sim_reset_counters": (C.c_int, [_SIM]),
    "cs_sim_add_cme_species": (C.c_int, [_SIM, C.c_char_p, C.c_int32]),
    "cs_sim_add_cme_reaction": (C.c_int, [_SIM, C.POINTER(C.c_char_p), C.c_uint32, C.POINTER(C.c_char_p),
                                          C.c_uint32, C.c_float]),
    "cs_sim_set_cme_seed": (C.c_int, [_SIM, C.c_uint32]),
    "cs_sim_read_cme_counts": (C.c_int, [_SIM, C.c_void_p, C.c_size_t]),
    "cs_sim_cme_species_name": (C.c_int, [_SIM, C.c_uint32, C.c_char_p, C.c_size_t]),
    "cs_sim_read_cme_tables": (C.c_int, [_SIM, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p,
                                         C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                         C.POINTER(C.c_uint32)]),
    "cs_sim_set_cme_audit": (C.c_int, [_SIM, C.c_int]),
    "cs_ens_create": (C.c_int, [C.POINTER(_SIM), C.c_uint32, C.POINTER(C.c_uint32), C.c_int, C.POINTER(_SIM)]),
    "cs_ens_destroy": (None, [_SIM]),
    "cs_ens_run": (C.c_int, [_SIM, C.c_uint32]),
    "cs_ens_run_graph": (C.c_int, [_SIM, C.c_uint32, C.c_uint32]),
    "cs_ens_run_timed": (C.c_int, [_SIM, C.c_uint32, C.c_uint32, C.POINTER(C.c_float)]),
    "cs_ens_synchronize": (C.c_int, [_SIM]),
    "cs_ens_set_tiling": (C.c_int, [_SIM, C.POINTER(EnsTilingC)]),
    "cs_ens_get_tiling": (C.c_int, [_SIM, C.POINTER(EnsTilingC)]),
    "cs_ens_autotune": (C.c_int, [_SIM, C.c_uint32, C.POINTER(C.c_float)]),
    "cs_ens_get_info": (C.c_int, [_SIM, C.POINTER(EnsInfoC)]),
    "cs_ens_cell_res": (C.c_int, [_SIM, C.c_uint32, C.POINTER(C.c_uint32)]),
    "cs_ens_read_lattice": (C.c_int, [_SIM, C.c_uint32, C.c_void_p, C.c_size_t]),
    "cs_ens_read_cme_counts": (C.c_int, [_SIM, C.c_uint32, C.c_void_p, C.c_size_t]),
    "cs_ens_read_concentrations": (C.c_int, [_SIM, C.c_uint32, C.c_void_p, C.c_size_t]),
    "cs_ens_read_occupancy": (C.c_int, [_SIM, C.c_uint32, C.c_void_p, C.c_size_t]),
    "cs_ens_read_counters": (C.c_int, [_SIM, C.c_void_p, C.c_size_t]),
    "cs_ens_reset_counters": (C.c_int, [_SIM]),
    "cs_sim_read_cme_audit": (C.c_int, [_SIM, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint64)]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libcellsim.so from the package directory (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m cell_simulation_amd.build` "
                "(there is no non-HIP fallback)")
        dll = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(dll, name)
            fn.restype = res
            fn.argtypes = args
        _lib = dll
    return _lib


def check(rc: int) -> None:
    if rc != CS_OK:
        raise CellSimError(rc, lib().cs_last_error().decode(errors="replace"))
