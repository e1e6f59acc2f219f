"""ctypes binding of the CPU oracle (libcsoracle.so). TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this module; it checks the HIP
path and is never part of it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libcsoracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


class Model(C.Structure):
    _fields_ = [
        ("res", C.c_uint32 * 3), ("S", C.c_uint32), ("ns", C.c_uint32), ("nr", C.c_uint32),
        ("nreg", C.c_uint32), ("tau", C.c_float), ("lambda_", C.c_float), ("mod_zero", C.c_uint32),
        ("diffusion", C.c_void_p), ("stoich", C.c_void_p), ("reactants", C.c_void_p), ("rates", C.c_void_p),
        ("regions", C.c_void_p), ("reservoir", C.c_void_p),
    ]


class State(C.Structure):
    _fields_ = [("lattice", C.c_void_p), ("occupancy", C.c_void_p), ("counts", C.c_void_p), ("stat", C.c_void_p)]


class Counters(C.Structure):
    _fields_ = [("moves", C.c_uint64), ("failed_claims", C.c_uint64), ("alias_writes", C.c_uint64),
                ("fired", C.c_uint64), ("rollbacks", C.c_uint64), ("near_threshold", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        d = C.CDLL(str(LIB))
        d.cso_pcg.restype = C.c_uint32
        d.cso_pcg.argtypes = [C.c_uint32]
        d.cso_xorshift.restype = C.c_uint32
        d.cso_xorshift.argtypes = [C.c_uint32]
        d.cso_uint_to_float.restype = C.c_float
        d.cso_uint_to_float.argtypes = [C.c_uint32]
        d.cso_uniform_float.restype = C.c_float
        d.cso_uniform_float.argtypes = [C.c_uint32]
        d.cso_expf.restype = C.c_float
        d.cso_expf.argtypes = [C.c_float]
        d.cso_step.restype = C.c_int
        d.cso_step.argtypes = [C.POINTER(Model), C.POINTER(State), C.c_uint32, C.POINTER(Counters)]
        d.cso_run.restype = C.c_int
        d.cso_run.argtypes = [C.POINTER(Model), C.POINTER(State), C.c_void_p, C.c_uint32, C.POINTER(Counters)]
        d.cso_probability_vector.restype = None
        d.cso_probability_vector.argtypes = [C.POINTER(Model), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                             C.POINTER(C.c_float)]
        d.cso_phase_r.restype = None
        d.cso_phase_r.argtypes = [C.POINTER(Model), C.c_void_p, C.c_void_p, C.POINTER(State), C.c_uint32,
                                  C.POINTER(Counters)]
        d.cso_phase_c.restype = None
        d.cso_phase_c.argtypes = [C.POINTER(Model), C.POINTER(State), C.c_uint32, C.POINTER(Counters)]
        d.cso_phase_f.restype = None
        d.cso_phase_f.argtypes = [C.POINTER(Model), C.POINTER(State)]
        _lib = d
    return _lib


class Oracle:
    """Serial reference step over explicit arrays.

    lattice: uint32[N*S] (or anything convertible), regions: uint32[N]; tables in the reference layout
    (diffusion f32[nreg,nreg,ns+1], stoich i32[nr+1,ns+1], reactants i32[nr+1,3], rates f32[nr+1]).
    Per-voxel counts start as the slot counts plus one unit of the reservoir species at reservoir voxels
    (src/lattice.rs:292-300); `stat` starts as their sum (src/simulation.rs:361-375)."""

    def __init__(self, res, S, tau, lam, diffusion, stoich, reactants, rates, regions, lattice,
                 reservoir=None, mod_zero=0xFFFFFFFF):
        res = tuple(int(r) for r in res)
        N = res[0] * res[1] * res[2]
        self.res, self.S, self.N = res, int(S), N
        self.diffusion = np.ascontiguousarray(diffusion, dtype=np.float32).ravel()
        self.stoich = np.ascontiguousarray(stoich, dtype=np.int32)
        self.nr = self.stoich.shape[0] - 1
        self.ns = self.stoich.shape[1] - 1
        self.stoich = self.stoich.ravel()
        self.reactants = np.ascontiguousarray(reactants, dtype=np.int32).ravel()
        self.rates = np.ascontiguousarray(rates, dtype=np.float32).ravel()
        self.regions = np.ascontiguousarray(regions, dtype=np.uint32).ravel()
        self.nreg = int(round((self.diffusion.size / (self.ns + 1)) ** 0.5))
        self.reservoir = (np.zeros(N, dtype=np.uint32) if reservoir is None
                          else np.ascontiguousarray(reservoir, dtype=np.uint32).ravel())
        lat = np.ascontiguousarray(lattice, dtype=np.uint32).ravel()
        assert lat.size == N * self.S and self.regions.size == N
        self.lattice = lat.copy()
        L2 = self.lattice.reshape(N, self.S)
        self.occupancy = np.count_nonzero(L2, axis=1).astype(np.uint32)
        nz = np.nonzero(self.lattice)[0]
        flat = (nz // self.S).astype(np.int64) * (self.ns + 1) + self.lattice[nz].astype(np.int64)
        counts = np.bincount(flat, minlength=N * (self.ns + 1)).astype(np.int32).reshape(N, self.ns + 1)
        rv = np.nonzero(self.reservoir)[0]
        counts[rv, self.reservoir[rv]] += 1
        self.counts = counts.ravel()
        self.stat = counts.sum(axis=0).astype(np.int32)
        self.stat[0] = 0
        m = Model()
        m.res = (C.c_uint32 * 3)(*res)
        m.S, m.ns, m.nr, m.nreg = self.S, self.ns, self.nr, self.nreg
        m.tau, m.lambda_, m.mod_zero = float(tau), float(lam), int(mod_zero)
        m.diffusion = self.diffusion.ctypes.data
        m.stoich = self.stoich.ctypes.data
        m.reactants = self.reactants.ctypes.data
        m.rates = self.rates.ctypes.data
        m.regions = self.regions.ctypes.data
        m.reservoir = self.reservoir.ctypes.data
        self.model = m
        st = State()
        st.lattice = self.lattice.ctypes.data
        st.occupancy = self.occupancy.ctypes.data
        st.counts = self.counts.ctypes.data
        st.stat = self.stat.ctypes.data
        self.state = st
        self.counters = Counters()

    @classmethod
    def from_simulation(cls, sim, mod_zero=0xFFFFFFFF) -> "Oracle":
        inf = sim.info()
        t = sim.tables()
        return cls(inf["res"], inf["max_particles_site"], inf["tau"], inf["lambda"], t["diffusion"], t["stoich"],
                   t["reactants"], t["rates"], sim.regions(), sim.lattice(), sim.reservoir(), mod_zero)

    def step(self, itime: int) -> None:
        rc = lib().cso_step(C.byref(self.model), C.byref(self.state), int(itime) & 0xFFFFFFFF,
                            C.byref(self.counters))
        if rc:
            raise RuntimeError(f"cso_step failed ({rc})")

    def run(self, itimes) -> None:
        arr = np.ascontiguousarray(np.asarray(itimes, dtype=np.uint32))
        rc = lib().cso_run(C.byref(self.model), C.byref(self.state), arr.ctypes.data, arr.size,
                           C.byref(self.counters))
        if rc:
            raise RuntimeError(f"cso_run failed ({rc})")

    def probability_vector(self, x, y, z, sp) -> np.ndarray:
        p = (C.c_float * 7)()
        lib().cso_probability_vector(C.byref(self.model), x, y, z, sp, p)
        return np.frombuffer(p, dtype=np.float32).copy()

    def lattice_u8(self) -> np.ndarray:
        return self.lattice.astype(np.uint8).reshape(*self.res, self.S)

    def counts_array(self) -> np.ndarray:
        return self.counts.reshape(self.N, self.ns + 1)
