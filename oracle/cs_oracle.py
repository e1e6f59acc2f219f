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


class CmeModel(C.Structure):
    _fields_ = [
        ("nc", C.c_uint32), ("ncr", C.c_uint32), ("reactants", C.c_void_p), ("delta", C.c_void_p),
        ("emit", C.c_void_p), ("emit_region", C.c_void_p), ("rates", C.c_void_p), ("region_voxels", C.c_void_p),
        ("region_len", C.c_void_p), ("seed", C.c_uint32),
    ]


class CmeCounters(C.Structure):
    _fields_ = [("events", C.c_uint64), ("near", C.c_uint64), ("emit_fail", C.c_uint64), ("exceptions", C.c_uint64)]

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
        d.cso_run_mt.restype = C.c_int
        d.cso_run_mt.argtypes = [C.POINTER(Model), C.POINTER(State), C.c_void_p, C.c_uint32, C.POINTER(Counters),
                                 C.c_uint32]
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
        d.cso_view.restype = None
        d.cso_view.argtypes = [C.POINTER(Model), C.POINTER(State), C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
        d.cso_cme_h.restype = C.c_double
        d.cso_cme_h.argtypes = [C.c_void_p, C.c_void_p]
        d.cso_cme_bits.restype = C.c_uint64
        d.cso_cme_bits.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        d.cso_cme_phase.restype = None
        d.cso_cme_phase.argtypes = [C.POINTER(Model), C.POINTER(CmeModel), C.POINTER(State), C.c_void_p,
                                    C.c_uint32, C.POINTER(CmeCounters)]
        for nm in ("cso_cme_total64", "cso_cme_total32"):
            getattr(d, nm).restype = C.c_double
            getattr(d, nm).argtypes = [C.POINTER(CmeModel), C.c_void_p]
        for nm in ("cso_cme_choice64", "cso_cme_choice32"):
            getattr(d, nm).restype = C.c_uint32
            getattr(d, nm).argtypes = [C.POINTER(CmeModel), C.c_void_p, C.c_double]
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
        self.cme = None
        self.cme_counters = CmeCounters()

    def set_cme(self, counts, reactants, delta, emit, emit_region, rates, seed) -> None:
        """Attach a well-mixed pool (the build's extension; see cso_cme_phase). Region voxel lists are the
        ascending voxel indices of each label."""
        self.cme_counts = np.ascontiguousarray(counts, dtype=np.int32).copy()
        self.cme_reactants = np.ascontiguousarray(reactants, dtype=np.int32).ravel()
        self.cme_delta = np.ascontiguousarray(delta, dtype=np.int32).ravel()
        self.cme_emit = np.ascontiguousarray(emit, dtype=np.uint8).ravel()
        self.cme_emit_region = np.ascontiguousarray(emit_region, dtype=np.int32).ravel()
        self.cme_rates = np.ascontiguousarray(rates, dtype=np.float32).ravel()
        ncr = self.cme_rates.size
        self._lists = [np.ascontiguousarray(np.nonzero(self.regions == r)[0], dtype=np.uint32)
                       for r in range(self.nreg)]
        self._list_ptrs = (C.c_void_p * self.nreg)(*[l.ctypes.data for l in self._lists])
        self._list_len = np.array([l.size for l in self._lists], dtype=np.uint32)
        cm = CmeModel()
        cm.nc, cm.ncr = self.cme_counts.size, ncr
        cm.reactants = self.cme_reactants.ctypes.data
        cm.delta = self.cme_delta.ctypes.data
        cm.emit = self.cme_emit.ctypes.data
        cm.emit_region = self.cme_emit_region.ctypes.data
        cm.rates = self.cme_rates.ctypes.data
        cm.region_voxels = C.cast(self._list_ptrs, C.c_void_p)
        cm.region_len = self._list_len.ctypes.data
        cm.seed = int(seed) & 0xFFFFFFFF
        self.cme = cm

    def cme_phase(self, itime: int) -> None:
        lib().cso_cme_phase(C.byref(self.model), C.byref(self.cme), C.byref(self.state), self.cme_counts.ctypes.data,
                            int(itime) & 0xFFFFFFFF, C.byref(self.cme_counters))

    @classmethod
    def from_simulation(cls, sim, mod_zero=0xFFFFFFFF) -> "Oracle":
        inf = sim.info()
        t = sim.tables()
        o = cls(inf["res"], inf["max_particles_site"], inf["tau"], inf["lambda"], t["diffusion"], t["stoich"],
                t["reactants"], t["rates"], sim.regions(), sim.lattice(), sim.reservoir(), mod_zero)
        if inf.get("num_cme_reactions", 0) > 0:
            c = sim.cme_tables()
            o.set_cme(sim.cme_counts(), c["reactants"], c["delta"], c["emit"], c["emit_region"], c["rates"],
                      c["seed"])
        return o

    def step(self, itime: int) -> None:
        if self.cme is not None and self.cme.ncr > 0:
            self.cme_phase(itime)
        rc = lib().cso_step(C.byref(self.model), C.byref(self.state), int(itime) & 0xFFFFFFFF,
                            C.byref(self.counters))
        if rc:
            raise RuntimeError(f"cso_step failed ({rc})")

    def run(self, itimes) -> None:
        if self.cme is not None and self.cme.ncr > 0:
            for t in np.asarray(itimes, dtype=np.uint32):
                self.step(int(t))
            return
        arr = np.ascontiguousarray(np.asarray(itimes, dtype=np.uint32))
        rc = lib().cso_run(C.byref(self.model), C.byref(self.state), arr.ctypes.data, arr.size,
                           C.byref(self.counters))
        if rc:
            raise RuntimeError(f"cso_run failed ({rc})")

    def run_mt(self, itimes, threads: int) -> None:
        """run() on `threads` host threads (cso_run_mt: the same result); lattice-only models."""
        if self.cme is not None and self.cme.ncr > 0:
            raise ValueError("run_mt steps lattice models only")
        arr = np.ascontiguousarray(np.asarray(itimes, dtype=np.uint32))
        rc = lib().cso_run_mt(C.byref(self.model), C.byref(self.state), arr.ctypes.data, arr.size,
                              C.byref(self.counters), int(threads))
        if rc:
            raise RuntimeError(f"cso_run_mt failed ({rc})")

    def probability_vector(self, x, y, z, sp) -> np.ndarray:
        p = (C.c_float * 7)()
        lib().cso_probability_vector(C.byref(self.model), x, y, z, sp, p)
        return np.frombuffer(p, dtype=np.float32).copy()

    def lattice_u8(self) -> np.ndarray:
        return self.lattice.astype(np.uint8).reshape(*self.res, self.S)

    def view(self, view: int, axis: int = 0, slice_: int = 0) -> np.ndarray:
        """final_stage.wgsl:35-67: float32[N] in voxel order."""
        out = np.empty(self.N, dtype=np.float32)
        lib().cso_view(C.byref(self.model), C.byref(self.state), int(view), int(axis), int(slice_),
                       out.ctypes.data)
        return out

    def counts_array(self) -> np.ndarray:
        return self.counts.reshape(self.N, self.ns + 1)
