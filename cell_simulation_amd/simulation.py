"""Python mirror of the reference crate's host API (`Simulation`, `LatticeParams`, `RegionType`).

Method names, argument meaning and order follow src/simulation.rs, src/lattice_params.rs and
src/region.rs; every call goes through libcellsim.so's C ABI. Errors the reference answers with a
panic raise CellSimError here.
"""
from __future__ import annotations

import ctypes as C
from collections import deque
from dataclasses import dataclass
from typing import Deque, Sequence

import numpy as np

from . import _lib
from ._lib import CellSimError, check, lib


def _f3(v: Sequence[float]):
    if len(v) != 3:
        raise ValueError("expected three components")
    return (C.c_float * 3)(*[float(x) for x in v])


@dataclass
class LatticeParams:
    """LatticeParams::new(dimensions, resolution, tau, lambda) (src/lattice_params.rs:61-81).

    `max_particles_site` is the reference's compile-time MAX_PARTICLES_SITE (src/lib.rs:15)."""
    dimensions: Sequence[float]
    resolution: Sequence[int]
    tau: float
    lambda_: float
    max_particles_site: int = 8

    def voxel_size(self):
        """lattice_params.rs:28-34 in f32."""
        d = np.asarray(self.dimensions, dtype=np.float32)
        r = np.asarray(self.resolution, dtype=np.float32)
        return d / r


class RegionType:
    """The variants of src/region.rs:9-20 accepted by Simulation.add_region."""

    @dataclass
    class Cube:
        name: str
        p0: Sequence[float]
        pf: Sequence[float]

    @dataclass
    class Sphere:
        name: str
        center: Sequence[float]
        radius: float

    @dataclass
    class SemiSphere:
        name: str
        center: Sequence[float]
        radius: float
        direction: Sequence[float]

    @dataclass
    class Cylinder:
        name: str
        p0: Sequence[float]
        pf: Sequence[float]
        radius: float

    @dataclass
    class SphericalShell:
        shell_name: str
        interior_name: str
        center: Sequence[float]
        internal_radius: float
        external_radius: float

    @dataclass
    class CylindricalShell:
        shell_name: str
        interior_name: str
        p0: Sequence[float]
        pf: Sequence[float]
        internal_radius: float
        external_radius: float

    @dataclass
    class Capsid:
        shell_name: str
        interior_name: str
        center: Sequence[float]
        dir: Sequence[float]
        internal_radius: float
        external_radius: float
        total_length: float


@dataclass
class SolverStatisticSample:
    """src/statistics.rs:4-9"""
    name: str
    value: int
    iteration_count: int


class Simulation:
    """src/simulation.rs:29-46. Builder calls mirror the reference; `prepare_for_gpu(device)` uploads to a
    HIP device and `step(frame_num, itime, write_freq)` enqueues one step there."""

    def __init__(self, lattice_params: LatticeParams | None = None, *, _handle=None):
        self._h = C.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            p = _lib.LatticeParamsC()
            p.dims = (C.c_float * 3)(*lattice_params.dimensions)
            p.res = (C.c_uint32 * 3)(*lattice_params.resolution)
            p.tau = lattice_params.tau
            p.lambda_ = lattice_params.lambda_
            p.max_particles_site = lattice_params.max_particles_site
            check(lib().cs_sim_create(C.byref(p), C.byref(self._h)))
        self.stats: Deque[SolverStatisticSample] = deque()

    # ---- construction from the JSON model format (src/simulation.rs:210-232) ----
    @classmethod
    def from_file(cls, path: str, seed: int = 0, max_particles_site: int = 0) -> "Simulation":
        h = C.c_void_p()
        check(lib().cs_sim_from_json_file(str(path).encode(), seed, max_particles_site, C.byref(h)))
        return cls(_handle=h)

    @classmethod
    def from_json(cls, text: str, seed: int = 0, max_particles_site: int = 0) -> "Simulation":
        h = C.c_void_p()
        check(lib().cs_sim_from_json_string(text.encode(), seed, max_particles_site, C.byref(h)))
        return cls(_handle=h)

    def close(self) -> None:
        if self._h:
            lib().cs_sim_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- model building ----
    def set_seed(self, seed: int) -> None:
        check(lib().cs_sim_set_seed(self._h, seed))

    def add_region(self, reg, diffusion_rate: float) -> None:
        """Simulation::add_region (src/simulation.rs:428-478)."""
        L, h, d = lib(), self._h, float(diffusion_rate)
        T = RegionType
        if isinstance(reg, T.Cube):
            rc = L.cs_sim_add_region_cube(h, reg.name.encode(), _f3(reg.p0), _f3(reg.pf), d)
        elif isinstance(reg, T.Sphere):
            rc = L.cs_sim_add_region_sphere(h, reg.name.encode(), _f3(reg.center), reg.radius, d)
        elif isinstance(reg, T.Cylinder):
            rc = L.cs_sim_add_region_cylinder(h, reg.name.encode(), _f3(reg.p0), _f3(reg.pf), reg.radius, d)
        elif isinstance(reg, T.SemiSphere):
            rc = L.cs_sim_add_region_semisphere(h, reg.name.encode(), _f3(reg.center), reg.radius,
                                                _f3(reg.direction), d)
        elif isinstance(reg, T.SphericalShell):
            rc = L.cs_sim_add_region_spherical_shell(h, reg.shell_name.encode(), reg.interior_name.encode(),
                                                     _f3(reg.center), reg.internal_radius, reg.external_radius, d)
        elif isinstance(reg, T.CylindricalShell):
            rc = L.cs_sim_add_region_cylindrical_shell(h, reg.shell_name.encode(), reg.interior_name.encode(),
                                                       _f3(reg.p0), _f3(reg.pf), reg.internal_radius,
                                                       reg.external_radius, d)
        elif isinstance(reg, T.Capsid):
            rc = L.cs_sim_add_region_capsid(h, reg.shell_name.encode(), reg.interior_name.encode(),
                                            _f3(reg.center), _f3(reg.dir), reg.internal_radius,
                                            reg.external_radius, reg.total_length, d)
        else:
            raise CellSimError(_lib.CS_ERR_ARG, f"region type not implemented: {reg!r}")
        check(rc)

    def add_sparse_region(self, name: str, base_region, to_region: str, max_volume: int,
                          diffusion_rate: float) -> None:
        """simulation.rs:483-535. `base_region` is a RegionType.Sphere (only its radius is used) or a radius."""
        radius = base_region if isinstance(base_region, (int, float)) else getattr(base_region, "radius", None)
        if radius is None:
            raise CellSimError(_lib.CS_ERR_ARG, "only spheres can be added as sparse regions")
        check(lib().cs_sim_add_sparse_region(self._h, name.encode(), float(radius), to_region.encode(),
                                             int(max_volume), float(diffusion_rate)))

    def prepare_regions(self) -> None:
        check(lib().cs_sim_prepare_regions(self._h))

    def region_index(self, name: str) -> int:
        return lib().cs_sim_region_index(self._h, name.encode())

    def add_particle_count(self, name: str, to_region: str, count: int, logging: bool, is_reservoir: bool) -> None:
        check(lib().cs_sim_add_particle_count(self._h, name.encode(), to_region.encode(), int(count),
                                              int(bool(logging)), int(bool(is_reservoir))))

    def add_particle_concentration(self, name: str, to_region: str, concentration: float, logging: bool,
                                   is_reservoir: bool) -> None:
        check(lib().cs_sim_add_particle_concentration(self._h, name.encode(), to_region.encode(),
                                                      float(concentration), int(bool(logging)),
                                                      int(bool(is_reservoir))))

    def fill_region(self, name: str, to_region: str, logging: bool) -> None:
        check(lib().cs_sim_fill_region(self._h, name.encode(), to_region.encode(), int(bool(logging))))

    def particle_random_walk(self, name: str, to_region: str, total_length: float, block_length: float,
                             radius: float, logging: bool) -> None:
        """simulation.rs:842-864: a new species laid as a chain of cylinders inside `to_region`."""
        check(lib().cs_sim_particle_random_walk(self._h, name.encode(), to_region.encode(), float(total_length),
                                                float(block_length), float(radius), int(bool(logging))))

    def particle_index(self, name: str) -> int:
        return lib().cs_sim_particle_index(self._h, name.encode())

    def particle_name(self, index: int) -> str:
        buf = C.create_string_buffer(256)
        check(lib().cs_sim_particle_name(self._h, index, buf, 256))
        return buf.value.decode()

    def set_diffusion_rate(self, region: str, rate: float) -> None:
        check(lib().cs_sim_set_diffusion_rate(self._h, region.encode(), float(rate)))

    def set_diffusion_rate_particle(self, particle: str, region: str, rate: float) -> None:
        check(lib().cs_sim_set_diffusion_rate_particle(self._h, particle.encode(), region.encode(), float(rate)))

    def set_transition_rate(self, from_region: str, to_region: str, rate: float) -> None:
        check(lib().cs_sim_set_transition_rate(self._h, from_region.encode(), to_region.encode(), float(rate)))

    def set_transition_rate_particle(self, particle: str, from_region: str, to_region: str, rate: float) -> None:
        check(lib().cs_sim_set_transition_rate_particle(self._h, particle.encode(), from_region.encode(),
                                                        to_region.encode(), float(rate)))

    def add_reaction(self, reactants: Sequence[str], products: Sequence[str], k: float) -> None:
        r = (C.c_char_p * max(1, len(reactants)))(*[s.encode() for s in reactants])
        p = (C.c_char_p * max(1, len(products)))(*[s.encode() for s in products])
        check(lib().cs_sim_add_reaction(self._h, r, len(reactants), p, len(products), float(k)))

    # ---- well-mixed (CME) species: no reference counterpart ----
    def add_cme_species(self, name: str, count: int) -> None:
        """A well-mixed species with an initial count (one pool per simulation)."""
        check(lib().cs_sim_add_cme_species(self._h, name.encode(), int(count)))

    def add_cme_reaction(self, reactants: Sequence[str], products: Sequence[str], k: float) -> None:
        """A well-mixed reaction; products are well-mixed names or lattice particles as "name@region"."""
        r = (C.c_char_p * max(1, len(reactants)))(*[s.encode() for s in reactants])
        p = (C.c_char_p * max(1, len(products)))(*[s.encode() for s in products])
        check(lib().cs_sim_add_cme_reaction(self._h, r, len(reactants), p, len(products), float(k)))

    def set_cme_seed(self, seed: int) -> None:
        check(lib().cs_sim_set_cme_seed(self._h, seed & 0xFFFFFFFF))

    def cme_counts(self) -> np.ndarray:
        n = self.info()["num_cme_species"]
        out = np.zeros(max(n, 1), dtype=np.int32)
        check(lib().cs_sim_read_cme_counts(self._h, out.ctypes.data, n))
        return out[:n]

    def cme_species_name(self, index: int) -> str:
        buf = C.create_string_buffer(256)
        check(lib().cs_sim_cme_species_name(self._h, index, buf, len(buf)))
        return buf.value.decode()

    def cme_tables(self) -> dict:
        inf = self.info()
        nc, ncr = inf["num_cme_species"], inf["num_cme_reactions"]
        reac = np.zeros(max(3 * ncr, 1), dtype=np.int32)
        delta = np.zeros(max(ncr * nc, 1), dtype=np.int32)
        emit = np.zeros(max(4 * ncr, 1), dtype=np.uint8)
        region = np.zeros(max(ncr, 1), dtype=np.int32)
        rates = np.zeros(max(ncr, 1), dtype=np.float32)
        seed = C.c_uint32(0)
        check(lib().cs_sim_read_cme_tables(self._h, reac.ctypes.data, 3 * ncr, delta.ctypes.data, ncr * nc,
                                           emit.ctypes.data, 4 * ncr, region.ctypes.data, ncr, rates.ctypes.data,
                                           ncr, C.byref(seed)))
        return {"reactants": reac[:3 * ncr].reshape(ncr, 3), "delta": delta[:ncr * nc].reshape(ncr, nc),
                "emit": emit[:4 * ncr].reshape(ncr, 4), "emit_region": region[:ncr], "rates": rates[:ncr],
                "seed": int(seed.value), "init": self.cme_counts() if not inf["device_ready"] else None}

    def set_cme_audit(self, on: bool) -> None:
        check(lib().cs_sim_set_cme_audit(self._h, 1 if on else 0))

    def cme_audit(self) -> dict:
        rel, div, ev = C.c_double(0.0), C.c_uint64(0), C.c_uint64(0)
        check(lib().cs_sim_read_cme_audit(self._h, C.byref(rel), C.byref(div), C.byref(ev)))
        return {"max_rel_error": float(rel.value), "divergences": int(div.value), "events": int(ev.value)}

    # ---- device ----
    def prepare_for_gpu(self, device: int = 0) -> None:
        """Simulation::prepare_for_gpu (src/simulation.rs:152-208) on HIP device `device`."""
        check(lib().cs_sim_prepare_device(self._h, int(device)))

    def step(self, frame_num: int, itime: int, write_freq: int = 0) -> None:
        """Simulation::step (src/simulation.rs:99-150); `itime` is the per-step RNG seed."""
        check(lib().cs_sim_step(self._h, frame_num & 0xFFFFFFFF, itime & 0xFFFFFFFF, write_freq))
        self._drain_stats()

    def run(self, first_frame: int, itimes: Sequence[int], write_freq: int = 0) -> None:
        arr = np.ascontiguousarray(np.asarray(itimes, dtype=np.uint32))
        check(lib().cs_sim_run(self._h, first_frame, arr.ctypes.data_as(C.POINTER(C.c_uint32)), len(arr),
                               write_freq))
        self._drain_stats()

    def run_timed(self, itimes: Sequence[int]) -> float:
        """Steps without statistics sampling; returns the summed step-kernel time in ms (HIP events on the
        simulation's stream: stamped by the wave kernel's own dispatch, or recorded around each step for the
        other kernels)."""
        arr = np.ascontiguousarray(np.asarray(itimes, dtype=np.uint32))
        ms = C.c_float(0.0)
        check(lib().cs_sim_run_timed(self._h, arr.ctypes.data_as(C.POINTER(C.c_uint32)), len(arr), C.byref(ms)))
        return float(ms.value)

    def synchronize(self) -> None:
        check(lib().cs_sim_synchronize(self._h))

    def _drain_stats(self) -> None:
        s = _lib.StatSampleC()
        while lib().cs_sim_pop_stat(self._h, C.byref(s)) == _lib.CS_OK:
            self.stats.append(SolverStatisticSample(s.name.decode(), int(s.value), int(s.iteration_count)))

    # ---- inspection ----
    def info(self) -> dict:
        i = _lib.SimInfoC()
        check(lib().cs_sim_get_info(self._h, C.byref(i)))
        return {
            "res": tuple(i.res), "n_voxels": int(i.n_voxels), "max_particles_site": int(i.max_particles_site),
            "num_species": int(i.num_species), "num_reactions": int(i.num_reactions),
            "num_regions": int(i.num_regions), "tau": float(i.tau), "lambda": float(i.lambda_),
            "dims": tuple(i.dims), "device_ready": bool(i.device_ready),
            "num_cme_species": int(i.num_cme_species), "num_cme_reactions": int(i.num_cme_reactions),
        }

    def lattice(self) -> np.ndarray:
        """Slots as uint8[R0, R1, R2, S] (species ids, compacted)."""
        inf = self.info()
        S = inf["max_particles_site"]
        out = np.empty(inf["n_voxels"] * S, dtype=np.uint8)
        check(lib().cs_sim_read_lattice(self._h, out.ctypes.data, out.nbytes))
        return out.reshape(*inf["res"], S)

    def write_lattice(self, lat: np.ndarray) -> None:
        arr = np.ascontiguousarray(lat, dtype=np.uint8)
        check(lib().cs_sim_write_lattice(self._h, arr.ctypes.data, arr.nbytes))

    def regions(self) -> np.ndarray:
        inf = self.info()
        out = np.empty(inf["n_voxels"], dtype=np.uint8)
        check(lib().cs_sim_read_regions(self._h, out.ctypes.data, out.nbytes))
        return out.reshape(*inf["res"])

    def reservoir(self) -> np.ndarray:
        inf = self.info()
        out = np.empty(inf["n_voxels"], dtype=np.uint32)
        check(lib().cs_sim_read_reservoir(self._h, out.ctypes.data, out.size))
        return out.reshape(*inf["res"])

    def concentrations(self) -> np.ndarray:
        """Per-voxel species counts int32[N][ns+1], the reference's `concentrations` buffer
        (src/lattice.rs:14-23): the slot contents plus the reservoir units no slot holds."""
        inf = self.info()
        n1 = inf["num_species"] + 1
        out = np.empty(inf["n_voxels"] * n1, dtype=np.int32)
        check(lib().cs_sim_read_concentrations(self._h, out.ctypes.data, out.size))
        return out.reshape(inf["n_voxels"], n1)

    def occupancy(self) -> np.ndarray:
        """Per-voxel occupancy word uint32[N] (src/lattice.rs:14-23; rdme.wgsl:61-80, cme.wgsl:114)."""
        inf = self.info()
        out = np.empty(inf["n_voxels"], dtype=np.uint32)
        check(lib().cs_sim_read_occupancy(self._h, out.ctypes.data, out.size))
        return out

    def render_view(self, view: int = 0, slice_axis: int = 0, slice_: int = 0) -> np.ndarray:
        """final_stage.wgsl:35-67 on the device: float32[R0, R1, R2] (0 = majority species, 1 = region,
        2 = reservoir, each / 255; voxels with res[axis] - coord[axis] < slice are 0)."""
        inf = self.info()
        out = np.empty(inf["n_voxels"], dtype=np.float32)
        check(lib().cs_sim_render_view(self._h, int(view), int(slice_axis), int(slice_), out.ctypes.data, out.size))
        return out.reshape(*inf["res"])

    def species_counts(self) -> np.ndarray:
        n = self.info()["num_species"] + 1
        out = np.empty(n, dtype=np.int32)
        check(lib().cs_sim_read_species_counts(self._h, out.ctypes.data, n))
        return out

    def region_counts(self) -> np.ndarray:
        inf = self.info()
        n1, nreg = inf["num_species"] + 1, inf["num_regions"]
        out = np.empty(nreg * n1, dtype=np.int32)
        check(lib().cs_sim_read_region_counts(self._h, out.ctypes.data, out.size))
        return out.reshape(nreg, n1)

    def region_counts_noalloc(self) -> np.ndarray:
        """region_counts() through the entry point that allocates nothing (its bins are a fixed per-GPU array)."""
        inf = self.info()
        n1, nreg = inf["num_species"] + 1, inf["num_regions"]
        out = np.empty(nreg * n1, dtype=np.int32)
        check(lib().cs_sim_read_region_counts_noalloc(self._h, out.ctypes.data, out.size))
        return out.reshape(nreg, n1)

    def tables(self) -> dict:
        inf = self.info()
        n1, nr1, nreg = inf["num_species"] + 1, inf["num_reactions"] + 1, inf["num_regions"]
        d = np.empty(nreg * nreg * n1, dtype=np.float32)
        st = np.empty(nr1 * n1, dtype=np.int32)
        ix = np.empty(nr1 * 3, dtype=np.int32)
        k = np.empty(nr1, dtype=np.float32)
        check(lib().cs_sim_read_tables(self._h, d.ctypes.data, d.size, st.ctypes.data, st.size, ix.ctypes.data,
                                       ix.size, k.ctypes.data, k.size))
        return {"diffusion": d.reshape(nreg, nreg, n1), "stoich": st.reshape(nr1, n1),
                "reactants": ix.reshape(nr1, 3), "rates": k}

    def counters(self) -> np.ndarray:
        out = np.zeros(_lib.CS_NUM_COUNTERS, dtype=np.uint64)
        check(lib().cs_sim_read_counters(self._h, out.ctypes.data, out.size))
        return out

    def reset_counters(self) -> None:
        check(lib().cs_sim_reset_counters(self._h))


def default_itime_schedule(seed: int, nsteps: int, first: int = 0) -> np.ndarray:
    """Per-step seeds: the low 32 bits of SplitMix64 applied to (seed << 32) + t.

    The reference reads the wall clock in microseconds instead (examples/no_window.rs:127-129). Seeds
    of consecutive steps must not differ by small amounts: the hop draw of slot s in voxel v hashes
    itime + (x % y) + z + v*S + s (rdme.wgsl:220), so itime_{t+1} = itime_t + 1 would hand slot s the
    draw slot s+1 had one step earlier, and particles sharing a voxel would keep hopping together."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed & 0xFFFFFFFF) << np.uint64(32)) + np.arange(first, first + nsteps, dtype=np.uint64)
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z & np.uint64(0xFFFFFFFF)).astype(np.uint32)
