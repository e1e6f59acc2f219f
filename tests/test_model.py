"""Host model: JSON format, reaction/diffusion tables, region rasterisation, placement (CPU only).

Expected tables are written out by hand from the reference's rules (src/simulation.rs:74-76 dummy row,
:904-947 add_reaction, :321 sorted reaction keys, :537-548 diffusion matrix growth); region volumes are
checked against closed forms and an independent float32 evaluation of the sphere predicate.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from cell_simulation_amd import CellSimError, LatticeParams, RegionType, Simulation, configs
from cell_simulation_amd import _lib

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_easy_json_tables():
    sim = Simulation.from_file(GOLDEN / "easy.json", seed=1)
    inf = sim.info()
    assert inf["res"] == (32, 32, 64)
    assert inf["num_species"] == 4 and inf["num_reactions"] == 2 and inf["num_regions"] == 3
    assert [sim.particle_name(i) for i in range(5)] == ["void", "A", "B", "C", "IeX"]
    t = sim.tables()
    assert t["stoich"].tolist() == [[0, 0, 0, 0, 0], [0, -1, -1, 1, 0], [0, 1, 1, -1, 0]]
    assert t["reactants"].tolist() == [[0, 0, 0], [1, 2, 0], [3, 0, 0]]
    assert t["rates"].tolist() == [0.0, np.float32(5.82), np.float32(0.351)]
    D = t["diffusion"]
    base = np.float32(1.358e-14)
    bg = np.float32(8.15e-14) / np.float32(6.0)
    assert D.shape == (3, 3, 5)
    assert np.all(D[0, 0] == bg)                     # background column copied into every species
    assert np.all(D[1, 1] == base) and np.all(D[2, 2] == base)
    off = D[~np.eye(3, dtype=bool)]
    assert np.all(off == 0)                          # transition rates start at 0
    assert sim.region_index("membrane") == 1 and sim.region_index("interior") == 2
    assert sim.region_index("cap1") == -1            # joined into the shell


def test_easy_json_particles_and_reservoir():
    sim = Simulation.from_file(GOLDEN / "easy.json", seed=1)
    lat = sim.lattice()
    reg = sim.regions()
    assert int(np.count_nonzero(lat == 1)) == 5000 and int(np.count_nonzero(lat == 2)) == 5000
    assert int(np.count_nonzero(lat == 3)) == 0
    interior = reg == 2
    assert np.all(interior[np.any(lat != 0, axis=-1)])  # placed only in "interior"
    res = sim.reservoir()
    membrane_volume = int(np.count_nonzero(reg == 1))
    n_res = int(np.float32(0.1) * np.float32(membrane_volume))
    assert int(np.count_nonzero(res == 4)) == n_res
    assert np.all(reg[res == 4] == 1)
    counts = sim.species_counts()
    assert counts.tolist() == [0, 5000, 5000, 0, n_res]


def test_easy_json_capsid_volume_closed_form():
    sim = Simulation.from_file(GOLDEN / "easy.json", seed=1)
    reg = sim.regions()
    vox = 0.025 * 0.025 * 0.03125  # voxel volume: 0.8/32 x 0.8/32 x 2.0/64
    r_in, r_out, length = 0.37, 0.4, 2.0 - 0.8
    interior = (math.pi * r_in ** 2 * length + 4 / 3 * math.pi * r_in ** 3) / vox
    outer = (math.pi * r_out ** 2 * length + 4 / 3 * math.pi * r_out ** 3) / vox
    n_int = int(np.count_nonzero(reg == 2))
    n_mem = int(np.count_nonzero(reg == 1))
    assert abs(n_int - interior) / interior < 0.05
    assert abs((n_int + n_mem) - outer) / outer < 0.05


def test_reaction_keys_sorted_and_numeric_tokens():
    model = json.loads((GOLDEN / "easy.json").read_text())
    model["particles"] = [p for p in model["particles"] if not p["is_reservoir"]]
    model["reactions"] = {"C -> A + B": 0.5, "2 A -> B": 1.5, "A + B -> C": 4.0}
    sim = Simulation.from_json(json.dumps(model), seed=3)
    t = sim.tables()
    # byte-wise key order: "2 A -> B" < "A + B -> C" < "C -> A + B"
    assert t["reactants"].tolist() == [[0, 0, 0], [1, 0, 0], [1, 2, 0], [3, 0, 0]]
    assert t["rates"].tolist() == [0.0, 3.0, 4.0, 0.5]
    assert t["stoich"].tolist() == [[0, 0, 0, 0], [0, -1, 1, 0], [0, -1, -1, 1], [0, 1, 1, -1]]


def test_sphere_volume_matches_predicate():
    lam = 32e-9
    R = 40
    sim = Simulation(LatticeParams([R * lam] * 3, [R] * 3, 1e-5, lam, 8))
    centre = [20.5 * lam, 19.0 * lam, 21.25 * lam]
    radius = 13.3 * lam
    sim.add_region(RegionType.Sphere("ball", centre, radius), 1e-12)
    reg = sim.regions()
    f = np.float32
    vs = f(R * lam) / f(R)
    c = [f(v) / vs for v in centre]
    x = np.arange(R, dtype=np.float32)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    d = ((X - c[0]) * (X - c[0])) * (vs * vs)
    d = d + ((Y - c[1]) * (Y - c[1])) * (vs * vs)
    d = d + ((Z - c[2]) * (Z - c[2])) * (vs * vs)
    inside = d < f(radius) * f(radius)
    assert np.array_equal(reg == 1, inside)
    assert abs(inside.sum() - 4 / 3 * math.pi * 13.3 ** 3) / (4 / 3 * math.pi * 13.3 ** 3) < 0.02


def test_cube_and_shell_regions():
    # simulation.rs:627-658: voxels [trunc(p0 / vs), trunc(pf / vs)) on each axis, vs = dims / res. With a
    # power-of-two voxel edge every quotient below is exact in f32, so the box is fixed voxel for voxel:
    # whole multiples start and end on that voxel, fractions truncate toward the lower one.
    lam = 2.0 ** -25
    sim = Simulation(LatticeParams([20 * lam] * 3, [20] * 3, 1e-5, lam, 8))
    sim.add_region(RegionType.Cube("box", [2 * lam, 3.5 * lam, 4.75 * lam], [10 * lam, 9.25 * lam, 8.5 * lam]),
                   1e-12)
    reg = sim.regions()
    exp = np.zeros((20, 20, 20), bool)
    exp[2:10, 3:9, 4:8] = True
    assert np.array_equal(reg == 1, exp) and np.all(reg[~exp] == 0)
    # a degenerate box (p0 == pf after truncation) holds no voxel; a later box overwrites an earlier one
    sim.add_region(RegionType.Cube("flat", [5 * lam, 5 * lam, 5 * lam], [5.5 * lam, 12 * lam, 12 * lam]), 1e-12)
    sim.add_region(RegionType.Cube("corner", [0.0, 0.0, 0.0], [3 * lam, 4 * lam, 5 * lam]), 1e-12)
    reg = sim.regions()
    assert int((reg == 2).sum()) == 0
    corner = np.zeros((20, 20, 20), bool)
    corner[0:3, 0:4, 0:5] = True
    assert np.array_equal(reg == 3, corner) and np.array_equal(reg == 1, exp & ~corner)
    sim2 = Simulation(LatticeParams([30 * lam] * 3, [30] * 3, 1e-5, lam, 8))
    sim2.add_region(RegionType.SphericalShell("mem", "cyto", [15 * lam] * 3, 10 * lam, 12 * lam), 1e-12)
    r2 = sim2.regions()
    assert sim2.info()["num_regions"] == 3
    n_shell, n_core = int((r2 == 1).sum()), int((r2 == 2).sum())
    assert abs(n_core - 4 / 3 * math.pi * 1000) / (4 / 3 * math.pi * 1000) < 0.03
    assert abs((n_shell + n_core) - 4 / 3 * math.pi * 1728) / (4 / 3 * math.pi * 1728) < 0.03


def test_placement_is_seeded():
    a = configs.spherical_cell(seed=42, edge=32, counts=(300, 300)).lattice()
    b = configs.spherical_cell(seed=42, edge=32, counts=(300, 300)).lattice()
    c = configs.spherical_cell(seed=43, edge=32, counts=(300, 300)).lattice()
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    nz = a.reshape(-1, a.shape[-1]) != 0
    assert not np.any(~nz[:, :-1] & nz[:, 1:])  # appended at slot `occupancy` (lattice.rs:284)


def test_builder_errors():
    lam = 32e-9
    sim = Simulation(LatticeParams([4 * lam] * 3, [4] * 3, 1e-5, lam, 4))
    with pytest.raises(CellSimError) as e:
        sim.add_particle_count("A", "background", 1, False, False)  # before prepare_regions
    assert e.value.code == _lib.CS_ERR_STATE
    sim.prepare_regions()
    with pytest.raises(CellSimError) as e:
        sim.add_particle_count("A", "nowhere", 1, False, False)
    assert e.value.code == _lib.CS_ERR_NOT_FOUND
    with pytest.raises(CellSimError) as e:
        sim.add_particle_count("A", "background", 4 * 64 + 1, False, False)  # more than S per voxel
    assert e.value.code == _lib.CS_ERR_FULL
    assert int(np.count_nonzero(sim.lattice())) == 0                    # placement rolled back
    sim.add_particle_count("A", "background", 10, False, False)
    with pytest.raises(CellSimError) as e:
        sim.add_reaction(["A", "A", "A", "A"], [], 1.0)
    assert e.value.code == _lib.CS_ERR_ARG
    with pytest.raises(CellSimError) as e:
        sim.add_reaction(["Z"], [], 1.0)
    assert e.value.code == _lib.CS_ERR_NOT_FOUND
    with pytest.raises(CellSimError) as e:
        sim.step(0, 1, 0)
    assert e.value.code == _lib.CS_ERR_STATE


def test_json_errors():
    model = json.loads((GOLDEN / "easy.json").read_text())
    with pytest.raises(CellSimError) as e:
        Simulation.from_json("{ not json")
    assert e.value.code == _lib.CS_ERR_PARSE
    m2 = dict(model)
    m2["regions"] = model["regions"] + [{"type": "sparse", "name": "s", "base_diffusion_rate": 1e-14}]
    with pytest.raises(CellSimError) as e:
        Simulation.from_json(json.dumps(m2))
    assert e.value.code == _lib.CS_ERR_UNSUPPORTED
    m3 = dict(model)
    m3["particles"] = [{"name": "A", "to_region": "interior", "logging": True, "is_reservoir": False}]
    with pytest.raises(CellSimError) as e:
        Simulation.from_json(json.dumps(m3))
    assert e.value.code == _lib.CS_ERR_PARSE


# ---------------------------------------------- sparse regions, random walks ----
# The reference draws both from rand::thread_rng (unseedable), so their placement is checked by invariants:
# the geometry each routine guarantees, the bookkeeping of volumes, labels and tables, and the sizes.
def _cell(edge=40, S=8, seed=5):
    lam = 32e-9
    sim = Simulation(LatticeParams([edge * lam] * 3, [edge] * 3, 50e-6, lam, S))
    sim.set_seed(seed)
    sim.add_region(RegionType.Sphere("cell", [edge * lam / 2] * 3, 0.45 * edge * lam), 1e-12)
    sim.prepare_regions()
    return sim, lam


def _ball_size(radius, dims, res):
    # simulation.rs:507-520 in f32: offsets d of [-rv, rv) per axis (rv = radius / voxel size, truncated)
    # with sum(d_i^2 * voxel_size_i^2) < radius^2
    vs = np.float32(dims) / np.float32(res)
    r = np.float32(radius)
    rv = int(r / vs)
    r2, vs2 = r * r, vs * vs
    n = 0
    for i in range(-rv, rv):
        for j in range(-rv, rv):
            for k in range(-rv, rv):
                d = np.float32(i * i) * vs2
                d = d + np.float32(j * j) * vs2
                d = d + np.float32(k * k) * vs2
                n += bool(d < r2)
    return n


def test_sparse_region_balls_inside_target():
    sim, lam = _cell()
    before = sim.regions().copy()
    cell = sim.region_index("cell")
    vol_cell = int(np.count_nonzero(before == cell))
    sim.add_sparse_region("granules", RegionType.Sphere("g", [0, 0, 0], 3 * lam), "cell", 400, 2e-13)
    after = sim.regions()
    g = sim.region_index("granules")
    assert g == sim.info()["num_regions"] - 1
    new = after == g
    ball = _ball_size(3 * lam, 40 * lam, 40)
    n = int(np.count_nonzero(new))
    assert n % ball == 0 and 400 <= n < 400 + ball          # whole balls until max_volume is reached
    assert np.all(before[new] == cell)                       # every ball lies inside the target region
    assert int(np.count_nonzero(after == cell)) == vol_cell - n
    assert np.array_equal(after[~new], before[~new])
    D = sim.tables()["diffusion"]
    assert np.all(D[g, g] == np.float32(2e-13)) and np.all(D[g, :g] == 0) and np.all(D[:g, g] == 0)
    # particles can be placed into it once regions are prepared again (it re-lists them itself)
    sim.add_particle_count("E", "granules", 50, True, False)
    lat = sim.lattice()
    assert np.all(after[np.any(lat == sim.particle_index("E"), axis=-1)] == g)


def test_sparse_region_rejects_oversized_volume():
    sim, lam = _cell(edge=16)
    with pytest.raises(CellSimError):
        sim.add_sparse_region("g", 2 * lam, "cell", 10 ** 6, 0.0)


def test_random_walk_places_a_chain_inside_the_region():
    # the start keeps 25 * 0.4 blocks of room on every side (lattice.rs:121-123): a box of +-15 voxels here,
    # whose corners must stay inside the sphere of radius 28.8
    sim, lam = _cell(edge=64, S=16)
    regions = sim.regions()
    cell = sim.region_index("cell")
    block, radius, total = 1.5 * lam, lam, 40 * lam
    sim.particle_random_walk("W", "cell", total, block, radius, True)
    w = sim.particle_index("W")
    assert w == sim.info()["num_species"]
    lat = sim.lattice()
    has_w = np.any(lat == w, axis=-1)
    assert np.all(regions[has_w] == cell)                    # no block leaves the region
    count = int(np.count_nonzero(lat == w))
    # about pi r^2 L voxels along the kept blocks (voxel rasterisation of a thin tube: generous bounds)
    tube = math.pi * 1.0 ** 2 * 40
    assert 0.5 * tube < count < 2.0 * tube, (count, tube)
    # the slots stay compacted and the counts agree with the lattice
    occ = np.count_nonzero(lat, axis=-1)
    assert np.all(np.sum(lat[..., :] != 0, axis=-1) == occ)
    assert sim.species_counts()[w] == count


def test_random_walk_fails_cleanly_in_a_region_too_small():
    sim, lam = _cell(edge=12)
    with pytest.raises(CellSimError):
        sim.particle_random_walk("W", "cell", 100 * lam, 4 * lam, lam, False)
