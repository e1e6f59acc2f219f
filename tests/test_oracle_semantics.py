"""Pins of the CPU oracle that do not go through the oracle's own formulas.

* RNG: known answers derived from the shader text (tests/golden/rng_kat.json) and algebraic identities
  of xorshift32 and the u32 -> f32 map.
* Tiny lattices whose outcome is derived by hand from res/shader/rdme.wgsl, cme.wgsl and
  final_stage.wgsl: forced hops (direction order), the capacity hazards (slot-S alias, rejected claim),
  one reaction per voxel per step, the rollback at S slots, the zero-propensity firing at seed 0, the choice
  between two channels by the cumulative propensity (exact powers of two, r2 from the pinned RNG).
* Closed forms: hop probabilities, acceptance probability 1 - exp(-a*tau) (binomial 3-sigma).
* Invariants on dense random lattices: conservation, occupancy == non-empty slots, counts == slots.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def make(oracle_mod, res, S, lattice, ns, D=None, regions=None, stoich=None, reactants=None, rates=None,
         tau=1.0, lam=1.0, nreg=1):
    N = res[0] * res[1] * res[2]
    if D is None:
        D = np.zeros((nreg, nreg, ns + 1), np.float32)
    if regions is None:
        regions = np.zeros(N, np.uint32)
    if stoich is None:
        stoich = np.zeros((1, ns + 1), np.int32)
        reactants = np.zeros((1, 3), np.int32)
        rates = np.zeros(1, np.float32)
    return oracle_mod.Oracle(res, S, tau, lam, D, stoich, reactants, rates, regions, np.asarray(lattice, np.uint32))


# ---------------------------------------------------------------- RNG ----
def test_rng_known_answers(oracle_mod):
    lib = oracle_mod.lib()
    for c in json.loads((GOLDEN / "rng_kat.json").read_text())["cases"]:
        assert lib.cso_pcg(c["seed"]) == c["pcg"]
        assert lib.cso_xorshift(c["pcg"]) == c["xs"]
        assert lib.cso_uniform_float(c["pcg"]) == np.float32(c["uf"])
        assert lib.cso_uniform_float((c["pcg"] + 1) & 0xFFFFFFFF) == np.float32(c["uf_next"])


def test_rng_identities(oracle_mod):
    lib = oracle_mod.lib()
    rng = np.random.default_rng(0)
    xs = lambda v: lib.cso_xorshift(int(v))
    assert xs(1) == 0x42021
    for a, b in rng.integers(0, 2**32, size=(200, 2), dtype=np.uint64):
        assert xs(int(a) ^ int(b)) == xs(a) ^ xs(b)          # GF(2)-linear
    for p in rng.integers(0, 2**31, size=200, dtype=np.uint64) * 2:
        assert xs(int(p) + 1) == xs(p) ^ 0x42021              # even P: P+1 = P ^ 1
    # u32 -> f32 scaled by 2^-32 reaches 1.0 exactly from 4294967168 upward (round to nearest even)
    assert lib.cso_uint_to_float(4294967167) < 1.0
    assert lib.cso_uint_to_float(4294967168) == 1.0
    assert lib.cso_uint_to_float(0xFFFFFFFF) == 1.0
    assert lib.cso_uint_to_float(1) == 2.0 ** -32
    assert lib.cso_pcg(0) == 0
    assert lib.cso_expf(0.0) == 1.0
    assert lib.cso_expf(-1.0) == np.float32(math.exp(-1.0))


# ----------------------------------------------------- hop probabilities ----
def test_probability_vector_interior_and_boundary(oracle_mod):
    # spherical-cell numbers: D = 1e-12, tau = 50us, lambda = 32nm -> p = 0.048828125 per direction
    D = np.full((1, 1, 2), 1e-12, np.float32)
    o = make(oracle_mod, (4, 4, 4), 4, np.zeros(64 * 4), 1, D=D, tau=50e-6, lam=32e-9)
    p = o.probability_vector(2, 2, 2, 1)
    step = np.float32(1e-12) * np.float32(50e-6) / (np.float32(32e-9) * np.float32(32e-9))
    assert step == np.float32(0.048828125)
    assert np.array_equal(p[:6], np.cumsum(np.full(6, step, np.float32), dtype=np.float32))
    assert p[6] == 1.0
    q = o.probability_vector(0, 3, 0, 1)  # -x, +y, -z missing: entries repeat
    assert q[0] == 0.0 and q[1] == step and q[2] == 2 * step and q[3] == q[2] and q[4] == q[3]
    assert q[5] == 3 * step


# ------------------------------------------------------------ hops ----
def test_forced_minus_x_walk(oracle_mod):
    # D*tau/lambda^2 = 1: p[0] = 1 where -x exists, so the particle hops -x. At x == 0 the missing -x
    # entry has zero width (p[0] = 0, p[1] = 1, rdme.wgsl:143-156), so the draw selects +x instead of
    # reflecting: the particle bounces 2 -> 1 -> 0 -> 1.
    D = np.ones((1, 1, 2), np.float32)
    lat = np.zeros(3 * 4, np.uint32)
    lat[2 * 4] = 1
    o = make(oracle_mod, (3, 1, 1), 4, lat, 1, D=D)
    assert np.array_equal(o.probability_vector(0, 0, 0, 1), [0, 1, 1, 1, 1, 1, 1])
    for want_x in (1, 0, 1):
        o.step(12345)
        L = o.lattice.reshape(3, 4)
        assert L[want_x, 0] == 1 and np.count_nonzero(L) == 1


def test_direction_order(oracle_mod):
    # Only the +y neighbour of the centre of a 3^3 lattice is region 1 and only region 0 -> 1 has D = 1,
    # so p = [0,0,0,1,1,1]: the particle must hop +y (index 3 in the order -x,+x,-y,+y,-z,+z).
    regions = np.zeros((3, 3, 3), np.uint32)
    regions[1, 2, 1] = 1
    D = np.zeros((2, 2, 2), np.float32)
    D[0, 1, 1] = 1.0
    lat = np.zeros((3, 3, 3, 2), np.uint32)
    lat[1, 1, 1, 0] = 1
    o = make(oracle_mod, (3, 3, 3), 2, lat, 1, D=D, regions=regions, nreg=2)
    assert np.array_equal(o.probability_vector(1, 1, 1, 1)[:6], [0, 0, 0, 1, 1, 1])
    o.step(777)
    L = o.lattice.reshape(3, 3, 3, 2)
    assert L[1, 2, 1, 0] == 1 and np.count_nonzero(L) == 1


def _column(oracle_mod, lat):
    # 1 x 1 x 3 column, S = 2: species 1 never moves (D = 0), species 2 has p[5] = 1 (always +z).
    D = np.zeros((1, 1, 3), np.float32)
    D[0, 0, 2] = 1.0
    return make(oracle_mod, (1, 1, 3), 2, lat, 2, D=D)


def test_claim_into_full_voxel_aliases_next_voxel(oracle_mod):
    # z0 = [2,0] hops +z into z1 = [1,1] (full). old = O'[1] = 2 <= S admits the claim (H3) and the
    # first-fit over slots 0..S reaches slot S = z2's slot 0, which is empty (H2): the particle lands in z2
    # while O[1] becomes 3.
    o = _column(oracle_mod, [2, 0, 1, 1, 0, 0])
    o.step(1000)
    assert list(o.lattice) == [0, 0, 1, 1, 2, 0]
    assert list(o.occupancy) == [0, 3, 0]
    C = o.counts_array()
    assert list(C[1, 1:]) == [2, 1] and list(C[2, 1:]) == [0, 0]
    assert o.counters.alias_writes == 1
    # Next step z1 walks slots 0..O[1]-1 = 0..2, i.e. into z2's slot 0, and treats that particle as its
    # own: at z1 the -z neighbour exists, so p[4] = 1 and it hops -z into z0. The books balance again.
    o.step(2000)
    assert list(o.lattice) == [2, 0, 1, 1, 0, 0]
    assert list(o.occupancy) == [1, 2, 0]
    C = o.counts_array()
    assert list(C[0, 1:]) == [0, 1] and list(C[1, 1:]) == [2, 0] and list(C[2, 1:]) == [0, 0]


def test_claim_rejected_when_alias_slot_taken(oracle_mod):
    o = _column(oracle_mod, [2, 0, 1, 1, 1, 0])
    o.step(1000)
    assert list(o.lattice) == [2, 0, 1, 1, 1, 0]
    assert list(o.occupancy) == [1, 2, 1]
    assert o.counters.failed_claims == 1 and o.counters.moves == 0


# -------------------------------------------------------- reactions ----
def _one_voxel(oracle_mod, slots, stoich_rows, reactants, rates, S=4, ns=2):
    stoich = np.zeros((len(stoich_rows) + 1, ns + 1), np.int32)
    for i, r in enumerate(stoich_rows):
        stoich[i + 1] = r
    rx = np.zeros((len(reactants) + 1, 3), np.int32)
    for i, r in enumerate(reactants):
        rx[i + 1, :len(r)] = r
    k = np.concatenate([[0.0], rates]).astype(np.float32)
    return make(oracle_mod, (1, 1, 1), S, slots, ns, stoich=stoich, reactants=rx, rates=k)


def test_one_reaction_per_voxel_per_step(oracle_mod):
    # A -> B with a0*tau huge: 1 - exp(-a0*tau) = 1, so it fires every step, once, until A is gone.
    o = _one_voxel(oracle_mod, [1, 1, 1, 0], [[0, -1, 1]], [[1]], [1e30])
    seen = []
    for t in range(5):
        o.step(100 + t)
        seen.append(list(o.lattice))
    assert seen == [[1, 1, 2, 0], [1, 2, 2, 0], [2, 2, 2, 0], [2, 2, 2, 0], [2, 2, 2, 0]]
    assert list(o.stat[1:]) == [0, 3]


def _pcg(state):
    # random.wgsl:14-20 (pinned by tests/golden/rng_kat.json above)
    state &= 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    return (word >> 22) ^ word


def _uf(state):
    # random.wgsl:32-45 then 27-29: xorshift32, u32 -> f32 times 2^-32
    s = state & 0xFFFFFFFF
    s ^= (s << 13) & 0xFFFFFFFF
    s ^= s >> 17
    s ^= (s << 5) & 0xFFFFFFFF
    return float(np.float32(s) * np.float32(2.0 ** -32))


def test_two_channels_split_by_cumulative_propensity(oracle_mod):
    # One voxel holding two A; A -> B at k and A -> C at 3k, k = 2^90 so that a1 = 2^91, a2 = 3 * 2^91 and
    # a0 = 2^93 are exact and 1 - exp(-a0 * tau) = 1: every step fires. cme.wgsl:50-52, 68-73 pick the first i
    # with r2 <= cum_i / a0, i.e. A -> B for r2 <= 1/4 and A -> C above, with r2 = UF(PCG(seed) + 1) and seed =
    # itime + x*y + z = itime. The rewrite (cme.wgsl:85-114) then writes A, B or A, C.
    seen = {"B": [], "C": []}
    for itime in range(1, 400):
        r2 = _uf(_pcg(itime) + 1)
        o = _one_voxel(oracle_mod, [1, 1, 0, 0], [[0, -1, 1, 0], [0, -1, 0, 1]], [[1], [1]],
                       [2.0 ** 90, 3 * 2.0 ** 90], ns=3)
        o.step(itime)
        assert o.counters.fired == 1
        want = [1, 2, 0, 0] if r2 <= 0.25 else [1, 3, 0, 0]
        assert list(o.lattice) == want, (itime, r2)
        seen["B" if r2 <= 0.25 else "C"].append(r2)
    # both channels taken, from either side of the split and far from it
    assert len(seen["B"]) > 50 and len(seen["C"]) > 200
    assert min(seen["C"]) < 0.3 and max(seen["B"]) > 0.2 and max(seen["C"]) > 0.9


def test_rollback_when_rewrite_reaches_s(oracle_mod):
    # A -> 2B: from [A,A,0,0] the rewrite writes A,B,B (3 < S=4) and keeps it; the next firing would write
    # B,B,B,B and reach S, so it is undone (the effective capacity after a reaction is S-1).
    o = _one_voxel(oracle_mod, [1, 1, 0, 0], [[0, -1, 2]], [[1]], [1e30])
    o.step(100)
    assert list(o.lattice) == [1, 2, 2, 0] and list(o.occupancy) == [3]
    o.step(101)
    assert list(o.lattice) == [1, 2, 2, 0] and list(o.occupancy) == [3]
    assert o.counters.rollbacks == 1 and list(o.stat[1:]) == [1, 2]
    full = _one_voxel(oracle_mod, [1, 1, 1, 0], [[0, -1, 2]], [[1]], [1e30])
    full.step(100)
    assert list(full.lattice) == [1, 1, 1, 0] and full.counters.rollbacks == 1


# ------------------------------------------------------- reservoirs ----
# A reservoir site holds one unit of its species in the per-voxel counts and in no slot
# (lattice.rs:292-300). cme.wgsl:80 skips the species reservoirs[v*S] -- the lattice index of the voxel's
# slot 0, which names the voxel's own entry only for v = 0; an index past the end reads 0 here.
def _reservoir_column(oracle_mod, slots, reservoir, stoich_rows, reactants, rates, S=4, ns=3):
    n = len(reservoir)
    stoich = np.zeros((len(stoich_rows) + 1, ns + 1), np.int32)
    for i, r in enumerate(stoich_rows):
        stoich[i + 1] = r
    rx = np.zeros((len(reactants) + 1, 3), np.int32)
    for i, r in enumerate(reactants):
        rx[i + 1, :len(r)] = r
    k = np.concatenate([[0.0], rates]).astype(np.float32)
    return oracle_mod.Oracle((1, 1, n), S, 1.0, 1.0, np.zeros((1, 1, ns + 1), np.float32), stoich, rx, k,
                             np.zeros(n, np.uint32), np.asarray(slots, np.uint32), np.asarray(reservoir, np.uint32))


def test_reservoir_at_voxel_zero_is_a_constant_source(oracle_mod):
    # voxel 0 reads its own entry: A -> A + B never rewrites A, so the unit stays out of the slots while B
    # piles up until the rewrite would reach S (then it is undone)
    o = _reservoir_column(oracle_mod, [0, 0, 0, 0], [1], [[0, 0, 1, 0]], [[1]], [1e30])
    for t in range(3):
        o.step(100 + t)
    assert list(o.lattice) == [2, 2, 2, 0] and list(o.occupancy) == [3]
    assert list(o.counts_array()[0]) == [1, 1, 3, 0]
    o.step(103)
    assert list(o.lattice) == [2, 2, 2, 0] and o.counters.rollbacks == 1
    assert list(o.stat[1:]) == [1, 3, 0]


def test_reservoir_unit_joins_the_slots_where_the_index_misses(oracle_mod):
    # voxel 1 of 2 reads reservoirs[4], past the end: nothing is skipped, so the first firing writes C[A] = 1
    # copy of A next to the new B, while the occupancy word only adds the stoichiometry (+1)
    o = _reservoir_column(oracle_mod, [0] * 8, [0, 1], [[0, 0, 1, 0]], [[1]], [1e30])
    o.step(100)
    assert list(o.lattice[4:]) == [1, 2, 0, 0] and list(o.occupancy) == [0, 1]
    o.step(200)
    assert list(o.lattice[4:]) == [1, 2, 2, 0] and list(o.occupancy) == [0, 2]
    assert list(o.counts_array()[1]) == [1, 1, 2, 0] and list(o.stat[1:]) == [1, 2, 0]


def test_reservoir_index_skips_another_species_at_voxel_one(oracle_mod):
    # voxel 4 is a reservoir of A, so voxel 1 (index 1*S = 4) skips A: B -> C there drops its real A from
    # the slots but keeps it in the counts, and the occupancy word keeps counting it
    slots = [0] * 32
    slots[4:8] = [1, 2, 0, 0]
    o = _reservoir_column(oracle_mod, slots, [0, 0, 0, 0, 1, 0, 0, 0], [[0, 0, -1, 1]], [[2]], [1e30])
    o.step(100)
    assert list(o.lattice[4:8]) == [3, 0, 0, 0] and o.occupancy[1] == 2
    assert list(o.counts_array()[1]) == [1, 1, 0, 1] and list(o.counts_array()[4]) == [1, 1, 0, 0]
    assert list(o.stat[1:]) == [2, 0, 1]


def test_rollback_subtracts_the_skipped_species(oracle_mod):
    # A + B -> 3B at voxel 1 (skips A): B would reach S, so the rewrite is undone, and the undo subtracts
    # stoich[A] = -1 from C[A] although A was never added: C[A] grows by one per firing
    slots = [0] * 32
    slots[4:8] = [1, 2, 2, 0]
    o = _reservoir_column(oracle_mod, slots, [0, 0, 0, 0, 1, 0, 0, 0], [[0, -1, 2, 0]], [[1, 2]], [1e30])
    for t in range(3):
        o.step(100 + t)
    assert list(o.lattice[4:8]) == [1, 2, 2, 0] and o.occupancy[1] == 3
    assert o.counters.rollbacks == 3
    assert list(o.counts_array()[1]) == [1, 4, 2, 0] and list(o.stat[1:]) == [5, 2, 0]


def test_zero_propensity_fires_only_at_seed_zero(oracle_mod):
    # No reactions: a0 = 0, threshold 1 - exp(0) = 0, so it "fires" only when UF(PCG(seed)) == 0, i.e. seed 0
    # (itime + x*y + z = itime here). cum = 0/0 = NaN selects the dummy row 0 and the rewrite sorts the
    # slots by species.
    o = _one_voxel(oracle_mod, [2, 1, 0, 0], [], [], [])
    o.step(5)
    assert list(o.lattice) == [2, 1, 0, 0] and o.counters.fired == 0
    o.step(0)
    assert list(o.lattice) == [1, 2, 0, 0] and o.counters.fired == 1


def test_acceptance_probability_matches_closed_form(oracle_mod):
    # 4096 isolated voxels (no diffusion), one A each, A -> B at k: per step P(fire) = 1 - exp(-k*tau).
    n = 4096
    S = 4
    lat = np.zeros(n * S, np.uint32)
    lat[::S] = 1
    k, tau = 300.0, 1e-3
    stoich = np.array([[0, 0, 0], [0, -1, 1]], np.int32)
    rx = np.array([[0, 0, 0], [1, 0, 0]], np.int32)
    o = oracle_mod.Oracle((16, 16, 16), S, tau, 1.0, np.zeros((1, 1, 3), np.float32), stoich, rx,
                          np.array([0, k], np.float32), np.zeros(n, np.uint32), lat)
    o.step(0x9E3779B9)
    fired = int(o.stat[2])
    p = 1 - math.exp(-k * tau)
    # CME seeds itime + x*y + z repeat across voxels (x*y collides), so count distinct seeds only.
    x, y, z = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    distinct = len(np.unique((x * y + z).ravel()))
    sigma = math.sqrt(n * p * (1 - p) * n / distinct)
    assert abs(fired - n * p) < 3 * sigma, (fired, n * p, sigma)


def test_hop_fractions_match_probabilities(oracle_mod):
    # One particle at every (3i+1, 3j+1, 3k+1) of a 60^3 box (8000 particles, no two sharing a neighbour),
    # p = 0.04 per direction: after one step the number found one hop away in direction d is
    # binomial(n, 0.04) for each d, and the number that stayed binomial(n, 0.76).
    R, S = 60, 4
    c = np.arange(1, R, 3)
    xi, yi, zi = (a.ravel() for a in np.meshgrid(c, c, c, indexing="ij"))
    lat = np.zeros((R, R, R, S), np.uint32)
    lat[xi, yi, zi, 0] = 1
    D = np.full((1, 1, 2), 0.04, np.float32)
    o = make(oracle_mod, (R, R, R), S, lat.ravel(), 1, D=D)
    o.step(0xC0FFEE)
    occ = o.lattice.reshape(R, R, R, S)[..., 0] != 0
    n = len(xi)
    offsets = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
    for dx, dy, dz in offsets:
        got = int(np.count_nonzero(occ[xi + dx, yi + dy, zi + dz]))
        sigma = math.sqrt(n * 0.04 * 0.96)
        assert abs(got - n * 0.04) < 3 * sigma, ((dx, dy, dz), got, n * 0.04, sigma)
    stayed = int(np.count_nonzero(occ[xi, yi, zi]))
    assert abs(stayed - n * 0.76) < 3 * math.sqrt(n * 0.76 * 0.24), stayed
    assert int(np.count_nonzero(occ)) == n


# ------------------------------------------------------- invariants ----
def test_invariants_dense_random(oracle_mod):
    rng = np.random.default_rng(7)
    R, S, ns = 12, 8, 3
    occ = rng.integers(0, 4, size=R ** 3)
    lat = np.zeros((R ** 3, S), np.uint32)
    for v, n in enumerate(occ):
        lat[v, :n] = rng.integers(1, ns + 1, size=n)
    regions = (np.arange(R ** 3) % 2).astype(np.uint32)
    D = np.zeros((2, 2, ns + 1), np.float32)
    D[0, 0, 1:] = [0.05, 0.08, 0.02]
    D[1, 1, 1:] = [0.03, 0.1, 0.01]
    D[0, 1, 1] = 0.04  # only species 1 crosses from region 0 to 1; nothing crosses back
    o = make(oracle_mod, (R, R, R), S, lat.ravel(), ns, D=D, regions=regions, nreg=2)
    before = np.bincount(lat.ravel(), minlength=ns + 1)
    in1_before = [int(np.count_nonzero(lat[regions == 1] == sp)) for sp in range(1, ns + 1)]
    for t in range(30):
        o.step(int(rng.integers(2**32)))
    L = o.lattice.reshape(R ** 3, S)
    # rejected claims keep every invariant; a slot-S alias would break occupancy == non-empty slots
    assert o.counters.alias_writes == 0
    assert np.array_equal(np.bincount(L.ravel(), minlength=ns + 1)[1:], before[1:])
    assert np.array_equal(np.count_nonzero(L, axis=1), o.occupancy)
    C = o.counts_array()
    for sp in range(1, ns + 1):
        assert np.array_equal(np.count_nonzero(L == sp, axis=1), C[:, sp])
    in1_after = [int(np.count_nonzero(L[regions == 1] == sp)) for sp in range(1, ns + 1)]
    assert in1_after[1] == in1_before[1] and in1_after[2] == in1_before[2]  # species 2, 3 cannot cross
    assert in1_after[0] >= in1_before[0]
    nz = L != 0
    assert not np.any(~nz[:, :-1] & nz[:, 1:])  # compacted


# ------------------------------------------------------------- view ----
def test_view_majority_ties_slicing_and_label_views(oracle_mod):
    # final_stage.wgsl:35-67 on a 1x1x4 column, S=4: majority species (ties keep the first in slot order,
    # empty voxels give 0) / 255; res[axis] - coord[axis] < slice hides a voxel; views 1 and 2 show labels
    slots = [2, 3, 3, 0,  3, 2, 0, 0,  0, 0, 0, 0,  1, 1, 2, 2]
    regions = np.array([0, 1, 1, 0], np.uint32)
    o = oracle_mod.Oracle((1, 1, 4), 4, 1.0, 1.0, np.zeros((2, 2, 4), np.float32), np.zeros((1, 4), np.int32),
                          np.zeros((1, 3), np.int32), np.zeros(1, np.float32), regions, np.asarray(slots, np.uint32),
                          np.array([0, 0, 3, 1], np.uint32))
    f = np.float32
    assert list(o.view(0)) == [f(3) / f(255), f(3) / f(255), 0.0, f(1) / f(255)]
    assert list(o.view(0, axis=2, slice_=2)) == [f(3) / f(255), f(3) / f(255), 0.0, 0.0]  # z = 3: 4 - 3 < 2
    assert list(o.view(0, axis=2, slice_=5)) == [0.0] * 4
    assert list(o.view(1)) == [0.0, f(1) / f(255), f(1) / f(255), 0.0]
    assert list(o.view(2)) == [0.0, 0.0, f(3) / f(255), f(1) / f(255)]


def test_view_follows_the_occupancy_word(oracle_mod):
    # voxel 1 of the reservoir case above holds [C] with an occupancy word of 2: the majority runs over
    # slots 0..1 = [C, 0], a tie that keeps C
    slots = [0] * 32
    slots[4:8] = [1, 2, 0, 0]
    o = _reservoir_column(oracle_mod, slots, [0, 0, 0, 0, 1, 0, 0, 0], [[0, 0, -1, 1]], [[2]], [1e30])
    o.step(100)
    assert o.occupancy[1] == 2 and list(o.lattice[4:8]) == [3, 0, 0, 0]
    assert o.view(0)[1] == np.float32(3) / np.float32(255)


def _run_pair(oracle_mod, sim, seed, nsteps, threads):
    """The serial oracle and its multi-threaded driver from the same state: identical whole state and counters."""
    from cell_simulation_amd import default_itime_schedule

    a = oracle_mod.Oracle.from_simulation(sim)
    b = oracle_mod.Oracle.from_simulation(sim)
    its = default_itime_schedule(seed, nsteps)
    a.run(its)
    b.run_mt(its, threads)
    assert np.array_equal(a.lattice, b.lattice)
    assert np.array_equal(a.occupancy, b.occupancy)
    assert np.array_equal(a.counts, b.counts)
    assert np.array_equal(a.stat, b.stat)
    assert a.counters.as_dict() == b.counters.as_dict()
    return a


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_oracle_matches_serial(oracle_mod, threads):
    """cso_run_mt (hop draws, reactions and compaction on several threads, claims serial) gives the serial
    oracle's state bit for bit: a dense S=4 lattice with rejected claims and slot-S aliases, a reservoir model
    and a homodimer row that takes counts below zero."""
    from cell_simulation_amd import LatticeParams, RegionType, Simulation

    lam = 32e-9
    sim = Simulation(LatticeParams([14 * lam] * 3, [14] * 3, 50e-6, lam, 4))
    sim.set_seed(3)
    sim.add_region(RegionType.Cube("left", [0, 0, 0], [7 * lam, 14 * lam, 14 * lam]), 2e-12)
    sim.prepare_regions()
    sim.add_particle_count("A", "background", 1200, True, False)
    sim.add_particle_count("B", "left", 1200, True, False)
    sim.add_particle_count("R", "left", 60, True, True)
    sim.set_diffusion_rate_particle("A", "background", 3e-12)
    sim.set_transition_rate_particle("A", "background", "left", 1e-12)
    sim.add_reaction(["A", "B"], ["B", "B"], 60.0)
    sim.add_reaction(["A", "A"], ["B"], 400.0)
    sim.add_reaction(["R"], ["R", "A"], 900.0)
    a = _run_pair(oracle_mod, sim, 3, 25, threads)
    assert a.counters.alias_writes > 0 and a.counters.failed_claims > 0 and a.counters.fired > 0
    assert np.any(a.counts_array()[:, 1:] < 0)
