"""The oracle's lattice reactions against an f64 closed form: isomerisation statistics (mean and variance
within 3 sigma) over replicate cells, on the CPU.

Each replicate is a 32^3 lattice with 40 particles that diffuse and isomerise, A <-> B (first order,
k1 = 200 /s, k2 = 100 /s, dt = 50 us), stepped with its own seed and itime schedule. A particle alone in its
voxel converts with probability 1 - exp(-k dt) per step (cme.wgsl:66: r1 <= 1 - exp(-tot * tau), tot = k for one
particle), so each particle's species is a two-state chain with a = 1 - exp(-k1 dt), b = 1 - exp(-k2 dt): after n
steps it is A with probability p = b / (a + b) + a / (a + b) * (1 - a - b)^n, and a cell's A count is
Binomial(40, p) (voxels holding two particles, about 1% of the particle-steps here, differ from independent
flips by O((k dt)^2)). A sign or index slip in the propensity, the rewrite or the draw moves the mean or the
variance by many sigma. The GPU ensemble runs the same cells (tests/test_gpu_ensemble_stats.py).
"""
import math

import numpy as np

from cell_simulation_amd import LatticeParams, RegionType, Simulation, default_itime_schedule

N0, EDGE, SLOTS = 40, 32, 4
K1, K2, TAU, N_STEPS = 200.0, 100.0, 50e-6, 100


def isomerisation_cell(seed: int) -> Simulation:
    lam = 32e-9
    sim = Simulation(LatticeParams([EDGE * lam] * 3, [EDGE] * 3, TAU, lam, SLOTS))
    sim.set_seed(seed)
    sim.add_region(RegionType.Cube("cell", [0.0] * 3, [EDGE * lam] * 3), 1e-12)
    sim.prepare_regions()
    sim.add_particle_count("A", "cell", N0, True, False)
    sim.add_particle_count("B", "cell", 0, True, False)
    sim.add_reaction(["A"], ["B"], K1)
    sim.add_reaction(["B"], ["A"], K2)
    return sim


def cell_seed(i: int) -> int:
    return 3000 + i


def itime_seed(i: int) -> int:
    return 90001 + 17 * i


def check_a_counts(n_a: np.ndarray) -> str:
    """Sample mean and variance of the A counts within 3 sigma of Binomial(N0, p) (f64 closed form)."""
    n = len(n_a)
    a, b = 1.0 - math.exp(-K1 * TAU), 1.0 - math.exp(-K2 * TAU)
    p = b / (a + b) + a / (a + b) * (1.0 - a - b) ** N_STEPS
    mean, var = N0 * p, N0 * p * (1.0 - p)
    m4 = N0 * p * (1 - p) * (1 + 3 * (N0 - 2) * p * (1 - p))  # binomial fourth central moment
    sigma_mean = math.sqrt(var / n)
    sigma_var = math.sqrt((m4 - (n - 3) / (n - 1) * var * var) / n)  # of the sample variance
    got_mean, got_var = float(n_a.mean()), float(n_a.var(ddof=1))
    msg = (f"A count over {n} cells: mean {got_mean:.3f} (closed form {mean:.3f} +- {sigma_mean:.3f}), "
           f"variance {got_var:.3f} (closed form {var:.3f} +- {sigma_var:.3f})")
    assert abs(got_mean - mean) < 3 * sigma_mean, msg
    assert abs(got_var - var) < 3 * sigma_var, msg
    return msg


def species_counts(lat: np.ndarray) -> tuple:
    n_a, n_b = int(np.count_nonzero(lat == 1)), int(np.count_nonzero(lat == 2))
    assert n_a + n_b == N0 and int(np.count_nonzero(lat > 2)) == 0, "particles lost or gained"
    return n_a, n_b


def test_oracle_isomerisation_mean_and_variance(oracle_mod):
    n_cells = 150
    n_a = np.empty(n_cells, np.int64)
    for i in range(n_cells):
        orc = oracle_mod.Oracle.from_simulation(isomerisation_cell(cell_seed(i)))
        orc.run(default_itime_schedule(itime_seed(i), N_STEPS))
        n_a[i] = species_counts(orc.lattice_u8())[0]
    print(check_a_counts(n_a))


def test_closed_form_limits():
    """The chain's closed form: p = 1 at n = 0 and the stationary b / (a + b) as n grows, and it matches the
    continuous-time two-state solution to O(k dt)."""
    a, b = 1.0 - math.exp(-K1 * TAU), 1.0 - math.exp(-K2 * TAU)
    pn = lambda n: b / (a + b) + a / (a + b) * (1.0 - a - b) ** n  # noqa: E731
    assert pn(0) == 1.0
    assert abs(pn(10 ** 6) - b / (a + b)) < 1e-12
    t = N_STEPS * TAU
    cont = K2 / (K1 + K2) + K1 / (K1 + K2) * math.exp(-(K1 + K2) * t)
    assert abs(pn(N_STEPS) - cont) < (K1 + K2) * TAU
