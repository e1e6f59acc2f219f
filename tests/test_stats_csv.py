"""Statistics time series and CSV output (examples/no_window.rs:17-62, 147-151): window bound, mean, last, the
';'-delimited `iteration;value` records with f32 values printed as Rust prints them, and the hand-off from the
solver's drained samples."""
import numpy as np

from cell_simulation_amd import SolverStatisticSample, StatisticContainer, collect_stats, make_all_stats
from cell_simulation_amd.stats import _f32_text


def test_window_keeps_the_newest_samples():
    c = StatisticContainer(3)
    for i in range(5):
        c.add(i * 10, float(i))
    assert list(c.x) == [20, 30, 40] and list(c.y) == [2.0, 3.0, 4.0]
    assert c.last() == 4.0
    assert c.mean() == 3.0


def test_csv_records(tmp_path):
    c = StatisticContainer(10)
    c.add(0, 2000)
    c.add(25, 1999)
    c.add(50, 0.1)
    c.add(75, 1.5e-7)
    p = tmp_path / "A.csv"
    c.to_csv(str(p))
    assert p.read_text() == "0;2000\n25;1999\n50;0.1\n75;0.00000015\n"
    assert not c.x and not c.y  # written samples leave the window, as the reference's to_csv drains it


def test_f32_text_round_trips():
    rng = np.random.default_rng(3)
    for v in rng.uniform(-1e6, 1e6, 200).astype(np.float32).tolist() + [0.1, 1e-7, 3.0, -2.5]:
        t = _f32_text(v)
        assert "e" not in t
        assert np.float32(float(t)) == np.float32(v)


def test_collect_from_solver_samples():
    class FakeSim:
        def __init__(self):
            from collections import deque
            self.stats = deque([SolverStatisticSample("A", 7, 0), SolverStatisticSample("B", 3, 0),
                                SolverStatisticSample("A", 6, 25)])

    sim = FakeSim()
    stats = make_all_stats(["A"])
    assert collect_stats(sim, stats) == 2
    assert list(stats["A"].x) == [0, 25] and list(stats["A"].y) == [7.0, 6.0]
    assert not sim.stats
