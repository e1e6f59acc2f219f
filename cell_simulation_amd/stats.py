"""Statistics time series and their CSV output, as the reference's `no_window` example keeps and writes them.

`StatisticContainer` mirrors examples/no_window.rs:17-62: a bounded window of (iteration_count, value) samples per
metric (the oldest leave first once the window is full), its mean and last value, and `to_csv`, which writes the
samples as `iteration;value` records (';' delimiter, no header) and empties the window, as the reference's
`to_csv` drains its deques. Values are f32 in the reference and are printed as Rust prints an f32 (`{}`): the
shortest decimal that reads back to the same f32, without an exponent.

`collect_stats(simulation, containers)` moves the simulation's drained solver samples (src/statistics.rs:4-9,
read back every `write_freq` steps) into the containers by metric name, as no_window.rs:147-151 does.
"""
from __future__ import annotations

from collections import deque
from typing import Dict, Iterable

import numpy as np


def _f32_text(y: float) -> str:
    v = np.float32(y)
    if np.isfinite(v) and v == np.floor(v) and abs(v) < 2.0 ** 63:
        return str(int(v))
    if np.isnan(v):
        return "NaN"
    if np.isinf(v):
        return "inf" if v > 0 else "-inf"
    return np.format_float_positional(v, unique=True, trim="-")


class StatisticContainer:
    """examples/no_window.rs:17-62"""

    def __init__(self, capacity: int = 500):
        if capacity < 1:
            raise ValueError("capacity must be positive")
        self.capacity = int(capacity)
        self.x: deque = deque()
        self.y: deque = deque()

    def add(self, x: int, y: float) -> None:
        while len(self.x) >= self.capacity:
            self.x.popleft()
            self.y.popleft()
        self.x.append(int(x) & 0xFFFFFFFF)
        self.y.append(float(np.float32(y)))

    def mean(self) -> float:
        if not self.y:
            return float("nan")
        acc = np.float32(0.0)
        for v in self.y:  # f32 running sum, as the reference's loop
            acc = np.float32(acc + np.float32(v))
        return float(np.float32(acc / np.float32(len(self.y))))

    def last(self) -> float:
        return self.y[-1] if self.y else 0.0

    def to_csv(self, path: str) -> None:
        with open(path, "w", newline="") as f:
            while self.x and self.y:
                f.write(f"{self.x.popleft()};{_f32_text(self.y.popleft())}\n")


def make_all_stats(metrics: Iterable[str], capacity: int = 500) -> Dict[str, StatisticContainer]:
    """examples/no_window.rs:64-70"""
    return {m: StatisticContainer(capacity) for m in metrics}


def collect_stats(simulation, containers: Dict[str, StatisticContainer]) -> int:
    """Moves `simulation.stats` (drained solver samples) into the containers of their metric names; samples of
    metrics without a container are dropped, as `HashMap::entry(..).and_modify` leaves absent keys alone.
    Returns the number of samples moved."""
    moved = 0
    while simulation.stats:
        s = simulation.stats.popleft()
        c = containers.get(s.name)
        if c is not None:
            c.add(s.iteration_count, s.value)
            moved += 1
    return moved
