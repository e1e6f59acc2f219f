import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libcellsim.so")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import cs_oracle

    if not cs_oracle.LIB.exists():
        cs_oracle.build()
    return cs_oracle


def announce(msg: str) -> None:
    """Print a progress line that names the case before GPU work (flushed)."""
    print(f"\n[case] {msg}", flush=True)


def check_state(sim, orc, tag: str) -> None:
    """The whole reference state after a step, against the oracle: the slots (species per lattice word), the
    per-voxel counts C[v][sp], the occupancy words and the global counts (concentrations_stat)."""
    import numpy as np

    g = sim.lattice()
    o = orc.lattice_u8()
    if not np.array_equal(g, o):
        diff = np.argwhere(np.any(g != o, axis=-1))
        v = tuple(diff[0])
        raise AssertionError(f"{tag}: {len(diff)} voxels differ; first {v}: gpu {g[v]} oracle {o[v]}")
    # per-voxel counts: up to 50M entries (a 256^3 lattice of 20 species would take 1.4 GB per copy)
    gc, oc = (sim.concentrations(), orc.counts_array()) if orc.counts.size <= 50_000_000 else (None, None)
    if gc is not None and not np.array_equal(gc, oc):
        bad = np.argwhere(np.any(gc != oc, axis=1))[:, 0]
        raise AssertionError(f"{tag}: counts of {len(bad)} voxels differ; first {bad[0]}: gpu {gc[bad[0]]} "
                             f"oracle {oc[bad[0]]}")
    go = sim.occupancy()
    if not np.array_equal(go, orc.occupancy):
        bad = np.nonzero(go != orc.occupancy)[0]
        raise AssertionError(f"{tag}: occupancy of {len(bad)} voxels differs; first {bad[0]}: gpu {go[bad[0]]} "
                             f"oracle {orc.occupancy[bad[0]]}")
    assert np.array_equal(sim.species_counts()[1:], orc.stat[1:]), (tag, sim.species_counts(), orc.stat)


def skewed_voxels(orc) -> int:
    """Voxels whose occupancy word or per-voxel counts differ from their slots (aliases, negative counts,
    reservoir units)."""
    import numpy as np

    N, S = orc.N, orc.S
    lat = orc.lattice.reshape(N, S).astype(np.int64)
    occ = np.count_nonzero(lat, axis=1)
    cnt = np.zeros((N, orc.ns + 1), np.int64)
    for s in range(S):
        np.add.at(cnt, (np.arange(N), lat[:, s]), 1)
    c = orc.counts_array().astype(np.int64)
    off = (occ != orc.occupancy.astype(np.int64)) | np.any(cnt[:, 1:] != c[:, 1:], axis=1)
    return int(np.count_nonzero(off))
