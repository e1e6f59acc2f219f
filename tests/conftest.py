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
