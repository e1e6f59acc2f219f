"""bench.py's command line on the CPU: a run on more than one GPU (by flag or by a launcher's WORLD_SIZE) is refused
with a message and exit status 2 before any device work; the defaults follow the workloads' run lengths."""
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _bench(args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True, env=e,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    for args, env in ((["--gpus", "2"], None), (["--gpus", "8", "--steps", "1"], None), ([], {"WORLD_SIZE": "2"})):
        r = _bench(args, env)
        assert r.returncode == 2, (args, env, r.returncode, r.stderr[-500:])
        assert "exactly one GPU" in r.stderr and r.stdout == ""


def test_defaults_follow_the_run_lengths():
    r = _bench(["--help"])
    assert r.returncode == 0
    text = " ".join(r.stdout.split())
    for frag in ("run length: 2000", "5000 steps in 100-step graphs", "its run length: 1000"):
        assert frag in text, frag
