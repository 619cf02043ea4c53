"""bench.py's command line without a GPU: a request for more than one GPU
ends at once with a message and a non-zero status, before any device work."""
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _run(args, env=None):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, env=env)


def test_more_than_one_gpu_is_refused():
    p = _run(["--gpus", "2"])
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_a_distributed_launch_is_refused():
    env = dict(os.environ, WORLD_SIZE="8")
    p = _run(["--gpus", "1"], env=env)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""
