"""CPU: bench.py's command line (no GPU is touched on these paths)."""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_options_listed():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--photons", "--tile", "--legs", "--autotune",
                "--calibrate-copy", "--no-cpu-baseline"):
        assert opt in r.stdout
