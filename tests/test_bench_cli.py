"""bench.py's command line on the CPU: more than one GPU is refused with a message and exit
status 2 before anything touches a GPU (the build measures one GPU; the path shards over
particles, which it does not run)."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_is_refused(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert f"exactly one GPU (asked for {n})" in r.stderr
    assert r.stdout == ""
