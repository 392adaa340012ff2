"""bench.py's command line on the CPU: a request for more than one GPU (by
--gpus or by a multi-rank launcher's WORLD_SIZE) exits with status 2 and a
message before anything touches a GPU; the timestep is measured on one GPU."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("argv,world", [(["--gpus", "2"], "1"), (["--gpus", "8"], "1"), ([], "4")])
def test_more_than_one_gpu_exits(argv, world):
    env = dict(os.environ, WORLD_SIZE=world)
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + argv, cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 2
    assert "measures one GPU" in p.stderr and p.stdout == ""
