"""bench.py's command line on the CPU: the run is one GPU's, so --gpus other
than 1 exits with status 2 and says why before anything touches a device;
the help text lists the documented options."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


@pytest.mark.parametrize("n", [0, 2, 8])
def test_bench_refuses_other_gpu_counts(n):
    p = subprocess.run([sys.executable, BENCH, "--gpus", str(n)], cwd=ROOT,
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 2
    assert "runs on one GPU" in p.stderr and ("--gpus %d" % n) in p.stderr
    assert p.stdout == ""


def test_bench_help_lists_the_options():
    p = subprocess.run([sys.executable, BENCH, "--help"], cwd=ROOT,
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--n", "--poisson-n", "--chocg-n",
                "--learned-n", "--copy-probe", "--no-cpu-baseline"):
        assert opt in p.stdout, opt
