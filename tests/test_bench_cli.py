"""The benchmark's command-line contract on the CPU: one GPU only, every workload listed."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    p = run("--gpus", "2")
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_workloads_are_listed():
    p = run("--help")
    assert p.returncode == 0
    for w in ("hit256", "batch", "adjoint", "relax", "dataset", "stats", "horizon", "configs", "onestep"):
        assert w in p.stdout
