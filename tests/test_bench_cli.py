"""bench.py's command line on the CPU: the one-GPU contract (--gpus above 1
exits with a message before anything touches a GPU) and the options of the
accelerated forms."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    p = run("--gpus", "2")
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_options_of_the_accelerated_forms():
    p = run("--help")
    assert p.returncode == 0
    for opt in ("--gcr", "--dc-sweeps", "--level2-sweeps", "--coarse-sweeps", "--guess", "--mode", "--steps",
                "--warmup"):
        assert opt in p.stdout, opt
