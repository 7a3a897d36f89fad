"""CPU checks of bench.py's command line: the runs it refuses before any
device work (more than one GPU; a size group the batched form does not have)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2 and "measures one GPU" in r.stderr and not r.stdout


def test_group_alone_needs_a_size_of_the_batched_form():
    r = run("--group-alone", "100", "--steps", "2")
    assert r.returncode == 2 and "no 100^3 solves" in r.stderr and "32, 64, 128, 192" in r.stderr
