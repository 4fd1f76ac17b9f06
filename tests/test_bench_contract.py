"""CPU checks of bench.py's command line: a multi-GPU request ends with a message and exit
status 2 before anything touches a GPU, and the reconstruction and solver choices parse."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run_bench("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert r.returncode == 2, r.stderr
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_help_lists_the_stage_choices():
    r = run_bench("--help")
    assert r.returncode == 0
    for opt in ("--recon", "--riemann", "--yz-shape", "--no-cpu", "--steps", "--warmup"):
        assert opt in r.stdout
    for choice in ("weno5z", "plm", "mp5", "constant", "hll", "llf"):
        assert choice in r.stdout
