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


def test_stage_profile_digest_covers_every_sweep_unit():
    """The digest that dates the committed stage profiles covers every translation unit
    the stage's kernels are compiled from: each .hip file that includes the sweeps header
    or defines the stage's update, CT or dt kernels."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    listed = set(bench.STAGE_SOURCES)
    csrc = os.path.join(ROOT, "phoebus_amd", "csrc")
    units = {"phx_stage.hip", "phx_kernels.hip"}  # the update, EMF, dt and p2c kernels
    for name in sorted(os.listdir(csrc)):
        if name.endswith(".hip") and \
                '#include "phx_flux_sweeps.hpp"' in open(os.path.join(csrc, name)).read():
            units.add(name)
    assert len(units) >= 6
    for name in sorted(units):
        assert "phoebus_amd/csrc/" + name in listed, name
    for f in listed:
        assert os.path.exists(os.path.join(ROOT, f)), f
    assert len(bench.sources_digest()) == 64
