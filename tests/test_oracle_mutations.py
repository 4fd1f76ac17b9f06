"""The oracle pins catch the mistakes they are meant to: each planted mutation of the
oracle's source (a dropped WENO5-Z / MC blend, the blend's MC slope 2 turned to 1.5, LLF's
dissipation sign flipped, HLL's sqrt(gamma) dissipation factor turned to sqrt(-g), the
Valencia energy source without its lapse-gradient term) is built into a library of its own
in a temporary folder, and the pin test named for it must fail against that library (and
pass against the real oracle).  CPU only: gcc and the oracle's source."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "phx_oracle.c")

# (name, text in phx_oracle.c, replacement, pin tests that must fail)
MUTATIONS = [
    ("weno_blend_dropped",
     "const double alpha_lin = 2.0 * alpha_l * alpha_r / (alpha_l + alpha_r);",
     "const double alpha_lin = 1.0;",
     ["test_weno5z_blend_exact_rationals"]),
    ("weno_mc_slope_1p5",
     "dq = mc_lim(q2 - q1, dq, 2.0);",
     "dq = mc_lim(q2 - q1, dq, 1.5);",
     ["test_weno5z_blend_exact_rationals"]),
    ("llf_dissipation_sign",
     "flux[m] = 0.5 * ((Fl[m] + Fr[m]) * g.detg - cmax * ((Ur[m] - Ul[m]) * g.detgamma));",
     "flux[m] = 0.5 * ((Fl[m] + Fr[m]) * g.detg + cmax * ((Ur[m] - Ul[m]) * g.detgamma));",
     ["test_llf_curved_textbook_form", "test_hll_equals_llf_on_mirror_symmetric_faces"]),
    ("hll_sqrt_gamma_to_sqrt_g",
     "crcl * (Ur[m] - Ul[m]) * g.detgamma) * inv;",
     "crcl * (Ur[m] - Ul[m]) * g.detg) * inv;",
     ["test_hll_curved_textbook_form"]),
    ("energy_source_lapse_term_dropped",
     "src[4] = gdet * alpha * (Ta - TGam);",
     "src[4] = gdet * alpha * (0.0 - TGam);",
     ["test_energy_source_matches_extrinsic_curvature_form"]),
]


def _build(tmp, text, repl):
    src = open(SRC).read()
    assert src.count(text) == 1, f"mutation anchor not found once: {text}"
    mut = os.path.join(tmp, "phx_oracle.c")
    open(mut, "w").write(src.replace(text, repl))
    shutil.copy(os.path.join(ROOT, "oracle", "phx_oracle.h"), tmp)
    so = os.path.join(tmp, "libphx_oracle.so")
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
                    "-std=c99", "-shared", "-o", so, mut, "-lm"], check=True)
    return so


def _run_pins(tests, lib):
    env = dict(os.environ, PHX_ORACLE_LIB=lib, OMP_NUM_THREADS="2")
    k = " or ".join(tests)
    return subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                           os.path.join(ROOT, "tests", "test_oracle_pins.py"), "-k", k],
                          env=env, capture_output=True, text=True, cwd=ROOT)


@pytest.mark.parametrize("name,text,repl,tests", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_pins_fail_on_planted_mutation(tmp_path, name, text, repl, tests):
    lib = _build(str(tmp_path), text, repl)
    r = _run_pins(tests, lib)
    assert r.returncode == 1, f"{name}: the pins did not fail\n{r.stdout[-2000:]}"
    assert "failed" in r.stdout


def test_pins_pass_on_the_oracle(tmp_path):
    lib = _build(str(tmp_path), "dq = mc_lim(q2 - q1, dq, 2.0);", "dq = mc_lim(q2 - q1, dq, 2.0);")
    r = _run_pins(sorted({t for m in MUTATIONS for t in m[3]}), lib)
    assert r.returncode == 0, r.stdout[-2000:]
