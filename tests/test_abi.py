"""CPU checks of the C ABI: libphx.so loads and exports every symbol include/*.h declares.

No compute call is made (there may be no GPU here)."""
import ctypes as C
import glob
import os
import re
import subprocess

from phoebus_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(phx_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_symbols()
    for must in ("phx_calculate_fluxes", "phx_flux_ct", "phx_flux_divergence_update",
                 "phx_con2prim", "phx_estimate_dt", "phx_prim_to_con",
                 "phx_fluid_source_terms", "phx_con2prim_batch"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = A.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", A.LIBPATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    for n in declared_symbols():
        assert n in exported, n


def test_library_is_gfx950_code_object():
    blob = open(A.LIBPATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_struct_layouts_match_header():
    # sizes implied by include/phx.h on LP64
    assert C.sizeof(A.Mgr) == 8 + 8 * 8  # int + pad, 8 doubles
    assert C.sizeof(A.Geometry) == 16 + C.sizeof(A.Mgr)
    assert C.sizeof(A.Bounds) == 72
    assert C.sizeof(A.Block) == 4 + 12 + 4 + 4 + 48  # int, int[3], int, pad, double[6]
    assert C.sizeof(A.BlockFields) == 80
