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
    assert C.sizeof(A.BlockFields) == 88  # 10 field pointers + the vars table pointer
    assert C.sizeof(A.BlockVars) == 8 * (10 + 8 + 24 + 8)
    assert C.sizeof(A.StageArgs) == 8 * 10
    assert C.sizeof(A.FluidParams) % 8 == 0 and A.FluidParams.flux_arith.offset > 0


def test_ctypes_layouts_match_the_c_compiler(tmp_path):
    """sizeof / offsetof of the ABI structs as gcc lays them out from include/*.h, against
    the ctypes mirror the tests and the bench use."""
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stddef.h>
#include <stdio.h>
#include "phx.h"
#include "phx_pack.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(phx_fluid_params), offsetof(phx_fluid_params, flux_arith),
         sizeof(phx_block_fields), offsetof(phx_block_fields, vars));
  printf("%zu %zu %zu\n", sizeof(phx_block_vars), offsetof(phx_block_vars, flux),
         offsetof(phx_block_vars, src));
  printf("%zu %zu %zu\n", sizeof(phx_stage_args), offsetof(phx_stage_args, base_vars),
         offsetof(phx_stage_args, out_vars));
  printf("%zu\n", sizeof(phx_pack_block));
  return 0;
}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = [C.sizeof(A.FluidParams), A.FluidParams.flux_arith.offset,
            C.sizeof(A.BlockFields), A.BlockFields.vars.offset,
            C.sizeof(A.BlockVars), A.BlockVars.flux.offset, A.BlockVars.src.offset,
            C.sizeof(A.StageArgs), A.StageArgs.base_vars.offset, A.StageArgs.out_vars.offset,
            C.sizeof(A.PackBlock)]
    assert got == want
