"""The C-ABI library loads and exports every entry point include/*.h declares,
and the ctypes structs match the C layouts (no GPU calls)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from tenex_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tenex_hip.h")


def declared_symbols():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"TNX_API\s+[\w\s\*]+?\b(tnx_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 15
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert set(syms) == set(_lib.SIGNATURES)


def test_abi_version_and_error_reporting():
    lib = _lib.load()
    assert lib.tnx_abi_version() == _lib.ABI_VERSION == 6
    info = _lib.PlanInfo()
    rc = lib.tnx_te_plan(None, 3, 10, None, 0, None, 0, 0, 0, None, None, None, 0, C.byref(info))
    assert rc == -1
    assert b"null pointer" in lib.tnx_last_error()


def test_struct_layouts_match_header(tmp_path):
    prog = tmp_path / "sz.c"
    prog.write_text(f'''#include <stdio.h>
#include <stddef.h>
#include "{HEADER}"
int main(void) {{
  printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(tnx_device_props_t), sizeof(tnx_work_t),
         sizeof(tnx_plan_info_t), sizeof(tnx_te_args_t), sizeof(tnx_surr_args_t),
         sizeof(tnx_mlp_t), sizeof(tnx_adam_t));
  printf("%zu %zu %zu\\n", offsetof(tnx_te_args_t, out), offsetof(tnx_surr_args_t, sum_te),
         offsetof(tnx_plan_info_t, est_cost_wave));
  return 0;
}}''')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", str(prog), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [C.sizeof(_lib.DeviceProps), C.sizeof(_lib.Work), C.sizeof(_lib.PlanInfo),
           C.sizeof(_lib.TeArgs), C.sizeof(_lib.SurrArgs), C.sizeof(_lib.Mlp),
           C.sizeof(_lib.Adam), _lib.TeArgs.out.offset, _lib.SurrArgs.sum_te.offset, _lib.PlanInfo.est_cost_wave.offset]
    assert [int(v) for v in out] == got
