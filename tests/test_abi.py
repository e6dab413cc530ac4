"""The C-ABI library loads and exports every symbol include/pybrown_amd.h
declares; the ctypes mirror of pbx_config matches the C layout.  No compute
calls (these run without a GPU)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "pybrown_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("typedef", "int", "void", "double", "char")))


def test_header_declares_reference_entry_points():
    names = _declared()
    for ref in ("RPY_M_matrix", "RPY_Smith_M_matrix", "JO_2B_R_matrix", "JO_R_lubrication_correction_matrix",
                "pointer_pbc_matrix", "pointer_immobile_pbc_matrix", "check_overlaps"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    from pybrown_amd import _lib
    missing = [n for n in _declared() if not hasattr(_lib.lib, n)]
    assert not missing, missing
    assert set(_declared()) <= set(_lib.EXPORTED) | set(_declared())


def test_config_layout_matches_c():
    from pybrown_amd import _lib
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "pybrown_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(pbx_config), offsetof(pbx_config, max_move_attempts),
         offsetof(pbx_config, box_length), offsetof(pbx_config, flux_normal), offsetof(pbx_config, external_force));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        r = subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), c, "-o", exe], capture_output=True)
        if r.returncode != 0:
            pytest.skip("gcc unavailable")
        out = subprocess.run([exe], capture_output=True, text=True).stdout.split()
    C = _lib.Config
    want = [ctypes.sizeof(C), C.max_move_attempts.offset, C.box_length.offset, C.flux_normal.offset,
            C.external_force.offset]
    assert [int(x) for x in out] == want


def test_library_is_gfx950_code_object():
    from pybrown_amd import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data
