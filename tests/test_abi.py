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
  printf("%zu %zu %zu %zu\n", offsetof(pbx_config, divergence_term), offsetof(pbx_config, overlap_exempt),
         offsetof(pbx_config, n_immobile), offsetof(pbx_config, immobile_radii));
  printf("%zu %zu %zu\n", sizeof(pbx_forces), offsetof(pbx_forces, angle_params), offsetof(pbx_forces, unit));
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
    C, Fo = _lib.Config, _lib.Forces
    want = [ctypes.sizeof(C), C.max_move_attempts.offset, C.box_length.offset, C.flux_normal.offset,
            C.external_force.offset, C.divergence_term.offset, C.overlap_exempt.offset, C.n_immobile.offset,
            C.immobile_radii.offset, ctypes.sizeof(Fo), Fo.angle_params.offset, Fo.unit.offset]
    assert [int(x) for x in out] == want


def test_box_mobility_checks_on_the_host():
    """Box rejects a system without mobile beads, and flux counting for a
    mobile bead whose label first appears on an immobile bead (the reference
    would fail with a KeyError on its first crossing), before any device call."""
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    inp = {"box_length": 20.0, "T": 298.0, "viscosity": 0.01, "seed": 1, "hydrodynamics": "nohi"}
    beads = [Bead([0.0, 0.0, 0.0], 1.0, label="W"), Bead([3.0, 0.0, 0.0], 1.0, label="W")]
    with pytest.raises(ValueError, match="mobile bead"):
        Box(beads, dict(inp, immobile_labels=["W"]))
    beads = [Bead([0.0, 0.0, 0.0], 1.0, label="A", mobile=False), Bead([3.0, 0.0, 0.0], 1.0, label="A")]
    with pytest.raises(ValueError, match="measure_flux"):
        Box(beads, dict(inp, measure_flux={"normal": [1, 0, 0], "plane_point": [0, 0, 0]}))


def test_library_is_gfx950_code_object():
    from pybrown_amd import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data
