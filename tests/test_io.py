"""Driver I/O in the reference's formats (pyBrown/input.py, output.py).
Host code only."""
import json

import numpy as np
import pytest

from pybrown_amd import io

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def test_read_str_file(tmp_path):
    p = tmp_path / "t.str"
    p.write_text("sub AB 7 1.5 -2.0 3.25 4.0 0.0 6.0 0.1 1.0\n"
                 "sub C 9 0.0 0.0 1e1 2.0 0.0 3.0 0.0 1.0\n"
                 "bond 7 9 5.0 x 100.0\n"
                 "angle x 7 9 7 120.0 3.0\n")
    b = io.read_str_file(str(p))
    assert [x.label for x in b] == ["AB", "C"] and [x.bead_id for x in b] == [7, 9]
    assert np.array_equal(b[0].r, [1.5, -2.0, 3.25]) and b[0].a == 4.0
    assert b[0].hard_core_radius == 3.0 and b[0].epsilon_LJ == 0.1   # 2r column halved
    assert b[0].bonded_with == [9] and b[0].bonded_how[9] == [5.0, 100.0]
    assert b[0].angled_with == [[9, 7]] and b[0].angled_how[(9, 7)] == [120.0, 3.0]


def test_grid_fixture_parses():
    b = io.read_str_file(GOLDEN + "/grid27.str")
    assert len(b) == 27 and {x.label for x in b} == {"A", "B"}


def test_input_data(tmp_path):
    p = tmp_path / "in.json"
    p.write_text(json.dumps({"dt": 1.0, "T": 293.15}))
    i = io.InputData(str(p), ["dt"], {"seed": None, "T": 1.0}, ["dt", "T", "seed"])
    assert i.input_data == {"dt": 1.0, "T": 293.15, "seed": None}
    with pytest.raises(AssertionError, match="Missing box_length"):
        io.InputData(str(p), ["box_length"])
    with pytest.raises(AssertionError, match="Unrecognized T"):
        io.InputData(str(p), [], {}, ["dt"])


class _B:
    def __init__(self, label, r):
        self.label, self.r = label, np.array(r, dtype=np.float64)


def test_xyz_and_flux_text(tmp_path):
    f = tmp_path / "o.xyz"
    with open(f, "w") as h:
        io.write_to_xyz_file(h, "o.xyz", 3, 0.5, [_B("A", [1.0, -2.5, 1e-7]), _B("B", [0.1, 0.2, 0.3])])
    assert f.read_text() == "2\no.xyz time [ps] 1.5\nA 1.0 -2.5 1e-07\nB 0.1 0.2 0.3\n"
    g = tmp_path / "o.flx"
    with open(g, "w") as h:
        io.write_to_flux_file(h, 0, 2.0, {"A": 0, "B": 0})
        io.write_to_flux_file(h, 1, 2.0, {"A": -1, "B": 3})
    assert g.read_text() == "time/ps A B\n2.0 -1 3\n"


def test_truncate(tmp_path):
    f = tmp_path / "x.txt"
    f.write_text("a\nb\nc\nd\n")
    io.truncate_output_file_during_restart(str(f), 2)
    assert f.read_text() == "a\nb\n"


# ---------------------------------------------------------------------------
# Output text against literal known answers in the reference's formats, so
# the writers are pinned independently of themselves.  The expected strings
# spell out what Python's `str.format` with '{}' fields produces for these
# inputs, field by field, as output.py builds the lines:
#   .xyz  output.py:63-68   'N', '<file> time [ps] <j*dt>', '<label> <x> <y> <z>'
#   .enr  output.py:108-120 header 'time/ps energy/<unit>' at j = 0, then '<j*dt> <E>'
#   .flx  output.py:122-140 header 'time/ps <labels>' at j = 0, else '<j*dt> <counts>'
# The inputs include the float spellings '{}' gives: shortest round-trip
# digits, exponent forms, -0.0, numpy scalars, and j*dt rounding
# (3 * 0.1 = 0.30000000000000004).

class _Bead:
    def __init__(self, label, r):
        self.label, self.r = label, r


def test_xyz_text_known_answer():
    import io as pyio
    beads = [_Bead("A", np.array([1.0, -2.5, 1e-20])), _Bead("Bq", np.array([123456789.125, -0.0, 0.1])),
             _Bead("C", [np.float64(3.0), 7, 2.5e16])]
    f = pyio.StringIO()
    io.write_to_xyz_file(f, "traj.xyz", 3, 0.1, beads)
    assert f.getvalue() == ("3\n"
                            "traj.xyz time [ps] 0.30000000000000004\n"
                            "A 1.0 -2.5 1e-20\n"
                            "Bq 123456789.125 -0.0 0.1\n"
                            "C 3.0 7 2.5e+16\n")


def test_flux_text_known_answer():
    import io as pyio
    f = pyio.StringIO()
    flux = {"XXX": 0, "Y": -3}
    io.write_to_flux_file(f, 0, 0.5, flux)
    io.write_to_flux_file(f, 1, 0.5, {"XXX": 2, "Y": -3})
    io.write_to_flux_file(f, 7, 0.1, {"XXX": np.int64(5), "Y": 0})
    assert f.getvalue() == ("time/ps XXX Y\n"
                            "0.5 2 -3\n"
                            "0.7000000000000001 5 0\n")


def test_enr_text_known_answer():
    import io as pyio
    from pybrown_amd.bd import write_to_enr_file
    f = pyio.StringIO()
    write_to_enr_file(f, 0, 0.1, 0.0, "kcal/mol")
    write_to_enr_file(f, 3, 0.1, -1.25e-05, "kcal/mol")
    assert f.getvalue() == ("time/ps energy/kcal/mol\n"
                            "0.0 0.0\n"
                            "0.30000000000000004 -1.25e-05\n")
