"""Transcribe the known-answer constants of the reference's test suite into
tests/golden/known_answers.json.

Reads the reference test files as text (never imports them) and extracts the
numeric literals of the named arrays:

  test_diffusion_vs_hydrolib.py  R_hydrolib_infty, HYDROLIB two-sphere
                                 resistance matrices at s = 3 .. 2.001, the
                                 three-in-line 3000 matrix, pecnut RPY
                                 mobilities
  test_diffusion_vs_bdbox.py     BD_BOX Ewald diffusion coefficients

    python tests/golden/extract_known_answers.py [/path/to/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
TESTS = os.path.join(REF, "src", "bdsim", "tests")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")

NUM = r"[-+]?\d+\.\d*(?:[eE][-+]?\d+)?|[-+]?\d+(?:[eE][-+]?\d+)?"


def array_after(text, name, start=0):
    """Numbers of the np.array( [...] ) literal assigned to `name` at or after `start`."""
    i = text.index(name + " = np.array(", start)
    j = text.index("] ] )", i)
    body = text[text.index("[", i):j + 3]
    rows = re.findall(r"\[([^\[\]]+)\]", body)
    return [[float(x) for x in re.findall(NUM, r)] for r in rows], j


def func_body(text, fname):
    i = text.index("def " + fname + "(")
    k = text.find("\tdef ", i + 5)
    return text[i:k if k > 0 else len(text)]


def main():
    hyd = open(os.path.join(TESTS, "test_diffusion_vs_hydrolib.py")).read()
    out = {"source": "reference src/bdsim/tests (values transcribed as data)"}
    out["R_hydrolib_infty"] = float(re.search(r"self\.R_hydrolib_infty = (" + NUM + ")", hyd).group(1))
    two = {}
    for fname, s, var, tol in [("test_JO_R2B_3", 3.0, "R_hydrolib_3", 0.001),
                               ("test_JO_R2B_2_5", 2.5, "R_hydrolib_2_5", 0.001),
                               ("test_JO_R2B_2_1", 2.1, "R_hydrolib_2_1", 0.002),
                               ("test_JO_R2B_2_01", 2.01, "R_hydrolib_2_01", 0.005),
                               ("test_JO_R2B_2_001", 2.001, "R_hydrolib_2_001", 0.005)]:
        body = func_body(hyd, fname)
        arr, _ = array_after(body, var)
        two[str(s)] = {"separation": s, "radius": 1.0, "R": arr, "delta": tol,
                       "cite": "test_diffusion_vs_hydrolib.py " + fname}
    out["hydrolib_two_sphere"] = two
    body = func_body(hyd, "test_JO_R_3inline_3000")
    arr, _ = array_after(body, "R_hydrolib_3inline_3000")
    out["hydrolib_3inline_3000"] = {"x": [0.0, 3000.0, 6000.0], "radius": 1.0, "R": arr, "delta": 1e-7,
                                    "cite": "test_diffusion_vs_hydrolib.py test_JO_R_3inline_3000"}
    pec = {}
    for fname, xs in [("test_pecnut_RPY", [-5.0, 0.0, 7.0]), ("test_pecnut_RPY_3inline_3", [-3.0, 0.0, 3.0])]:
        body = func_body(hyd, fname)
        arr, _ = array_after(body, "M_infinity_pecnut")
        pec[fname] = {"x": xs, "radius": 1.0, "M": arr, "places": 7,
                      "cite": "test_diffusion_vs_hydrolib.py " + fname}
    out["pecnut_rpy"] = pec

    bd = open(os.path.join(TESTS, "test_diffusion_vs_bdbox.py")).read()
    cases = []
    for m in re.finditer(r"def (test_RPY_Smith_\w+)\(self\):(.*?)(?=\n\tdef |\Z)", bd, re.S):
        name, body = m.group(1), m.group(2)
        beads = re.findall(r"Bead\(\[(" + NUM + r"), (" + NUM + r"), (" + NUM + r")\], (" + NUM + r")\)", body)
        vals = dict(re.findall(r"(D_bdbox_\w+) = (" + NUM + ")", body))
        L = float(re.search(r"box_size = (" + NUM + ")", body).group(1))
        cases.append({"name": name, "box_length": L,
                      "beads": [[float(b[0]), float(b[1]), float(b[2]), float(b[3])] for b in beads],
                      "values": {k: float(v) for k, v in vals.items()}, "m": 2, "n": 2, "delta": 1e-4,
                      "cite": "test_diffusion_vs_bdbox.py " + name})
    out["bdbox_smith"] = cases
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT, "(%d hydrolib, %d pecnut, %d bdbox)" % (len(two), len(pec), len(cases)))


if __name__ == "__main__":
    main()
