"""Lubrication steps that rebuild D and factor it take both from one
factorisation of the reversed resistance matrix (pbx_inverse_and_factor).
The D and B a step leaves (box.py:537, :554) against the oracle's
D = kBT inv(R) and numpy's Cholesky of it, at orders of one and of several
tiles (pads of 43 and 12 rows), in both schemes; and the same with the
factor requested on its own (the cadence flags of BD.py:177), which takes
the separate inverse and factorisation."""
import numpy as np
import pytest

from oracle.box_oracle import OracleBox

pytestmark = pytest.mark.gpu


def _cluster(N, seed):
    from pybrown_amd.workloads import rsa_positions
    return rsa_positions(np.ones(N), 14.0, seed, min_gap=0.15)


@pytest.mark.parametrize("N", [7, 60])
@pytest.mark.parametrize("scheme", ["ermak", "midpoint"])
def test_fused_D_and_B_against_the_oracle(N, scheme):
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    xyz = _cluster(N, N)
    inp = {"box_length": 14.0, "T": 298.0, "viscosity": 0.01, "seed": 4, "hydrodynamics": "rpy_smith_lub",
           "ewald_alpha": np.sqrt(np.pi), "ewald_real": 1, "ewald_imag": 1, "lubrication_cutoff": 1.0,
           "cichocki_correction": True, "check_overlaps": False, "propagation_scheme": scheme}
    g = Box([Bead(p.copy(), 1.0) for p in xyz], inp)
    try:
        built_at = None
        for flags in [(True, True, True), (True, True, False), (False, False, True), (True, True, True)]:
            if flags[1]:
                built_at = g._pos.copy()  # D of this step: at its start positions
            g.propagate(1e-6, *flags)
            o = OracleBox(built_at, np.ones(N), ["XXX"] * N, inp)
            o._compute_Dff_matrix()
            o._compute_Dtot_matrix()
            D = g.D
            dD = np.abs(D - o.D).max() / np.abs(o.D).max()
            assert dD <= 1e-12, "flags %s: D error %.3e" % (flags, dD)
            if flags[2]:
                Bn = np.linalg.cholesky(o.D)
                dB = np.abs(g.B - Bn).max() / np.abs(Bn).max()
                assert dB <= 1e-12, "flags %s: B error %.3e" % (flags, dB)
            print("N=%d %s flags %s: D %.2e%s" % (N, scheme, flags, dD, (" B %.2e" % dB) if flags[2] else ""))
    finally:
        g.close()


def test_lean_midpoint_matches_the_explicit_path_at_71_tile_columns():
    """N = 1500 (n = 4500, 71 tile columns: the two-workgroups-per-CU
    factorisation with rolling far updates): the lean midpoint (B0 held as
    the factor of J R0 J, D_mid w from R_mid's factor and two chained solves)
    against the explicit path that a force field selects (explicit D, B and
    D_mid; here the field is a zero external force over the whole box, so the
    two steps are the same in exact arithmetic).  Per-step displacements agree
    to 1e-10 relative (rounding: solves against inverses); flux and attempts
    are identical."""
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    from pybrown_amd.workloads import rsa_positions, box_for
    a = np.full(1500, 5.0)
    Lb = box_for(a, 0.10)
    xyz = rsa_positions(a, Lb, 21, min_gap=0.5)
    base = {"box_length": Lb, "T": 293.15, "viscosity": 0.01005, "seed": 9, "hydrodynamics": "rpy_smith_lub",
            "ewald_alpha": np.sqrt(np.pi), "ewald_real": 1, "ewald_imag": 1, "lubrication_cutoff": 1.0,
            "propagation_scheme": "midpoint", "m_midpoint": 100, "check_overlaps": False,
            "measure_flux": {"normal": [1.0, 0.0, 0.0], "plane_point": [0.0, 0.0, 0.0]}}
    forced = dict(base, external_force=[0.0, 0.0, 0.0], external_force_region={"x": [-1e9, 1e9]})
    boxes = [Box([Bead(p.copy(), 5.0) for p in xyz], inp) for inp in (base, forced)]
    try:
        prev = [np.array([b.r for b in bx.beads]) for bx in boxes]
        for step in range(3):
            for bx in boxes:
                bx.propagate(1.0)
            cur = [np.array([b.r for b in bx.beads]) for bx in boxes]
            d = [c - p for c, p in zip(cur, prev)]
            d = [x - Lb * np.round(x / Lb) for x in d]
            rel = np.abs(d[0] - d[1]).max() / np.abs(d[1]).max()
            print("step %d: displacement difference %.3e" % (step, rel))
            assert rel <= 1e-10
            assert boxes[0].net_flux == boxes[1].net_flux
            prev = cur
    finally:
        for bx in boxes:
            bx.close()
