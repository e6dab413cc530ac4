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
