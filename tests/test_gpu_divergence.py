"""The divergence term (box.py:211-215, 560-610) on the device against the
oracle: trajectories of lubricated clusters with divergence_term on.  The
forward difference divides by eps = 1e-7, so matrix differences of ~1e-13
between the two paths become ~1e-9 in the displacement: positions are
compared at 1e-8 relative (the term itself moves the beads by far more)."""
import numpy as np
import pytest

from oracle.box_oracle import OracleBox

pytestmark = pytest.mark.gpu

# a cluster of unit spheres with gaps of 0.14-0.6 (the JO series needs
# separated spheres; an overlapping pair turns R into NaN)
XYZ = np.array([[0.0, 0.0, 0.0], [-0.955, -1.779, 1.347], [0.518, -2.31, -0.105], [-0.511, 1.809, 1.133],
                [0.531, 0.272, 2.335], [1.211, -2.273, -2.492], [1.22, 2.251, -0.518]])


@pytest.mark.parametrize("hydro,L", [("rpy_lub", 1.0e6), ("rpy_smith_lub", 12.0)])
def test_divergence_trajectory(hydro, L):
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    inp = {"box_length": L, "T": 298.0, "viscosity": 0.01, "seed": 4, "hydrodynamics": hydro,
           "ewald_alpha": np.sqrt(np.pi), "ewald_real": 1, "ewald_imag": 1, "lubrication_cutoff": 1.0,
           "cichocki_correction": True, "divergence_term": True, "check_overlaps": False}
    a = np.ones(len(XYZ))
    g = Box([Bead(p.copy(), 1.0) for p in XYZ], inp)
    o = OracleBox(XYZ.copy(), a, ["XXX"] * len(a), inp)
    off = OracleBox(XYZ.copy(), a, ["XXX"] * len(a), dict(inp, divergence_term=False))
    try:
        worst = 0.0
        for j in range(4):
            flags = (True, True, True) if j != 2 else (False, True, True)
            g.propagate(0.01, *flags)
            o.propagate(0.01, *flags)
            off.propagate(0.01, *flags)
            pos = np.array([b.r for b in g.beads])
            err = np.abs(pos - o.r).max() / np.abs(o.r).max()
            worst = max(worst, err)
            assert err <= 1e-8, "step %d: %.3e" % (j, err)
        effect = np.abs(o.r - off.r).max() / np.abs(o.r).max()
        print("%s: worst position error %.3e; divergence term moves beads by %.3e relative" % (hydro, worst,
                                                                                            effect))
        assert effect > 1e3 * max(worst, 1e-15)
    finally:
        g.close()
