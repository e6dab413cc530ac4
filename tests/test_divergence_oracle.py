"""The oracle's divergence term (box.py:560-610 restated in
oracle/box_oracle.py) against an independent central-difference divergence
of the same diffusion matrix, and its translation invariance in free space."""
import numpy as np
import pytest

from oracle.box_oracle import OracleBox

# separated unit spheres (closest gap 0.15)
XYZ = np.array([[0.0, 0.0, 0.0], [2.15, 0.3, 0.0], [0.4, 2.2, 0.5], [-1.0, -1.9, 1.3]])


def _box(xyz, hydro="rpy_lub"):
    inp = {"box_length": 1.0e6, "T": 298.0, "viscosity": 0.01, "seed": 1, "hydrodynamics": hydro,
           "lubrication_cutoff": 1.0, "cichocki_correction": True, "divergence_term": True,
           "check_overlaps": False}
    b = OracleBox(xyz, np.ones(len(xyz)), ["X"] * len(xyz), inp)
    b._compute_Dff_matrix()
    b._compute_Dtot_matrix()
    return b


def _D_at(b, xyz):
    b.r = np.array(xyz, dtype=float)
    b._compute_Dff_matrix()
    b._compute_Dtot_matrix()
    return np.copy(b.D)


def test_divergence_matches_central_differences():
    b = _box(XYZ.copy())
    div = b._compute_divergence(True, True)
    n = 3 * len(XYZ)
    h = 1e-4
    ref = np.zeros(n)
    probe = _box(XYZ.copy())
    for k in range(n):
        xp, xm = XYZ.copy(), XYZ.copy()
        xp.flat[k] += h
        xm.flat[k] -= h
        ref += (_D_at(probe, xp)[k] - _D_at(probe, xm)[k]) / (2 * h)
    print("divergence", div, "central", ref)
    assert np.abs(ref).max() > 1e-3          # lubrication makes it non-zero here
    assert np.allclose(div, ref, rtol=2e-3, atol=2e-3 * np.abs(ref).max())
    # positions are restored (up to the (r + eps) - eps rounding)
    assert np.allclose(b.r, XYZ, atol=1e-12)


def test_divergence_is_translation_invariant():
    d0 = _box(XYZ.copy())._compute_divergence(True, True)
    d1 = _box(XYZ + np.array([3.0, -7.0, 11.0]))._compute_divergence(True, True)
    assert np.allclose(d0, d1, rtol=1e-5, atol=1e-5 * np.abs(d0).max())
