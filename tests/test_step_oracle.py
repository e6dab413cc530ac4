"""Pins the numpy step restatement (oracle/box_oracle.py) with the reference
suite's own designs (src/bdsim/tests/test_box.py:215-511): a hand-written
restatement of one step, fed the same seeded stream, must reproduce the
engine's positions (bit-exact where the reference asserts equality)."""
import math

import numpy as np
import pytest

from oracle.box_oracle import BOLTZMANN, OracleBox
from oracle.orc import Oracle

BASE = {"box_length": 35.0, "T": 298.0, "viscosity": 0.01, "seed": 0}


def _grid(coords):
    return np.array([[x, y, z] for x in coords for y in coords for z in coords], float)


def test_random_stream_concatenates():
    # test_box.py:215-237: successive 3N draws equal one draw of the whole block
    rng = np.random.RandomState(0)
    parts = np.concatenate([rng.normal(0.0, 1.0, 9) for _ in range(7)])
    np.testing.assert_array_equal(parts, np.random.RandomState(0).normal(0.0, 1.0, 63))


def test_ermak_nohi_matches_restatement():
    # test_box.py:241-285
    xyz = _grid([-6.0, -3.0, 0.0, 3.0, 6.0])
    n = xyz.size
    box = OracleBox(xyz, np.ones(len(xyz)), ["X"] * len(xyz), dict(BASE, hydrodynamics="nohi"))
    dt, Nit = 1e-10, 50
    N = np.random.RandomState(0).normal(0.0, 1.0, Nit * n)
    r = xyz.copy()
    for it in range(Nit):
        box.propagate(dt)
        D = BOLTZMANN * 298.0 * 10**19 / (6.0 * np.pi * 0.01) * np.ones(n)
        BX = np.sqrt(D) * N[it * n:(it + 1) * n] * math.sqrt(2 * dt)
        r += BX.reshape(-1, 3)
        np.testing.assert_array_equal(box.xi, N[it * n:(it + 1) * n])
        np.testing.assert_array_equal(box.r, r)


def test_ermak_rpy_matches_restatement():
    # test_box.py:289-343 (bit-exact)
    xyz = _grid([-3.0, 0.0, 3.0])
    n = xyz.size
    orc = Oracle("orc")
    box = OracleBox(xyz, np.ones(len(xyz)), ["X"] * len(xyz), dict(BASE, hydrodynamics="rpy"))
    dt, Nit = 1e-10, 50
    N = np.random.RandomState(0).normal(0.0, 1.0, Nit * n)
    r = xyz.copy()
    for it in range(Nit):
        P = Oracle.packed(orc.pointer_pbc_matrix(r, 35.0))
        box.propagate(dt)
        D = BOLTZMANN * 298.0 * 10**19 / 0.01 * orc.rpy_m_matrix(np.ones(len(xyz)), P)
        BX = np.linalg.cholesky(D) @ N[it * n:(it + 1) * n] * math.sqrt(2 * dt)
        r += BX.reshape(-1, 3)
        np.testing.assert_allclose(box.r, r, rtol=0, atol=1e-12)


@pytest.mark.parametrize("mn,m", [(1, 2.0), (2, 10.0)])
def test_midpoint_rpy_smith_lub_matches_restatement(mn, m):
    # test_box.py:409-511 (places = 9)
    xyz = _grid([-2.2, 0.0, 2.2])
    n = xyz.size
    a = np.ones(len(xyz))
    orc = Oracle("orc")
    inp = dict(BASE, hydrodynamics="rpy_smith_lub", propagation_scheme="midpoint", lubrication_cutoff=10,
               ewald_alpha=math.sqrt(math.pi), ewald_real=mn, ewald_imag=mn, m_midpoint=m)
    box = OracleBox(xyz, a, ["X"] * len(xyz), inp)
    dt, Nit = 1e-10, 10
    N = np.random.RandomState(0).normal(0.0, 1.0, Nit * n)

    def Dtot(r):
        P = Oracle.packed(orc.pointer_pbc_matrix(r, 35.0))
        Mff = orc.rpy_smith_m_matrix(a, P, 35.0, math.sqrt(math.pi), mn, mn) * 10**19 / 0.01
        Rlc = orc.jo_r_lubrication_correction_matrix(a, P, 10, True) * 0.01 * 10**(-19)
        return BOLTZMANN * 298.0 * np.linalg.inv(Rlc + np.linalg.inv(Mff))

    r = xyz.copy()
    for it in range(Nit):
        box.propagate(dt)
        D0 = Dtot(r)
        B0 = np.linalg.cholesky(D0)
        Ni = N[it * n:(it + 1) * n]
        BX = B0 @ Ni * math.sqrt(2 * dt)
        r += BX.reshape(-1, 3) / m
        Dmid = Dtot(r)
        r += (BX * (1.0 - 1.0 / m)).reshape(-1, 3)
        BX_mid = Dmid @ np.linalg.inv(B0.T) @ Ni * math.sqrt(2 * dt)
        r += ((BX_mid - BX) * m / 2).reshape(-1, 3)
        np.testing.assert_allclose(box.r, r, rtol=0, atol=1e-9)


def test_flux_telescopes_and_counts_direction():
    # bead.py:103-132 sign convention, A1: net count = s(end) - s(start)
    xyz = np.array([[-0.5, 0.0, 0.0], [5.0, 5.0, 5.0]])
    inp = dict(BASE, hydrodynamics="nohi",
               measure_flux={"normal": [1.0, 0.0, 0.0], "plane_point": [0.0, 0.0, 0.0]})
    box = OracleBox(xyz, np.ones(2), ["A", "B"], inp)
    box._translate(np.array([1.0, 0, 0, 0, 0, 0]))
    box._translate(np.array([-2.0, 0, 0, 0, 0, 0]))
    box._translate(np.array([3.0, 0, 0, 0, 0, 0]))
    assert box.net_flux == {"A": 1, "B": 0}


def test_rejection_consumes_one_draw_per_attempt():
    # box.py:223-248: every attempt draws 3N; rejected moves are undone
    rng = np.random.RandomState(1)
    xyz = np.array([[0.0, 0, 0], [2.02, 0, 0], [10.0, 10, 10]])
    inp = dict(BASE, hydrodynamics="rpy", seed=4)
    box = OracleBox(xyz, np.ones(3), ["X"] * 3, inp)
    for _ in range(20):
        box.propagate(1e-2)
    assert box.draw_count == 9 * sum(box.attempt_log)
    assert max(box.attempt_log) > 1
    assert not box._check_overlaps()


def test_planted_mistake_in_oracle_is_detected():
    # dropping the sqrt(2 dt) factor must break the nohi restatement check
    xyz = _grid([-3.0, 0.0, 3.0])
    box = OracleBox(xyz, np.ones(len(xyz)), ["X"] * len(xyz), dict(BASE, hydrodynamics="nohi"))
    box.propagate(1e-10)
    N = np.random.RandomState(0).normal(0.0, 1.0, xyz.size)
    D = BOLTZMANN * 298.0 * 10**19 / (6.0 * np.pi * 0.01) * np.ones(xyz.size)
    wrong = xyz + (np.sqrt(D) * N).reshape(-1, 3)
    assert not np.allclose(box.r, wrong)
