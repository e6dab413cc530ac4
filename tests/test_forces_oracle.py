"""The oracle's force field (oracle/forces.py) against the known answers of
the reference's tests/test_interactions.py and against -grad E by central
differences, so a wrong sign, factor or pairing fails here on the CPU."""
import math

import numpy as np
import pytest

from oracle import forces as of

BIG = 1.0e6  # box length of the reference's free-space cases
ALPHA, EPS, D = 4.0, 1.0, 2.0


def _fe(xyz, **kw):
    return of.forces_and_energy(np.array(xyz, dtype=float), BIG, **kw)


@pytest.mark.parametrize("shift", [0.0, 1.0, -1.0])
def test_harmonic_dumbbell(shift):
    # test_interactions.py:54-76, 183-205, 312-334: beads at z = 0 and z = d
    k, deq = 1.0, D + shift
    F, E = _fe([[0, 0, 0], [0, 0, D]], bonds=[(0, 1, deq, k)])
    assert E == 0.5 * k * (deq - D) ** 2
    assert list(F) == [0.0, 0.0, k * (D - deq), 0.0, 0.0, -k * (D - deq)]


def test_two_dumbbells():
    # test_interactions.py:441-480
    k1, deq1, k2, deq2 = 1.0, 3.0, 2.0, 1.0
    xyz = [[0, 0, 0], [0, 0, D], [0, 0, 10.0], [0, 0, 10.0 + D], [5, 5, 5]]
    F, E = _fe(xyz, bonds=[(0, 1, deq1, k1), (2, 3, deq2, k2)])
    assert E == 0.5 * k1 * (deq1 - D) ** 2 + 0.5 * k2 * (deq2 - D) ** 2
    want = [0, 0, -k1 * (deq1 - D), 0, 0, k1 * (deq1 - D), 0, 0, k2 * (D - deq2), 0, 0, -k2 * (D - deq2), 0, 0, 0]
    assert list(F) == want


def test_angle_at_equilibrium():
    # test_interactions.py:581-604: a right angle
    F, E = _fe([[0, 0, 0], [0.5, 0, 0.5], [1.0, 0, 0]], angles=[(0, 1, 2, 90.0, 1.0)])
    assert E == 0.0 and np.all(F == 0.0)


def test_angle_off_equilibrium():
    # test_interactions.py:865-895: energy, zero net force, forces on the end
    # beads perpendicular to their arms
    xyz = np.array([[0, 0, 0], [0.5, 0, 0.5], [1.0, 0, 0]])
    for aeq in np.linspace(0.0, 360.0, 100):
        F, E = _fe(xyz, angles=[(0, 1, 2, aeq, 1.0)])
        assert E == pytest.approx(0.5 * (aeq - 90.0) ** 2 * (np.pi / 180.0) ** 2, abs=1e-10)
        assert np.allclose(F[:3] + F[3:6] + F[6:], 0.0, atol=1e-10)
        p12, p23 = xyz[1] - xyz[0], xyz[2] - xyz[1]
        assert abs(F[:3] @ p12) < 1e-10 and abs(F[6:] @ p23) < 1e-10


@pytest.mark.parametrize("lj6,lj12,E_ref,f_ref", [
    (True, False, -ALPHA * EPS, 6.0 * ALPHA * EPS / D),      # test_interactions.py:1253-1267
    (False, True, ALPHA * EPS, -12.0 * ALPHA * EPS / D),     # :1369-1383
    (True, True, 0.0, -6.0 * ALPHA * EPS / D),               # :1485-1499
])
def test_lj_pair_at_sigma(lj6, lj12, E_ref, f_ref):
    F, E = _fe([[0, 0, 0], [0, 0, D]], lj6=lj6, lj12=lj12, alpha=ALPHA, hc=[1.0, 1.0], eps=[EPS, EPS])
    assert E == E_ref
    assert list(F) == [0.0, 0.0, f_ref, 0.0, 0.0, -f_ref]


def test_lj_pair_at_minimum():
    # test_interactions.py:1635-1657
    F, E = _fe([[0, 0, 0], [2 * 2.0 ** (1 / 6), 0, 0]], lj6=True, lj12=True, alpha=ALPHA, hc=[1.0, 1.0],
               eps=[EPS, EPS])
    assert E == -EPS
    assert np.allclose(F, 0.0, atol=1e-7)


@pytest.mark.parametrize("scale,E_ref", [(2 * 2.0 ** (1 / 6), -3 * EPS), (2.0, 0.0)])
def test_lj_equilateral_triangles(scale, E_ref):
    # test_interactions.py:1753-1779 (stable, F = 0) and 2024-2048 (unstable, E = 0)
    A = scale / (np.cos(5 * np.pi / 6) - np.cos(np.pi / 6))
    xyz = [[A * np.cos(t), A * np.sin(t), 0.0] for t in (np.pi / 6, 5 * np.pi / 6, 3 * np.pi / 2)]
    F, E = _fe(xyz, lj6=True, lj12=True, alpha=ALPHA, hc=[1.0] * 3, eps=[EPS] * 3)
    assert E == pytest.approx(E_ref, abs=1e-7)
    if E_ref != 0.0:
        assert np.allclose(F, 0.0, atol=1e-7)


def _random_system(seed, N=7, L=40.0):
    rng = np.random.RandomState(seed)
    xyz = rng.uniform(-L / 2, L / 2, (N, 3))
    hc = rng.uniform(0.8, 1.5, N)
    eps = rng.uniform(0.5, 2.0, N)
    bonds = [(0, 1, 3.0, 0.7), (2, 1, 2.5, 1.3), (4, 3, 5.0, 0.2)]
    angles = [(0, 1, 2, 100.0, 0.9), (5, 3, 4, 70.0, 1.7)]
    return xyz, L, hc, eps, bonds, angles


@pytest.mark.parametrize("kind", ["lj6", "lj12", "lj612", "bonds", "angles"])
def test_force_is_minus_gradient_of_energy(kind):
    xyz, L, hc, eps, bonds, angles = _random_system(3)
    kw = dict(lj6=kind in ("lj6", "lj612"), lj12=kind in ("lj12", "lj612"), alpha=ALPHA, hc=hc, eps=eps,
              bonds=bonds if kind == "bonds" else (), angles=angles if kind == "angles" else ())
    F, _ = of.forces_and_energy(xyz, L, **kw)
    h = 1e-6
    G = np.zeros_like(F)
    for q in range(F.size):
        xp, xm = xyz.copy(), xyz.copy()
        xp.flat[q] += h
        xm.flat[q] -= h
        G[q] = -(of.forces_and_energy(xp, L, **kw)[1] - of.forces_and_energy(xm, L, **kw)[1]) / (2 * h)
    assert np.allclose(F, G, rtol=1e-5, atol=1e-7 * np.abs(F).max())


def test_periodic_pointer_and_unit_and_region():
    # bond across the periodic boundary: the minimum image is used
    L = 10.0
    F, E = of.forces_and_energy(np.array([[4.5, 0, 0], [-4.5, 0, 0]]), L, bonds=[(0, 1, 2.0, 1.0)])
    assert E == pytest.approx(0.5, abs=1e-12) and F[0] == pytest.approx(-1.0)
    # energy units scale F, not E (box.py:414-428)
    F1, E1 = of.forces_and_energy(np.array([[0, 0, 0], [0, 0, D]]), BIG, Fex=(1.0, 0, 0), bonds=[(0, 1, 3.0, 1.0)],
                                  unit="kcal/mol")
    F0, E0 = of.forces_and_energy(np.array([[0, 0, 0], [0, 0, D]]), BIG, Fex=(1.0, 0, 0), bonds=[(0, 1, 3.0, 1.0)])
    assert E1 == E0 and np.array_equal(F1, of.KCAL_PER_MOL * F0)
    # the external force acts only inside the region (bounds inclusive)
    xyz = np.array([[0.0, 0, 0], [2.0, 0, 0], [5.0, 0, 0]])
    F, _ = of.forces_and_energy(xyz, BIG, Fex=(1.0, 2.0, 3.0), region={"x": [0.0, 2.0]})
    assert list(F) == [1.0, 2.0, 3.0, 1.0, 2.0, 3.0, 0.0, 0.0, 0.0]
