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


# --- immobile beads (interactions.py:185-416 with mobile_beads / immobile_beads)

def _fe_mob(xyz, mobile, **kw):
    return of.forces_and_energy(np.array(xyz, dtype=float), BIG, mobile=mobile, **kw)


@pytest.mark.parametrize("deq_shift,mobile,E_ref,F_ref", [
    # test_interactions.py:80-205: b1 (at z = 0) lists the bond to b2 (z = d)
    (0.0, [False, True], 0.0, [0.0, 0.0, 0.0]),
    (0.0, [True, False], 0.0, [0.0, 0.0, 0.0]),
    (0.0, [False, False], 0.0, []),
    (-1.0, [False, True], 0.5, [0.0, 0.0, -1.0]),   # :209-241
    (-1.0, [True, False], 0.5, [0.0, 0.0, 1.0]),    # :243-275
    (-1.0, [False, False], 0.0, []),                # :277-309
    (1.0, [False, True], 0.5, [0.0, 0.0, 1.0]),     # :338-370
    (1.0, [True, False], 0.5, [0.0, 0.0, -1.0]),    # :372-404
    (1.0, [False, False], 0.0, []),                 # :406-438
])
def test_harmonic_dumbbell_with_immobile_beads(deq_shift, mobile, E_ref, F_ref):
    F, E = _fe_mob([[0, 0, 0], [0, 0, D]], mobile, bonds=[(0, 1, D + deq_shift, 1.0)])
    assert E == E_ref
    assert list(F) == F_ref


def test_two_dumbbells_with_immobile_beads():
    # test_interactions.py:484-530 (b1, b4 immobile) and 532-578 (b1, b3, b5)
    k1, deq1, k2, deq2 = 1.0, D + 1.0, 2.0, D - 1.0
    xyz = [[0, 0, 0], [0, 0, D], [0, 2 * D, D / 2], [0, 4 * D, D / 2], [0, 5 * D, D / 2]]
    bonds = [(0, 1, deq1, k1), (4, 3, deq2, k2)]  # b1 lists b2, b5 lists b4
    E_ref = 0.5 * k1 * (deq1 - D) ** 2 + 0.5 * k2 * (deq2 - D) ** 2
    F, E = _fe_mob(xyz, [False, True, True, False, True], bonds=bonds)
    assert E == E_ref
    assert list(F) == [0.0, 0.0, k1 * (deq1 - D), 0.0, 0.0, 0.0, 0.0, -k2 * (D - deq2), 0.0]
    F, E = _fe_mob(xyz, [False, True, False, True, False], bonds=bonds)
    assert E == E_ref
    assert list(F) == [0.0, 0.0, k1 * (deq1 - D), 0.0, k2 * (D - deq2), 0.0]


@pytest.mark.parametrize("lj6,lj12,mobile,E_ref,f_ref", [
    # test_interactions.py:1271-1355, 1377-1459, 1497-1613 (LJ pair at sigma)
    (True, False, [False, True], -ALPHA * EPS, -6.0 * ALPHA * EPS / D),
    (True, False, [True, False], -ALPHA * EPS, 6.0 * ALPHA * EPS / D),
    (True, False, [False, False], 0.0, None),
    (False, True, [False, True], ALPHA * EPS, 12.0 * ALPHA * EPS / D),
    (False, True, [True, False], ALPHA * EPS, -12.0 * ALPHA * EPS / D),
    (False, True, [False, False], 0.0, None),
    (True, True, [False, True], 0.0, 6.0 * ALPHA * EPS / D),
    (True, True, [True, False], 0.0, -6.0 * ALPHA * EPS / D),
    (True, True, [False, False], 0.0, None),
])
def test_lj_pair_at_sigma_with_immobile_beads(lj6, lj12, mobile, E_ref, f_ref):
    F, E = _fe_mob([[0, 0, 0], [0, 0, D]], mobile, lj6=lj6, lj12=lj12, alpha=ALPHA, hc=[1.0, 1.0], eps=[EPS, EPS])
    assert E == E_ref
    assert list(F) == ([] if f_ref is None else [0.0, 0.0, f_ref])


@pytest.mark.parametrize("mobile,E_ref", [
    # test_interactions.py:1661-1751 (pair at the minimum) and 1783-2024
    # (stable triangle): pairs of two immobile beads do not count
    ([False, True, True], -3.0), ([True, False, True], -3.0), ([True, True, False], -3.0),
    ([False, False, True], -2.0), ([False, True, False], -2.0), ([True, False, False], -2.0),
    ([False, False, False], 0.0),
])
def test_lj_triangle_with_immobile_beads(mobile, E_ref):
    A = 2 * 2.0 ** (1 / 6) / (np.cos(5 * np.pi / 6) - np.cos(np.pi / 6))
    xyz = [[A * np.cos(t), A * np.sin(t), 0.0] for t in (np.pi / 6, 5 * np.pi / 6, 3 * np.pi / 2)]
    F, E = _fe_mob(xyz, mobile, lj6=True, lj12=True, alpha=ALPHA, hc=[1.0] * 3, eps=[EPS] * 3)
    assert E == E_ref
    assert F.size == 3 * sum(mobile) and np.allclose(F, 0.0, atol=1e-7)
    F, E = _fe_mob([[0, 0, 0], [2 * 2.0 ** (1 / 6), 0, 0]], mobile[:2], lj6=True, lj12=True, alpha=ALPHA,
                   hc=[1.0, 1.0], eps=[EPS, EPS])
    assert E == (-EPS if any(mobile[:2]) else 0.0) and np.allclose(F, 0.0, atol=1e-7)


def test_lj_with_dummy_immobile_beads():
    # test_interactions.py:2211-2247: epsilon 0 immobile beads change nothing
    A = 2 * 2.0 ** (1 / 6) / (np.cos(5 * np.pi / 6) - np.cos(np.pi / 6))
    xyz = [[A * np.cos(t), A * np.sin(t), 0.0] for t in (np.pi / 6, 5 * np.pi / 6, 3 * np.pi / 2)]
    xyz += [[5.0, 5.0, 0.0], [3.0, 3.0, 0.0], [0.0, 0.0, 0.0]]
    F, E = _fe_mob(xyz, [True] * 3 + [False] * 3, lj6=True, lj12=True, alpha=ALPHA, hc=[1.0] * 6,
                   eps=[EPS] * 3 + [0.0] * 3)
    assert E == -3 * EPS
    assert F.size == 9 and np.allclose(F, 0.0, atol=1e-7)


def _chain_system(seed, N=9, L=30.0):
    """Beads on a random walk (bond length ~2.5) with bonds and angles along
    it in both listing directions, plus LJ parameters."""
    rng = np.random.RandomState(seed)
    xyz = np.cumsum(rng.normal(size=(N, 3)) * 1.5, axis=0) - 5.0
    hc = rng.uniform(0.8, 1.2, N)
    eps = rng.uniform(0.5, 2.0, N)
    bonds = [(i, i + 1, 2.5, 0.7 + 0.1 * i) if i % 2 else (i + 1, i, 2.2, 1.1) for i in range(N - 1)]
    angles = [(i, i + 1, i + 2, 100.0 + 5 * i, 0.9) if i % 3 else (i + 2, i + 1, i, 80.0, 1.3) for i in range(N - 2)]
    return xyz, L, hc, eps, bonds, angles


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_immobile_beads_leave_the_other_forces_unchanged(seed):
    """Freezing beads drops their own force rows and the energy terms among
    frozen beads only; the forces on the mobile beads stay those of the
    all-mobile system (to rounding: some pair vectors are negated pointers or
    differences of two, as in the reference)."""
    xyz, L, hc, eps, bonds, angles = _chain_system(seed)
    N = len(xyz)
    rng = np.random.RandomState(100 + seed)
    mobile = rng.rand(N) < 0.5
    mobile[rng.randint(N)] = True
    kw = dict(lj6=True, lj12=True, alpha=ALPHA, hc=hc, eps=eps, bonds=bonds, angles=angles, Fex=(0.3, -0.2, 0.1))
    F_all, E_all = of.forces_and_energy(xyz, L, **kw)
    F_mix, E_mix = of.forces_and_energy(xyz, L, mobile=mobile, **kw)
    rows = np.repeat(np.nonzero(mobile)[0] * 3, 3) + np.tile([0, 1, 2], int(mobile.sum()))
    assert np.allclose(F_mix, F_all[rows], rtol=1e-12, atol=1e-12 * np.abs(F_all).max())
    # energy of the terms among immobile beads only
    imm = np.nonzero(~mobile)[0]
    pos = {b: n for n, b in enumerate(imm)}
    sub_b = [(pos[i], pos[j], d, k) for i, j, d, k in bonds if not mobile[i] and not mobile[j]]
    sub_a = [(pos[i], pos[j], pos[k], t, c) for i, j, k, t, c in angles
             if not (mobile[i] or mobile[j] or mobile[k])]
    E_imm = of.forces_and_energy(xyz[imm], L, lj6=True, lj12=True, alpha=ALPHA, hc=hc[imm], eps=eps[imm],
                                 bonds=sub_b, angles=sub_a)[1] if len(imm) else 0.0
    assert E_mix == pytest.approx(E_all - E_imm, rel=1e-12, abs=1e-12 * abs(E_all))


@pytest.mark.parametrize("seed", [5, 6])
def test_immobile_force_is_minus_gradient_over_mobile_coordinates(seed):
    xyz, L, hc, eps, bonds, angles = _chain_system(seed)
    mobile = np.array([i % 3 != 1 for i in range(len(xyz))])
    kw = dict(lj6=True, lj12=True, alpha=ALPHA, hc=hc, eps=eps, bonds=bonds, angles=angles, mobile=mobile)
    F, _ = of.forces_and_energy(xyz, L, **kw)
    h = 1e-6
    G = []
    for b in np.nonzero(mobile)[0]:
        for c in range(3):
            xp, xm = xyz.copy(), xyz.copy()
            xp[b, c] += h
            xm[b, c] -= h
            G.append(-(of.forces_and_energy(xp, L, **kw)[1] - of.forces_and_energy(xm, L, **kw)[1]) / (2 * h))
    assert np.allclose(F, G, rtol=1e-5, atol=1e-7 * np.abs(F).max())


def test_bonded_lists_follow_the_reference_loop_order():
    """Mobile beads list first, then immobile ones (interactions.py:189-245);
    ids resolve to mobile beads first (bead.py:447-453 on the mobile list)."""
    from types import SimpleNamespace as NS
    beads = [NS(bead_id=1, bonded_with=[2], bonded_how={2: (1.0, 1.0)}, angled_with=[(2, 3)],
                angled_how={(2, 3): (90.0, 1.0)}),
             NS(bead_id=2, bonded_with=[3], bonded_how={3: (2.0, 1.0)}, angled_with=[], angled_how={}),
             NS(bead_id=3, bonded_with=[], bonded_how={}, angled_with=[(1, 2)], angled_how={(1, 2): (80.0, 2.0)}),
             NS(bead_id=3, bonded_with=[], bonded_how={}, angled_with=[], angled_how={})]
    bonds, angles = of.bonded_lists(beads, mobile=[False, True, False, True])
    assert bonds == [(1, 3, 2.0, 1.0), (0, 1, 1.0, 1.0)]
    assert angles == [(0, 1, 3, 90.0, 1.0), (2, 0, 1, 80.0, 2.0)]
