"""The force field on the device (pbx_forces.hip) against the oracle's
restatement of the reference (oracle/forces.py): forces and energy at the
same positions, and trajectories with forces in the Ermak and midpoint
steps (positions within 1e-10 relative per step)."""
import math

import numpy as np
import pytest

from oracle import forces as of
from oracle.box_oracle import OracleBox

pytestmark = pytest.mark.gpu


def _chain(N, spacing, seed, L=1.0e6, min_sep=None):
    """A random walk with steps of `spacing` (a bead-spring chain), bonds
    between neighbours and angles over each three consecutive beads;
    non-neighbours at least `min_sep` apart (default 0.9 spacing)."""
    from pybrown_amd.bead import Bead
    rng = np.random.RandomState(seed)
    min_sep = 0.9 * spacing if min_sep is None else min_sep
    xyz = np.zeros((N, 3))
    for i in range(1, N):
        while True:
            v = rng.normal(size=3)
            p = xyz[i - 1] + spacing * v / np.linalg.norm(v)
            if i < 2 or np.linalg.norm(p - xyz[:i - 1], axis=1).min() > min_sep:
                break
        xyz[i] = p
    xyz -= xyz.mean(axis=0)
    beads = []
    for i in range(N):
        b = Bead(xyz[i].copy(), 1.0, hard_core_radius=rng.uniform(0.9, 1.1), epsilon_LJ=rng.uniform(0.5, 1.5),
                 bead_id=i + 1)
        beads.append(b)
    for i in range(N - 1):
        beads[i].bonded_with.append(i + 2)
        beads[i].bonded_how[i + 2] = [spacing * rng.uniform(0.9, 1.1), rng.uniform(0.5, 2.0)]
    for i in range(N - 2):
        beads[i].angled_with.append([i + 2, i + 3])
        beads[i].angled_how[(i + 2, i + 3)] = [rng.uniform(90.0, 150.0), rng.uniform(0.1, 1.0)]
    return beads, xyz


def _oracle_box(beads, xyz, inp):
    bonds, angles = of.bonded_lists(beads)
    return OracleBox(xyz, [b.a for b in beads], [b.label for b in beads], inp,
                     hard_core=[b.hard_core_radius for b in beads], epsilon_LJ=[b.epsilon_LJ for b in beads],
                     bonds=bonds, angles=angles)


@pytest.mark.parametrize("lj", [(True, False), (False, True), (True, True)], ids=["lj6", "lj12", "lj612"])
@pytest.mark.parametrize("unit", ["joule", "kcal/mol", "eV"])
def test_forces_and_energy_match_oracle(lj, unit):
    from pybrown_amd.box import Box
    beads, xyz = _chain(40, 2.4, 1)
    inp = {"box_length": 30.0, "T": 298.0, "viscosity": 0.01, "seed": 3, "hydrodynamics": "nohi",
           "lennard_jones_6": lj[0], "lennard_jones_12": lj[1], "lennard_jones_alpha": 3.0, "energy_unit": unit,
           "external_force": [0.3, -0.2, 0.1], "external_force_region": {"x": [-4.0, 6.0], "z": [-10.0, 3.0]},
           "check_overlaps": False}
    box = Box(beads, inp)
    try:
        box.propagate(1e-6)
        bonds, angles = of.bonded_lists(beads)
        Fo, Eo = of.forces_and_energy(xyz, 30.0, Fex=inp["external_force"], region=inp["external_force_region"],
                                      lj6=lj[0], lj12=lj[1], alpha=3.0, hc=[b.hard_core_radius for b in beads],
                                      eps=[b.epsilon_LJ for b in beads], bonds=bonds, angles=angles, unit=unit)
        Fg = box.F
        print("unit %s lj %s: |F| %.3e  max diff %.3e  E %.6e vs %.6e" % (unit, lj, np.abs(Fo).max(),
                                                                       np.abs(Fg - Fo).max(), box.E, Eo))
        assert np.abs(Fg - Fo).max() <= 1e-12 * np.abs(Fo).max()
        assert box.E == pytest.approx(Eo, rel=1e-12)
        # Newton's third law for the internal forces: only the external part remains
        ext = of.external(xyz, inp["external_force"], inp["external_force_region"]) * of.UNIT[unit]
        assert np.abs(Fg.reshape(-1, 3).sum(0) - ext.reshape(-1, 3).sum(0)).max() <= 1e-10 * np.abs(Fo).max()
    finally:
        box.close()


def _trajectory(beads, xyz, inp, steps, dt, tol=1e-10):
    from pybrown_amd.box import Box
    g = Box(beads, inp)
    o = _oracle_box(beads, xyz, inp)
    try:
        worst = 0.0
        for j in range(steps):
            g.propagate(dt)
            o.propagate(dt)
            if o.attempt_log:
                assert g.last_move_attempts == o.attempt_log[-1], j
            pos_g = np.array([b.r for b in g.beads])
            err = np.abs(pos_g - o.r).max() / np.abs(o.r).max()
            worst = max(worst, err)
            assert err <= tol, "step %d: position error %.3e" % (j, err)
            assert g.E == pytest.approx(o.E, rel=1e-9, abs=1e-12)
        return worst, g.E
    finally:
        g.close()


def test_ermak_chain_with_lj_bonds_angles():
    # parameters in kcal/mol (epsilon 0.5-1.5, bond constants 0.5-2 per A^2,
    # angle constants 0.1-1 per rad^2): drifts of a few hundredths of an
    # Angstrom per 0.05 ps step
    beads, xyz = _chain(30, 2.3, 2)
    inp = {"box_length": 1.0e6, "T": 298.0, "viscosity": 0.01, "seed": 5, "hydrodynamics": "rpy",
           "lennard_jones_6": True, "lennard_jones_12": True, "lennard_jones_alpha": 0.5,
           "energy_unit": "kcal/mol", "external_force": [0.0, 0.0, 0.0], "check_overlaps": False}
    worst, E = _trajectory(beads, xyz, inp, 20, 0.05)
    print("ermak chain: worst position error %.3e, final E %.6e" % (worst, E))
    assert abs(E) < 1e3  # the chain stays bound (kcal/mol)


def test_midpoint_with_region_force_and_bonds():
    beads, xyz = _chain(25, 2.3, 4)
    for b in beads:
        b.angled_with.clear()
    inp = {"box_length": 40.0, "T": 298.0, "viscosity": 0.01, "seed": 6, "hydrodynamics": "rpy_smith",
           "ewald_alpha": math.sqrt(math.pi), "ewald_real": 1, "ewald_imag": 1, "propagation_scheme": "midpoint",
           "m_midpoint": 100, "external_force": [2.0, 0.0, -1.0], "external_force_region": {"x": [-3.0, 5.0]},
           "energy_unit": "kcal/mol", "check_overlaps": False}
    worst, E = _trajectory(beads, xyz, inp, 10, 0.05)
    print("midpoint chain: worst position error %.3e, final E %.6e" % (worst, E))


def test_bonded_chain_with_overlap_checks():
    """Bonded neighbours closer than their radii' sum are exempt from the
    overlap test (bead.py:401-413, bead.c:180): the run starts and moves, with
    rejections of non-bonded contacts as in the oracle."""
    beads, xyz = _chain(24, 1.8, 7, min_sep=2.3)
    inp = {"box_length": 1.0e6, "T": 298.0, "viscosity": 0.01, "seed": 8, "hydrodynamics": "rpy",
           "energy_unit": "kcal/mol", "external_force": [0.0, 0.0, 0.0], "check_overlaps": True,
           "overlap_treshold": 0.3}
    d = np.linalg.norm(xyz[1:] - xyz[:-1], axis=1)
    assert d.max() < 2.0  # every bonded pair overlaps by the radii
    worst, E = _trajectory(beads, xyz, inp, 15, 0.05)
    print("bonded chain with overlap checks: worst position error %.3e, E %.4f" % (worst, E))
