"""Synthetic bead configurations for the benchmark and parity workloads.

Beads are placed by random sequential addition (RSA) in the periodic cube
[-L/2, L/2)^3 with minimum-image rejection; L follows from the volume
fraction phi: L = (sum 4 pi a^3 / 3 / phi)^(1/3).  Placement uses a cell list
so N = 4000 at phi = 0.25 places in well under a second.

Named workloads (input keywords of BD.py:37-64):

  c1  rpy, N=100, a=5, L=150, dt=1 ps, flux plane x=0, seed 0
  c2  rpy_smith, N=1000, a=10, phi=0.05 (L=437.6), alpha=sqrt(pi), m=n=2
  c3  rpy_smith_lub midpoint, N=500 bidisperse 20/10 (1:1), phi=0.25,
      lubrication_cutoff 0.5, overlap_treshold 0.1, dt=0.1 ps
  c4  rpy_smith, N=4000, a=10, phi=0.10 (L=551.3), overlap checks off: the
      noise factor at scale is what this case measures, and with ~24 expected
      overlapping pairs per 1 ps move of all beads from an RSA start an
      overlap-free attempt has probability ~1e-10
"""
import math

import numpy as np

T_DEFAULT = 293.15
ETA_DEFAULT = 0.01005


def box_for(radii, phi):
    vol = sum(4.0 * math.pi * a ** 3 / 3.0 for a in radii)
    return (vol / phi) ** (1.0 / 3.0)


def rsa_positions(radii, L, seed, min_gap=0.0, max_tries=2000000):
    """Place beads in the given order; bead i rejects positions closer than
    a_i + a_j + min_gap (minimum image) to any placed bead j."""
    rng = np.random.RandomState(seed)
    radii = np.asarray(radii, dtype=np.float64)
    N = radii.size
    reach = 2 * radii.max() + min_gap
    nc = max(1, int(L // reach))
    cell = L / nc
    grid = {}
    pos = np.zeros((N, 3))
    tries = 0
    for i in range(N):
        while True:
            tries += 1
            if tries > max_tries:
                raise RuntimeError("RSA failed to place %d beads (phi too high?)" % N)
            p = rng.uniform(-L / 2, L / 2, 3)
            c = tuple(int((p[k] + L / 2) // cell) % nc for k in range(3))
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        key = ((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc)
                        for j in grid.get(key, ()):
                            d = pos[j] - p
                            d -= L * np.round(d / L)
                            if d @ d < (radii[i] + radii[j] + min_gap) ** 2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pos[i] = p
                grid.setdefault(c, []).append(i)
                break
    return pos


def c1_uniform(seed=0, N=100, a=5.0, L=150.0, min_dist=10.0):
    """Uniform placement with RandomState(seed), rejecting centre distances < min_dist."""
    rng = np.random.RandomState(seed)
    pos = []
    while len(pos) < N:
        p = rng.uniform(-L / 2, L / 2, 3)
        ok = True
        for q in pos:
            d = q - p
            d -= L * np.round(d / L)
            if d @ d < min_dist * min_dist:
                ok = False
                break
        if ok:
            pos.append(p)
    return np.array(pos), np.full(N, a), L


def workload(name, seed=0):
    """(positions (N,3), radii (N,), input keywords, dt, steps) for a named workload."""
    base = {"T": T_DEFAULT, "viscosity": ETA_DEFAULT, "seed": seed, "check_overlaps": True,
            "overlap_treshold": 0.0}
    if name == "c1":
        xyz, a, L = c1_uniform(seed)
        inp = dict(base, hydrodynamics="rpy", box_length=L,
                   measure_flux={"normal": [1.0, 0.0, 0.0], "plane_point": [0.0, 0.0, 0.0]})
        return xyz, a, inp, 1.0, 1000
    if name in ("c2", "c4"):
        N, phi = (1000, 0.05) if name == "c2" else (4000, 0.10)
        a = np.full(N, 10.0)
        L = box_for(a, phi)
        xyz = rsa_positions(a, L, seed)
        inp = dict(base, hydrodynamics="rpy_smith", box_length=L, ewald_alpha=math.sqrt(math.pi),
                   ewald_real=2, ewald_imag=2,
                   measure_flux={"normal": [1.0, 0.0, 0.0], "plane_point": [0.0, 0.0, 0.0]})
        if name == "c4":
            inp["check_overlaps"] = False
        return xyz, a, inp, 1.0, 500 if name == "c2" else 100
    if name == "c3":
        a = np.array([20.0] * 250 + [10.0] * 250)
        L = box_for(a, 0.25)
        xyz = rsa_positions(a, L, seed, min_gap=0.1)
        inp = dict(base, hydrodynamics="rpy_smith_lub", box_length=L, ewald_alpha=math.sqrt(math.pi),
                   ewald_real=2, ewald_imag=2, lubrication_cutoff=0.5, cichocki_correction=True,
                   propagation_scheme="midpoint", m_midpoint=100, overlap_treshold=0.1,
                   measure_flux={"normal": [1.0, 0.0, 0.0], "plane_point": [0.0, 0.0, 0.0]})
        return xyz, a, inp, 0.1, 200
    raise ValueError(name)
