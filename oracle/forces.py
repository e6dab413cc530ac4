"""CPU restatement of the reference's force field on the hydrodynamic step's
beads (test infrastructure only: tests/ and the oracle box use it to check
the device kernels; the library never calls it).

Follows, in the reference's own accumulation order:
  box.py:404-428        _compute_forces: F = external force, then each
                        interaction adds into F, then the energy-unit
                        conversion scales F (not E)
  box.py:449-461, 685-723  external force, optionally zeroed outside a box
                        region per axis
  interactions.py:483-589  interaction order: Lennard-Jones, harmonic bonds,
                        harmonic angles
  interactions.py:185-247, 249-289, 291-416  loops over mobile beads, then
                        immobile ones, with the mobile-immobile pointers
                        (rik, bead.c:99) and their negations and differences
                        exactly where the reference uses them
  interactions.py:641-821  force and energy formulas
  interactions.py:823-841  unit factors
  bead.py:272-278       angle_pbc
  bead.c:44-50          minimum image with the `while` semantics

Pinned by tests/test_forces_oracle.py against the known answers of the
reference's tests/test_interactions.py (with and without immobile beads),
against -grad E by finite differences, and against the invariant that making
a bead immobile leaves the forces on the other beads unchanged.
"""
import math

import numpy as np

KCAL_PER_MOL = 10 ** (-21) * 6.9477   # interactions.py:823-827
EV = 10 ** (-19) * 1.602176634        # interactions.py:835-839
UNIT = {"joule": 1.0, "kcal/mol": KCAL_PER_MOL, "eV": EV}


def min_image(d, L):
    """One component of bead.c:44-50 (values beyond 64 L are first reduced
    with fmod, as on the device, so a runaway position cannot stall the loop)."""
    if not math.isfinite(d):
        return d
    if abs(d) > 64.0 * L:
        d = math.fmod(d, L)
    while d >= L / 2:
        d -= L
    while d <= -L / 2:
        d += L
    return d


def pointer(ri, rj, L):
    """rij[i][j] = wrap(r_j - r_i) (bead.c:75)."""
    return np.array([min_image(rj[c] - ri[c], L) for c in range(3)])


def _dist2(p):
    return p[0] * p[0] + p[1] * p[1] + p[2] * p[2]


def lj(p, hc_i, hc_j, eps_i, eps_j, alpha, lj6, lj12):
    """Force on i and pair energy (interactions.py:711-821)."""
    sigma = hc_i + hc_j
    epsilon = math.sqrt(eps_i * eps_j)
    dist2 = _dist2(p)
    dist = math.sqrt(dist2)
    s2 = sigma * sigma / dist2
    versor = p / dist
    if lj6 and lj12:
        s6 = s2 * s2 * s2
        s12 = s6 * s6
        return alpha * epsilon / dist * (6 * s6 - 12 * s12) * versor, alpha * epsilon * (s12 - s6)
    if lj6:
        s6 = s2 * s2 * s2
        return 6.0 * alpha * epsilon * s6 / dist * versor, -alpha * epsilon * s6
    s12 = s2 * s2 * s2 * s2 * s2 * s2
    return -12.0 * alpha * epsilon * s12 / dist * versor, alpha * epsilon * s12


def bond(p, dist_eq, k):
    """interactions.py:641-665."""
    dist = math.sqrt(_dist2(p))
    return p / dist * k * (dist - dist_eq), 0.5 * k * (dist - dist_eq) ** 2


def angle_deg(p12, p23):
    """bead.py:272-278."""
    d12 = math.sqrt(_dist2(p12))
    d23 = math.sqrt(_dist2(p23))
    return np.rad2deg(np.arccos(-(p12[0] * p23[0] + p12[1] * p23[1] + p12[2] * p23[2]) / d12 / d23))


def angle(p12, p23, angle_eq, k):
    """interactions.py:667-709: 9 force components (beads i, j, k) and energy."""
    th = angle_deg(p12, p23) * (np.pi / 180.0)
    th_eq = angle_eq * (np.pi / 180.0)
    d2_12, d2_23 = _dist2(p12), _dist2(p23)
    d12, d23 = math.sqrt(d2_12), math.sqrt(d2_23)
    sc = np.zeros(9)
    sc[:3] = (p23 / d23 + np.cos(th) * p12 / d12) / d12
    sc[3:6] = (p12 - p23) / d12 / d23 - np.cos(th) * (p12 / d2_12 - p23 / d2_23)
    sc[6:] = -p12 / d12 / d23 - np.cos(th) * p23 / d2_23
    return k * (th - th_eq) / np.sin(th) * sc, 0.5 * k * (th - th_eq) ** 2


def external(xyz, Fex, region=None):
    """box.py:449-461 and 708-723: Fex on every bead, or only on beads inside
    region = {"x": [lo, hi], ...} (bounds inclusive)."""
    N = len(xyz)
    F = np.array(list(Fex) * N, dtype=np.float64)
    if region:
        for i in range(N):
            for c, ax in enumerate("xyz"):
                if ax in region and (xyz[i][c] < region[ax][0] or xyz[i][c] > region[ax][1]):
                    F[3 * i:3 * i + 3] = 0.0
                    break
    return F


def forces_and_energy(xyz, L, Fex=(0.0, 0.0, 0.0), region=None, lj6=False, lj12=False, alpha=4.0, hc=None,
                      eps=None, bonds=(), angles=(), unit="joule", mobile=None):
    """F (3 per mobile bead, in bead order) and E of box.py:404-428.

    xyz, hc, eps: every bead; mobile: one flag per bead (default: all mobile)
    bonds:  (i, j, dist_eq, k) in the reference's loop order (bead i lists j)
    angles: (i, j, k, angle_eq_degrees, force_constant), likewise
    """
    N = len(xyz)
    mob = [i for i in range(N) if mobile is None or mobile[i]]
    imm = [i for i in range(N) if not (mobile is None or mobile[i])]
    is_mob = [mobile is None or bool(mobile[i]) for i in range(N)]
    row = {b: n for n, b in enumerate(mob)}  # force row of a mobile bead

    def add(b, f):
        F[3 * row[b]:3 * row[b] + 3] += f

    def P(a, b):  # rij[a][b] / rik[a][b] = wrap(r_b - r_a)
        return pointer(xyz[a], xyz[b], L)

    F = external(np.asarray(xyz)[mob], Fex, region) if mob else np.zeros(0)
    E = 0.0
    if lj6 or lj12:
        # interactions.py:249-289: mobile pairs, then mobile-immobile pairs
        e = 0.0
        for ii in range(1, len(mob)):
            for jj in range(ii):
                i, j = mob[ii], mob[jj]
                f, ee = lj(P(i, j), hc[i], hc[j], eps[i], eps[j], alpha, lj6, lj12)
                add(i, f)
                add(j, -f)
                e += ee
        for i in mob:
            for j in imm:
                f, ee = lj(P(i, j), hc[i], hc[j], eps[i], eps[j], alpha, lj6, lj12)
                add(i, f)
                e += ee
        E += e
    if len(bonds):
        # interactions.py:185-247
        e = 0.0
        for i, j, deq, k in bonds:
            if is_mob[i]:
                f, ee = bond(P(i, j), deq, k)
                add(i, f)
                if is_mob[j]:
                    add(j, -f)
            elif is_mob[j]:
                f, ee = bond(-P(j, i), deq, k)
                add(j, -f)
            else:
                continue
            e += ee
        E += e
    if len(angles):
        # interactions.py:291-416
        e = 0.0
        for i, j, k, aeq, kc in angles:
            mj, mk = is_mob[j], is_mob[k]
            if is_mob[i]:
                p12 = P(i, j)
                if mk:
                    p23 = P(j, k) if mj else -P(k, j)
                else:
                    p23 = P(j, k) if mj else P(i, k) - p12
            elif not (mj or mk):
                continue
            elif mj:
                p12 = -P(j, i)
                p23 = P(j, k)
            else:
                p23 = -P(k, j)
                p12 = -P(k, i) - p23
            f, ee = angle(p12, p23, aeq, kc)
            if is_mob[i]:
                add(i, f[:3])
            if mj:
                add(j, f[3:6])
            if mk:
                add(k, f[6:])
            e += ee
        E += e
    if unit != "joule":
        F = UNIT[unit] * F
    return F, E


def bonded_lists(beads, mobile=None):
    """Bond and angle lists (indices into `beads`) in the reference's loop
    order: beads listing them, mobile ones first, then immobile ones
    (interactions.py:189-245, 295-416); an id resolves to the first mobile
    bead carrying it, else the first immobile one (bead.py:447-453 applied to
    the mobile list, then the immobile list)."""
    is_mob = [(mobile is None or bool(mobile[n])) for n in range(len(beads))]
    order = [n for n in range(len(beads)) if is_mob[n]] + [n for n in range(len(beads)) if not is_mob[n]]
    idx = {}
    for n in order:
        idx.setdefault(beads[n].bead_id, n)
    bonds, angles = [], []
    for i in order:
        b = beads[i]
        for other in b.bonded_with:
            deq, k = b.bonded_how[other]
            bonds.append((i, idx[other], deq, k))
    for i in order:
        b = beads[i]
        for id1, id2 in b.angled_with:
            aeq, k = b.angled_how[(id1, id2)]
            angles.append((i, idx[id1], idx[id2], aeq, k))
    return bonds, angles
