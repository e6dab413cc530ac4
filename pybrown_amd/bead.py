"""Beads and the geometry entry points of pyBrown/bead.py, on the GPU.

``Bead`` keeps the reference constructor (pyBrown/bead.py:81-90) and the
translation / flux / wrap methods the step uses (:92-183).
``compute_pointer_pbc_matrix`` and ``check_overlaps`` keep the reference
signatures (:282-303, :417-443) and call the HIP implementations of
clib/bead.c:27-95 and :169-260 through libpybrown_amd.so.
"""
import numpy as np

from pybrown_amd import _lib
from pybrown_amd._lib import lib, ptr, iptr


class Bead:
    """Spherical bead (pyBrown/bead.py:29-90)."""

    def __init__(self, coords, hydrodynamic_radius, label="XXX", hard_core_radius=None, epsilon_LJ=0.0,
                 mobile=True, bead_id=None):
        self._r = np.array(coords, dtype=np.float64)
        self.a = hydrodynamic_radius
        self.label = label
        self.mobile = mobile
        self.hard_core_radius = hard_core_radius if hard_core_radius is not None else self.a
        self.epsilon_LJ = epsilon_LJ
        self.bead_id = bead_id
        self.bonded_with = []
        self.bonded_how = {}
        self.angled_with = []
        self.angled_how = {}

    @property
    def r(self):
        """Position (3,).  Inside a Box this is a view of the box's position
        array; assigning a new array copies it into that view."""
        return self._r

    @r.setter
    def r(self, value):
        if getattr(self, "_bound", False):
            self._r[:] = value
        else:
            self._r = np.array(value, dtype=np.float64)

    def _bind(self, view):
        self._r = view
        self._bound = True

    def translate(self, vector):
        """bead.py:92-99"""
        if self.mobile:
            self.r += vector

    def translate_and_return_flux(self, vector, plane):
        """bead.py:103-132: +1 crossing along the normal, -1 against, else 0."""
        f0 = np.dot(plane.normal_vector, (self.r - plane.plane_point)) > 0.0
        self.translate(vector)
        f1 = np.dot(plane.normal_vector, (self.r - plane.plane_point)) > 0.0
        if f0 == f1:
            return 0
        return -1 if f0 else 1

    def keep_in_box(self, box_length):
        """bead.py:170-183: wrap into [-L/2, L/2)."""
        for i in range(3):
            while self.r[i] < -box_length / 2:
                self.r[i] += box_length
            while self.r[i] >= box_length / 2:
                self.r[i] -= box_length

    def is_there_bond_between(self, bead):
        return bead.bead_id in self.bonded_with or self.bead_id in bead.bonded_with

    def __str__(self):
        return "{}: {}, radius = {}, {}".format(self.label, self.r, self.a, "mobile" if self.mobile else "immobile")


def _xyz(beads):
    return np.ascontiguousarray(np.array([b.r for b in beads], dtype=np.float64).reshape(-1))


def compute_pointer_pbc_matrix(beads, box_length):
    """bead.py:282-303: (N, N, 3) with rij[a][b] = wrap(r_b - r_a)."""
    N = len(beads)
    xyz = _xyz(beads)
    P = np.zeros(3 * N * N)
    lib.pointer_pbc_matrix(ptr(xyz), N, float(box_length), ptr(P))
    _lib.check_void("pointer_pbc_matrix")
    return P.reshape(N, N, 3)


def compute_pointer_immobile_pbc_matrix(mobile_beads, immobile_beads, box_length):
    """bead.py:307-335: (N_mob, N_immob, 3) with rik[i][j] = wrap(r_j - r_i)."""
    nm, ni = len(mobile_beads), len(immobile_beads)
    xm, xi = _xyz(mobile_beads), _xyz(immobile_beads)
    P = np.zeros(3 * nm * ni)
    if nm and ni:
        lib.pointer_immobile_pbc_matrix(ptr(xm), ptr(xi), nm, ni, float(box_length), ptr(P))
        _lib.check_void("pointer_immobile_pbc_matrix")
    return P.reshape(nm, ni, 3)


def build_connection_matrix(beads):
    """bead.py:401-413 (bonds and immobile pairs are never tested)."""
    N = len(beads)
    cm = np.zeros((N, N), dtype=np.int32)
    for i in range(N):
        for j in range(N):
            if beads[i].is_there_bond_between(beads[j]):
                cm[i][j] = cm[j][i] = 1
            if not (beads[i].mobile or beads[j].mobile):
                cm[i][j] = cm[j][i] = 1
    return cm


def check_overlaps(beads, box_length, overlap_treshold, connection_matrix):
    """bead.py:417-443: True if any non-bonded pair overlaps."""
    N = len(beads)
    xyz = _xyz(beads)
    a = np.ascontiguousarray([b.a for b in beads], dtype=np.float64)
    cm = np.ascontiguousarray(connection_matrix, dtype=np.int32).reshape(-1)
    r = lib.check_overlaps(ptr(xyz), ptr(a), N, float(box_length), float(overlap_treshold), iptr(cm))
    if r < 0:
        raise RuntimeError("check_overlaps: %s" % _lib.last_error())
    return r == 1


def bonded_lists(beads):
    """Harmonic bonds (i, j, dist_eq, k) and angles (i, j, k, angle_eq, k_angle)
    as bead indices, in the order of the reference's loops
    (interactions.py:189-205, 295-316); bead ids resolve to the first bead
    carrying them (bead.py:447-453).  Given mobile beads followed by immobile
    ones, this is also the reference's order with immobile beads (their
    loops come after the mobile ones, :232-245, :349-416) and its id lookup
    (mobile beads first)."""
    first = {}
    for n, b in enumerate(beads):
        first.setdefault(b.bead_id, n)
    bonds, angles = [], []
    for i, b in enumerate(beads):
        for other in b.bonded_with:
            deq, k = b.bonded_how[other]
            bonds.append((i, first[other], float(deq), float(k)))
    for i, b in enumerate(beads):
        for id1, id2 in b.angled_with:
            aeq, k = b.angled_how[(id1, id2)]
            angles.append((i, first[id1], first[id2], float(aeq), float(k)))
    return bonds, angles
