"""numpy restatement of the reference's per-step engine (pyBrown/box.py).

TEST INFRASTRUCTURE ONLY: the checker for pybrown_amd.Box.  Imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg only.

Covers the hydrodynamic step: hydrodynamics nohi / rpy / rpy_smith / rpy_lub /
rpy_smith_lub, Ermak-McCammon and midpoint propagation, overlap rejection with
ξ redraw, signed plane-crossing flux and the periodic wrap.  Arithmetic keeps
the reference's operation order (box.py:353-375, :486-556) so that, with the
restatement library and numpy's LAPACK, trajectories agree with the reference
to rounding.

Flux: the reference passes ``self.planes`` (a list that exists only with
walls) to ``Bead.translate_and_return_flux`` (box.py:383 against
bead.py:103-132), so any measure_flux run raises on its first step.  The
intended definition is implemented: a Plane(plane_point, normal) from the
measure_flux keywords (box.py:755-756) and, per bead and per translation,
[n.(r1-p) > 0] - [n.(r0-p) > 0] summed per label over one propagate().
"""
import math

import numpy as np

from oracle import forces
from oracle.orc import Oracle

BOLTZMANN = 1.380649e-23  # scipy.constants.Boltzmann (exact SI value)

# BD.py:42-56 defaults (the subset that reaches the step)
DEFAULTS = {
    "debug": False, "verbose": False, "hydrodynamics": "nohi",
    "ewald_alpha": math.sqrt(math.pi), "ewald_real": 0, "ewald_imag": 0,
    "diff_freq": 1, "lub_freq": 1, "chol_freq": 1,
    "lubrication_cutoff": 1, "seed": None, "immobile_labels": [],
    "propagation_scheme": "ermak", "m_midpoint": 100,
    "divergence_term": False, "check_overlaps": True,
    "external_force": [0.0, 0.0, 0.0], "cichocki_correction": True,
    "overlap_treshold": 0.0, "max_move_attempts": 1000000,
    "lennard_jones_6": False, "lennard_jones_12": False, "lennard_jones_alpha": 4.0, "energy_unit": "joule",
}


class OracleBox:
    """CPU twin of pyBrown.box.Box restricted to the hydrodynamic step.

    xyz: (Nt, 3) positions of every bead; radii: (Nt,); labels: Nt strings;
    mobile: Nt flags (default all mobile; box.py:649-668 also makes beads
    with a label in immobile_labels immobile).  Immobile beads stay put: D,
    the noise and the moves cover the N mobile beads (in bead order), the
    overlap test covers every bead, with pairs of two immobile beads exempt
    (bead.py:401-413), and the force field lets immobile beads act on mobile
    ones."""

    def __init__(self, xyz, radii, labels, inp, lib="orc", hard_core=None, epsilon_LJ=None, bonds=(), angles=(),
                 mobile=None):
        self.inp = dict(DEFAULTS)
        self.inp.update(inp)
        self.c = Oracle(lib)
        self.r = np.array(xyz, dtype=np.float64).reshape(-1, 3).copy()
        self.a_all = np.array(radii, dtype=np.float64).copy()
        self.labels_all = list(labels)
        Nt = self.a_all.size
        flags = [True] * Nt if mobile is None else [bool(m) for m in mobile]
        flags = [f and (self.labels_all[i] not in self.inp["immobile_labels"]) for i, f in enumerate(flags)]
        self.mobile = np.array(flags, dtype=bool)
        # rows of the mobile beads in self.r (a slice, hence views, when all move)
        self.mob = slice(None) if self.mobile.all() else np.nonzero(self.mobile)[0]
        self.a = self.a_all[self.mob]
        self.labels = [lab for lab, f in zip(self.labels_all, flags) if f]
        self.N = self.a.size
        self.box_length = float(self.inp["box_length"])
        # box.py:621-629
        if self.inp["seed"] is None:
            self.seed = np.random.randint(2**32 - 1)
        else:
            self.seed = self.inp["seed"]
        self.rng = np.random.RandomState(self.seed)
        self.draw_count = 0
        # box.py:633-637
        self.T = self.inp["T"]
        self.kBT = BOLTZMANN * self.T
        self.viscosity = self.inp["viscosity"]
        self.hydrodynamics = self.inp["hydrodynamics"]
        if self.hydrodynamics in ("rpy_smith", "rpy_smith_lub"):
            self.alpha = self.inp["ewald_alpha"]
            self.m_max = self.inp["ewald_real"]
            self.n_max = self.inp["ewald_imag"]
        if self.hydrodynamics in ("rpy_lub", "rpy_smith_lub"):
            self.lubrication_cutoff = self.inp["lubrication_cutoff"]
            self.cichocki_correction = self.inp["cichocki_correction"]
        self.propagation_scheme = self.inp["propagation_scheme"]
        self.m_midpoint = self.inp["m_midpoint"]
        self.overlaps = self.inp["check_overlaps"]
        # bead.py:401-413: bonded pairs and pairs of immobile beads are not
        # tested for overlaps
        self.mask = np.zeros((Nt, Nt), dtype=np.int32)
        for bond in bonds:
            self.mask[bond[0]][bond[1]] = self.mask[bond[1]][bond[0]] = 1
        imm = np.nonzero(~self.mobile)[0]
        self.mask[np.ix_(imm, imm)] = 1
        if self.hydrodynamics == "nohi":
            self._compute_Dff_matrix()
            self._decompose_D_matrix()
        if self.overlaps:
            assert not self._check_overlaps(), "starting configuration contains overlaps"
        self.Fex = np.array(self.inp["external_force"], dtype=np.float64)
        self.F = np.zeros(3 * self.N)
        # box.py:672-681: a label is mobile when its first bead is
        self.mobile_labels, seen = [], []
        for lab, f in zip(self.labels_all, flags):
            if lab not in seen:
                seen.append(lab)
                if f:
                    self.mobile_labels.append(lab)
        self.is_flux = "measure_flux" in self.inp
        if self.is_flux:
            self.flux_normal = np.array(self.inp["measure_flux"]["normal"], float)
            self.flux_plane_point = np.array(self.inp["measure_flux"]["plane_point"], float)
            self.net_flux = {lab: 0 for lab in self.mobile_labels}
        self.attempt_log = []
        # force field (oracle/forces.py restates box.py:404-461 and interactions.py)
        self.hard_core = np.array(hard_core if hard_core is not None else self.a_all, dtype=np.float64)
        self.epsilon_LJ = np.array(epsilon_LJ if epsilon_LJ is not None else np.zeros(Nt), dtype=np.float64)
        self.bonds, self.angles = list(bonds), list(angles)
        self.E = 0.0

    # -- geometry ---------------------------------------------------------
    @property
    def rm(self):
        """Positions of the mobile beads (a view when every bead moves)."""
        return self.r[self.mob]

    def _pointers(self):
        return self.c.pointer_pbc_matrix(self.rm, self.box_length)

    def _check_overlaps(self):
        return self.c.check_overlaps(self.r, self.a_all, self.box_length, self.inp["overlap_treshold"],
                                     self.mask) == 1

    # -- matrices (box.py:486-556) ---------------------------------------
    def _compute_Dff_matrix(self):
        h = self.hydrodynamics
        if h == "nohi":
            self.D = self.kBT * 10**19 / 6 / np.pi / np.repeat(self.a, 3) / self.viscosity
            return
        P = Oracle.packed(self._pointers())
        if h == "rpy":
            self.D = self.c.rpy_m_matrix(self.a, P)
            self.D *= self.kBT * 10**19 / self.viscosity
        elif h == "rpy_smith":
            self.D = self.c.rpy_smith_m_matrix(self.a, P, self.box_length, self.alpha, self.m_max, self.n_max)
            self.D *= self.kBT * 10**19 / self.viscosity
        elif h == "rpy_lub":
            self.Mff = self.c.rpy_m_matrix(self.a, P) * 10**19 / self.viscosity
            self.Rff = np.linalg.inv(self.Mff)
        elif h == "rpy_smith_lub":
            self.Mff = self.c.rpy_smith_m_matrix(self.a, P, self.box_length, self.alpha, self.m_max,
                                                 self.n_max) * 10**19 / self.viscosity
            self.Rff = np.linalg.inv(self.Mff)
        else:
            raise ValueError(h)

    def _compute_Dtot_matrix(self):
        P = Oracle.packed(self._pointers())
        C = Oracle("orc").jo_r_lubrication_correction_matrix(self.a, P, self.lubrication_cutoff,
                                                              self.cichocki_correction)
        self.R = C * self.viscosity / 10**19 + self.Rff
        self.D = self.kBT * np.linalg.inv(self.R)

    def _decompose_D_matrix(self):
        if self.hydrodynamics == "nohi":
            self.B = np.sqrt(self.D)
        else:
            self.B = np.linalg.cholesky(self.D)

    # -- moves (box.py:353-399) -------------------------------------------
    def _translate(self, v):
        v = np.asarray(v, dtype=np.float64).reshape(-1, 3)
        if self.is_flux:
            n, p = self.flux_normal, self.flux_plane_point
            d0 = self.rm - p
            s0 = (n[0] * d0[:, 0] + n[1] * d0[:, 1] + n[2] * d0[:, 2]) > 0.0
        self.r[self.mob] += v
        if self.is_flux:
            d1 = self.rm - p
            s1 = (n[0] * d1[:, 0] + n[1] * d1[:, 1] + n[2] * d1[:, 2]) > 0.0
            delta = s1.astype(np.int64) - s0.astype(np.int64)
            for k in np.nonzero(delta)[0]:
                self.net_flux[self.labels[k]] += int(delta[k])

    def _deterministic_step(self, dt, mult=1.0):
        if self.hydrodynamics != "nohi":
            FX = dt / self.kBT * self.D @ self.F * mult
        else:
            FX = dt / self.kBT * self.D * self.F * mult
        self._translate(FX)

    def _stochastic_step(self, dt, mult=1.0):
        if self.hydrodynamics == "nohi":
            BX = self.B * self.xi * math.sqrt(2 * dt) * mult
        else:
            BX = self.B @ self.xi * math.sqrt(2 * dt) * mult
        self._translate(BX)

    def _generate_random_vector(self):
        self.xi = self.rng.normal(0.0, 1.0, 3 * self.N)
        self.draw_count += 3 * self.N

    def _compute_divergence(self, build_Dff, build_Dnf):
        """box.py:560-610: forward differences of D, one coordinate at a time."""
        eps = 0.0000001
        lub = self.hydrodynamics in ("rpy_lub", "rpy_smith_lub")
        D0 = np.copy(self.D)
        if lub:
            Mff0, Rff0, R0 = np.copy(self.Mff), np.copy(self.Rff), np.copy(self.R)
        div = -np.sum(self.D, axis=1)
        rows = np.arange(len(self.r))[self.mob]
        for i in range(self.N):
            for j, vector in enumerate(([eps, 0.0, 0.0], [0.0, eps, 0.0], [0.0, 0.0, eps])):
                vector = np.array(vector)
                self.r[rows[i]] += vector
                if build_Dff:
                    self._compute_Dff_matrix()
                if lub and build_Dnf:
                    self._compute_Dtot_matrix()
                div += self.D[:][3 * i + j]
                self.r[rows[i]] += -vector
        self.D = D0
        if lub:
            self.Mff, self.Rff, self.R = Mff0, Rff0, R0
        div /= eps
        return div

    def _ermak_step(self, dt, build_Dff=True, build_Dnf=True):
        self._deterministic_step(dt)
        # box.py:211-215
        if self.inp.get("divergence_term") and self.hydrodynamics in ("rpy_lub", "rpy_smith_lub"):
            self._translate(self._compute_divergence(build_Dff, build_Dnf) * dt)
        attempts = 0
        while True:
            attempts += 1
            self._generate_random_vector()
            self._stochastic_step(dt)
            if self.overlaps:
                if self._check_overlaps():
                    self._stochastic_step(dt, mult=-1)
                else:
                    break
            else:
                break
            assert attempts < self.inp["max_move_attempts"], "maximal number of move attepts exceeded"
        self.attempt_log.append(attempts)

    def _midpoint_step(self, dt, build_Dff, build_Dnf, cholesky):
        lub = self.hydrodynamics in ("rpy_lub", "rpy_smith_lub")
        D0 = np.copy(self.D)
        B0 = np.copy(self.B)
        if lub:
            Mff0, Rff0, R0 = np.copy(self.Mff), np.copy(self.Rff), np.copy(self.R)
        m = self.m_midpoint
        self._deterministic_step(dt, mult=1.0 / m)
        attempts = 0
        while True:
            attempts += 1
            self._generate_random_vector()
            self._stochastic_step(dt, mult=1.0 / m)
            if self.overlaps and self._check_overlaps():
                self._stochastic_step(dt, mult=-1.0 / m)
                continue
            if self.hydrodynamics != "nohi":
                if build_Dff:
                    self._compute_Dff_matrix()
                if lub and build_Dnf:
                    self._compute_Dtot_matrix()
                if cholesky:
                    # box.py:303-304: computed, then discarded by the B0 restore.
                    self._decompose_D_matrix()
            det = m * dt / (2 * self.kBT) * (self.D - D0) @ self.F
            sto = m / 2 * math.sqrt(2 * dt) * (self.D @ np.linalg.inv(np.transpose(B0)) - B0) @ self.xi
            drift = sto + det
            self.D, self.B = D0, B0
            if lub:
                self.Mff, self.Rff, self.R = Mff0, Rff0, R0
            self._deterministic_step(dt, mult=1.0 - 1.0 / m)
            self._stochastic_step(dt, mult=1.0 - 1.0 / m)
            self._translate(drift)
            if self.overlaps and self._check_overlaps():
                self._deterministic_step(dt, mult=-(1.0 - 1.0 / m))
                self._stochastic_step(dt, mult=-1)
                self._translate(-drift)
            else:
                break
            assert attempts < self.inp["max_move_attempts"], "maximal number of move attepts exceeded"
        self.attempt_log.append(attempts)

    def _keep_in_box(self):
        L = self.box_length
        for i in np.arange(len(self.r))[self.mob]:
            for c in range(3):
                while self.r[i, c] < -L / 2:
                    self.r[i, c] += L
                while self.r[i, c] >= L / 2:
                    self.r[i, c] -= L

    # -- box.py:150-198 ---------------------------------------------------
    def propagate(self, dt, build_Dff=True, build_Dnf=True, cholesky=True):
        if self.is_flux:
            self.net_flux = {lab: 0 for lab in self.mobile_labels}
        if self.hydrodynamics != "nohi":
            if build_Dff:
                self._compute_Dff_matrix()
            if self.hydrodynamics in ("rpy_lub", "rpy_smith_lub") and build_Dnf:
                self._compute_Dtot_matrix()
            if cholesky:
                self._decompose_D_matrix()
        self.F, self.E = forces.forces_and_energy(
            self.r, self.box_length, Fex=self.Fex, region=self.inp.get("external_force_region"),
            lj6=bool(self.inp["lennard_jones_6"]), lj12=bool(self.inp["lennard_jones_12"]),
            alpha=self.inp["lennard_jones_alpha"], hc=self.hard_core, eps=self.epsilon_LJ, bonds=self.bonds,
            angles=self.angles, unit=self.inp["energy_unit"], mobile=self.mobile)
        if self.propagation_scheme == "ermak":
            self._ermak_step(dt, build_Dff, build_Dnf)
        elif self.propagation_scheme == "midpoint":
            self._midpoint_step(dt, build_Dff, build_Dnf, cholesky)
        else:
            raise ValueError(self.propagation_scheme)
        self._keep_in_box()
