"""``Box``: the reference engine's driver interface (pyBrown/box.py:42-198)
over the device-resident HIP simulation.

    box = Box(beads, input_data)
    box.propagate(dt, build_Dff, build_Dnf, cholesky)   # BD.py:177

Positions, D, B (and Rff for lubrication) live in HBM; after each propagate
the new positions are copied once into a shared host array that every
``Bead.r`` views, so BD.py-style loops read ``box.beads[k].r`` unchanged.
N(0,1) draws are the stream of ``numpy.random.RandomState(seed)``
(box.py:393-399 and :621-629: 3N per move attempt, in attempt order).  By
default the library generates that stream itself (``noise="native"``: the
same MT19937 and polar Gaussian, bit for bit), ahead of use on a producer
thread; ``pseudorandom_number_generator`` is then not advanced.
``noise="callback"`` draws from ``pseudorandom_number_generator`` instead,
one chunk of ``xi_pool_attempts`` attempts at a time.

Immobile beads (``immobile_labels`` or ``Bead.mobile = False``,
box.py:649-668) stay where they are: they are absent from D and the noise,
take part in the overlap test against mobile beads and act on mobile beads
through the force field.  On the device they follow the mobile beads in
index order; their positions are uploaded once, at construction.

Scope: the hydrodynamic step.  Keywords the step does not cover (walls,
reactions, concentration, custom interactions) raise NotImplementedError
instead of being ignored.
"""
import math

import numpy as np

from pybrown_amd import _lib
from pybrown_amd._lib import lib, ptr
from pybrown_amd.bead import bonded_lists
from pybrown_amd.plane import Plane

BOLTZMANN = 1.380649e-23  # scipy.constants.Boltzmann

# BD.py:42-56
DEFAULTS = {
    "debug": False, "verbose": False, "hydrodynamics": "nohi",
    "ewald_alpha": math.sqrt(math.pi), "ewald_real": 0, "ewald_imag": 0,
    "diff_freq": 1, "lub_freq": 1, "chol_freq": 1, "xyz_write_freq": 1,
    "lubrication_cutoff": 1, "progress_bar": False,
    "seed": None, "immobile_labels": [],
    "propagation_scheme": "ermak", "m_midpoint": 100,
    "divergence_term": False, "check_overlaps": True,
    "external_force": [0.0, 0.0, 0.0],
    "lennard_jones_6": False, "lennard_jones_12": False,
    "lennard_jones_alpha": 4.0, "energy_unit": "joule",
    "custom_interactions": False, "cichocki_correction": True,
    "overlap_treshold": 0.0, "max_move_attempts": 1000000,
}

_UNSUPPORTED = ("walls", "measure_concentration", "reactions")

# interactions.py:823-841: factors the reference applies to F for non-SI
# energy units (box.py:414-428)
UNIT_FACTORS = {"joule": 1.0, "kcal/mol": 10 ** (-21) * 6.9477, "eV": 10 ** (-19) * 1.602176634}


# Precision of the dense algebra: "fp64" throughout; "mixed": the noise
# factor's trailing updates and (lubrication modes) the SPD inverses' bulk on
# FP32 matrix cores, diagonal tiles FP64; "mixed_refined": as "mixed", each
# SPD inverse followed by one Newton-Schulz step with an FP64 residual;
# "mixed_factor": the noise factor as "mixed", the SPD inverses in FP64.
PRECISIONS = ("fp64", "mixed", "mixed_refined", "mixed_factor")
_INVERSE_PRECISION = {"fp64": 0, "mixed": 1, "mixed_refined": 2, "mixed_factor": 0}


class Box:
    """Simulation box (pyBrown/box.py:30-109)."""

    def __init__(self, beads, input_data, xi_pool_attempts=64, noise="native", attempt_batch=0, precision="fp64",
                 graphs=False):
        self.inp = dict(DEFAULTS)
        self.inp.update(input_data)
        inp = self.inp
        for k in _UNSUPPORTED:
            if k in inp:
                raise NotImplementedError("keyword %r is outside the hydrodynamic step" % k)
        if inp["custom_interactions"]:
            raise NotImplementedError("custom interactions are Python functions: not run on the device")
        if inp["energy_unit"] not in UNIT_FACTORS:
            raise ValueError("unknown energy_unit %r" % inp["energy_unit"])
        if inp["hydrodynamics"] not in _lib.HYDRO_CODES:
            raise ValueError("unknown hydrodynamics %r" % inp["hydrodynamics"])
        if inp["propagation_scheme"] not in _lib.SCHEME_CODES:
            raise ValueError("unknown propagation_scheme %r" % inp["propagation_scheme"])

        if precision not in PRECISIONS:
            raise ValueError("precision must be one of %s" % ", ".join(repr(p) for p in PRECISIONS))
        self.precision = precision
        self._init_args = (xi_pool_attempts, noise, attempt_batch, precision, bool(graphs))
        self.beads = beads
        self.box_length = inp["box_length"]
        # box.py:621-629
        if inp["seed"] is None:
            self.seed = np.random.randint(2**32 - 1)
        else:
            self.seed = inp["seed"]
        self.pseudorandom_number_generator = np.random.RandomState(self.seed)

        # box.py:649-668 mobility split
        for b in self.beads:
            if b.label in inp["immobile_labels"]:
                b.mobile = False
        self.mobile_beads = [b for b in self.beads if b.mobile]
        self.immobile_beads = [b for b in self.beads if not b.mobile]
        self.mobile_bead_indices = [i for i, b in enumerate(self.beads) if b.mobile]
        self.immobile_bead_indices = [i for i, b in enumerate(self.beads) if not b.mobile]
        if not self.mobile_beads:
            raise ValueError("Box needs at least one mobile bead")
        # box.py:672-681 labels in order of first appearance; a label counts
        # as mobile when the first bead carrying it is mobile
        self.labels, self.mobile_labels = [], []
        for b in self.beads:
            if b.label not in self.labels:
                self.labels.append(b.label)
                if b.mobile:
                    self.mobile_labels.append(b.label)
        if "measure_flux" in inp and any(b.label not in self.mobile_labels for b in self.mobile_beads):
            raise ValueError("measure_flux: a mobile bead carries a label first seen on an immobile bead")
        label_idx = np.array([self.mobile_labels.index(b.label) if b.label in self.mobile_labels else 0
                              for b in self.mobile_beads], dtype=np.int32)
        # device order: mobile beads, then immobile ones
        self._dev_beads = self.mobile_beads + self.immobile_beads

        # box.py:633-637
        self.T = inp["T"]
        self.kBT = BOLTZMANN * self.T
        self.viscosity = inp["viscosity"]
        self.hydrodynamics = inp["hydrodynamics"]
        self.propagation_scheme = inp["propagation_scheme"]
        if self.hydrodynamics in ("rpy_smith", "rpy_smith_lub"):
            self.alpha, self.m_max, self.n_max = inp["ewald_alpha"], inp["ewald_real"], inp["ewald_imag"]
        if self.hydrodynamics in ("rpy_lub", "rpy_smith_lub"):
            self.lubrication_cutoff = inp["lubrication_cutoff"]
            self.cichocki_correction = inp["cichocki_correction"]
        if self.propagation_scheme == "midpoint":
            self.m_midpoint = inp["m_midpoint"]
        self.overlaps = inp["check_overlaps"]

        # box.py:750-757 flux plane (see Plane; the reference passes a list here)
        self.is_flux = "measure_flux" in inp
        if self.is_flux:
            self.flux_normal = np.array(inp["measure_flux"]["normal"], float)
            self.flux_plane_point = np.array(inp["measure_flux"]["plane_point"], float)
            self.planes = Plane(self.flux_plane_point, self.flux_normal)
            self.net_flux = {lab: 0 for lab in self.mobile_labels}
        self.is_concentration = False
        self.is_energy = "enr_write_freq" in inp  # box.py:77-79; E stays 0 (no force field)
        self.is_wall = False
        self.end_simulation = False
        self.E = 0.0
        self.Fex = np.array(inp["external_force"], dtype=np.float64)

        cfg = _lib.Config()
        cfg.hydrodynamics = _lib.HYDRO_CODES[self.hydrodynamics]
        cfg.scheme = _lib.SCHEME_CODES[self.propagation_scheme]
        cfg.ewald_real = int(inp["ewald_real"])
        cfg.ewald_imag = int(inp["ewald_imag"])
        cfg.cichocki_correction = 1 if inp["cichocki_correction"] else 0
        cfg.check_overlaps = 1 if self.overlaps else 0
        cfg.measure_flux = 1 if self.is_flux else 0
        cfg.n_labels = max(1, len(self.mobile_labels))
        cfg.max_move_attempts = int(inp["max_move_attempts"])
        cfg.xi_pool_attempts = int(xi_pool_attempts)
        cfg.factor_precision = 0 if precision == "fp64" else 1
        cfg.divergence_term = 1 if inp["divergence_term"] else 0
        # bead.py:401-413 connection matrix: bonded pairs skip the overlap test
        self._exempt = _overlap_exemptions(self._dev_beads)
        cfg.n_overlap_exempt = len(self._exempt) // 2
        cfg.overlap_exempt = _lib.iptr(self._exempt) if len(self._exempt) else None
        cfg.box_length = float(self.box_length)
        cfg.T = float(self.T)
        cfg.viscosity = float(self.viscosity)
        cfg.ewald_alpha = float(inp["ewald_alpha"])
        cfg.lubrication_cutoff = float(inp["lubrication_cutoff"])
        cfg.m_midpoint = float(inp["m_midpoint"])
        cfg.overlap_treshold = float(inp["overlap_treshold"])
        if self.is_flux:
            for c in range(3):
                cfg.flux_normal[c] = self.flux_normal[c]
                cfg.flux_plane_point[c] = self.flux_plane_point[c]
        for c in range(3):
            cfg.external_force[c] = float(self.Fex[c])
        self._cfg = cfg

        N = len(self.mobile_beads)
        self.N = N
        Ni = len(self.immobile_beads)
        self._imm_xyz = np.ascontiguousarray(np.array([b.r for b in self.immobile_beads], dtype=np.float64)
                                             .reshape(-1))
        self._imm_radii = np.ascontiguousarray([b.a for b in self.immobile_beads], dtype=np.float64)
        cfg.n_immobile = Ni
        cfg.immobile_xyz = ptr(self._imm_xyz) if Ni else None
        cfg.immobile_radii = ptr(self._imm_radii) if Ni else None
        self._pos = np.ascontiguousarray(np.array([b.r for b in self.mobile_beads], dtype=np.float64).reshape(N, 3))
        self._bind_views()
        radii = np.ascontiguousarray([b.a for b in self.mobile_beads], dtype=np.float64)
        handle = ctypes_void_p()
        rc = lib.pbx_sim_create(cfg, N, ptr(radii), ptr(self._pos.reshape(-1)), _lib.iptr(label_idx), handle)
        _lib.check_rc(rc, "Box")
        self._sim = handle.value
        self._fill_cb = _lib.XI_FILL(self._fill)
        seed_ok = isinstance(self.seed, (int, np.integer)) and 0 <= int(self.seed) < 2**32
        if noise not in ("native", "callback"):
            raise ValueError("noise must be 'native' or 'callback'")
        if noise == "native" and seed_ok:
            _lib.check_rc(lib.pbx_sim_set_seed(self._sim, int(self.seed)), "set_seed")
            self.noise = "native"
        else:
            _lib.check_rc(lib.pbx_sim_set_noise_source(self._sim, self._fill_cb, None), "set_noise_source")
            self.noise = "callback"
        _lib.check_rc(lib.pbx_sim_set_attempt_batch(self._sim, int(attempt_batch)), "set_attempt_batch")
        # the factorisation and SPD-inverse launch sequences as HIP graphs
        _lib.check_rc(lib.pbx_sim_set_graphs(self._sim, 1 if graphs else 0), "set_graphs")
        _lib.check_rc(lib.pbx_sim_set_inverse_precision(self._sim, _INVERSE_PRECISION[precision]),
                      "set_inverse_precision")
        self._set_forces()
        self._synced = self._pos.copy()
        self._counts = np.zeros(max(1, len(self.mobile_labels)), dtype=np.int64)
        self._pos_ptr = ptr(self._pos.reshape(-1))
        self._flux_ptr = self._counts.ctypes.data_as(_lib._lp)

    # -- force field (box.py:404-461, 685-723; interactions.py:483-589) ----
    def _set_forces(self):
        import ctypes
        inp = self.inp
        f = _lib.Forces()
        f.lennard_jones_6 = 1 if inp["lennard_jones_6"] else 0
        f.lennard_jones_12 = 1 if inp["lennard_jones_12"] else 0
        f.lennard_jones_alpha = float(inp["lennard_jones_alpha"])
        self._hc = np.ascontiguousarray([b.hard_core_radius for b in self._dev_beads], dtype=np.float64)
        self._eps = np.ascontiguousarray([b.epsilon_LJ for b in self._dev_beads], dtype=np.float64)
        f.hard_core_radius = ptr(self._hc)
        f.epsilon_LJ = ptr(self._eps)
        # device order = the reference's loop order (mobile beads, then
        # immobile ones) and its id lookup (mobile beads first)
        bonds, angles = bonded_lists(self._dev_beads)
        self.bonds, self.angles = bonds, angles
        self._bond_beads = np.ascontiguousarray([[q[0], q[1]] for q in bonds] or [[0, 0]], dtype=np.int32)
        self._bond_params = np.ascontiguousarray([[q[2], q[3]] for q in bonds] or [[0.0, 0.0]], dtype=np.float64)
        self._angle_beads = np.ascontiguousarray([[q[0], q[1], q[2]] for q in angles] or [[0, 0, 0]], dtype=np.int32)
        self._angle_params = np.ascontiguousarray([[q[3], q[4]] for q in angles] or [[0.0, 0.0]], dtype=np.float64)
        f.n_bonds, f.n_angles = len(bonds), len(angles)
        f.bond_beads = _lib.iptr(self._bond_beads)
        f.bond_params = ptr(self._bond_params)
        f.angle_beads = _lib.iptr(self._angle_beads)
        f.angle_params = ptr(self._angle_params)
        region = inp.get("external_force_region") or {}
        for c, ax in enumerate("xyz"):
            if ax in region:
                f.region_on[c] = 1
                f.region[2 * c], f.region[2 * c + 1] = float(region[ax][0]), float(region[ax][1])
        f.unit = UNIT_FACTORS[inp["energy_unit"]]
        _lib.check_rc(lib.pbx_sim_set_forces(self._sim, ctypes.byref(f)), "set_forces")
        self._E = np.zeros(1)

    def _read_energy(self):
        lib.pbx_sim_get_energy(self._sim, ptr(self._E))
        self.E = float(self._E[0])

    # -- plumbing ---------------------------------------------------------
    def _bind_views(self):
        for i, b in enumerate(self.mobile_beads):
            if hasattr(b, "_bind"):
                b._bind(self._pos[i])
            else:
                b.r = self._pos[i]

    def _fill(self, user, out, count):
        try:
            buf = np.ctypeslib.as_array(out, shape=(count,))
            buf[:] = self.pseudorandom_number_generator.normal(0.0, 1.0, count)
            return 0
        except Exception:  # pragma: no cover - reported through the return code
            return 1

    def _push_host_positions(self):
        """Upload positions if the caller moved beads since the last step
        (bead positions are views of self._pos; see Bead.r)."""
        if not np.array_equal(self._pos, self._synced):
            rc = lib.pbx_sim_set_positions(self._sim, ptr(self._pos.reshape(-1)))
            _lib.check_rc(rc, "set_positions")

    # -- box.py:150-198 ---------------------------------------------------
    def propagate(self, dt, build_Dff=True, build_Dnf=True, cholesky=True):
        self._push_host_positions()
        rc = lib.pbx_sim_propagate_out(self._sim, float(dt), int(bool(build_Dff)), int(bool(build_Dnf)),
                                       int(bool(cholesky)), self._pos_ptr, self._flux_ptr)
        _lib.check_rc(rc, "propagate")
        self._synced[:] = self._pos
        self._read_energy()
        if self.is_flux:
            self.net_flux = {lab: int(self._counts[k]) for k, lab in enumerate(self.mobile_labels)}

    def run(self, dt, steps, j0=0, diff_freq=1, lub_freq=1, chol_freq=1, xyz_every=0):
        """`steps` iterations of the BD.py:164-185 loop (propagate with its
        cadence flags) inside the library, without returning per step.
        Returns (flux, frames): flux[t, l] is label l's flux of step j0+t
        (zeros without measure_flux); frames holds the positions before each
        step j with j % xyz_every == 0 (None when xyz_every is 0)."""
        self._push_host_positions()
        nl = max(1, len(self.mobile_labels))
        flux = np.zeros((int(steps), nl), dtype=np.int64)
        frames = None
        if xyz_every:
            nf = sum(1 for j in range(j0, j0 + steps) if j % xyz_every == 0)
            frames = np.zeros((nf, self.N, 3))
        rc = lib.pbx_sim_run(self._sim, float(dt), int(j0), int(steps), int(diff_freq), int(lub_freq),
                             int(chol_freq), flux.ctypes.data_as(_lib._lp),
                             ptr(frames.reshape(-1)) if frames is not None else None, int(xyz_every))
        _lib.check_rc(rc, "run")
        _lib.check_rc(lib.pbx_sim_get_positions(self._sim, ptr(self._pos.reshape(-1))), "get_positions")
        self._synced[:] = self._pos
        self._read_energy()
        if self.is_flux and steps:
            self.net_flux = {lab: int(flux[-1, k]) for k, lab in enumerate(self.mobile_labels)}
        return flux, frames

    @property
    def F(self):
        """The force vector (3N) the last step used (box.py:404-428)."""
        out = np.zeros(3 * self.N)
        _lib.check_rc(lib.pbx_sim_get_force(self._sim, ptr(out)), "get_force")
        return out

    def factor_residual(self):
        """||D - B B^T||_F / ||D||_F of the current matrices (on the device)."""
        import ctypes
        v = ctypes.c_double(0.0)
        _lib.check_rc(lib.pbx_sim_factor_residual(self._sim, ctypes.byref(v)), "factor_residual")
        return v.value

    @property
    def device_bytes(self):
        return int(lib.pbx_sim_device_bytes(self._sim))

    # -- state queries ----------------------------------------------------
    @property
    def draw_count(self):
        return int(lib.pbx_sim_draw_count(self._sim))

    @property
    def last_move_attempts(self):
        return int(lib.pbx_sim_last_attempts(self._sim))

    def _matrix(self, which):
        n = 3 * self.N
        if self.hydrodynamics == "nohi":
            out = np.zeros(n)
            _lib.check_rc(lib.pbx_sim_get_diag(self._sim, which, ptr(out)), "get_diag")
            return out
        out = np.zeros((n, n))
        _lib.check_rc(lib.pbx_sim_get_matrix(self._sim, which, ptr(out.reshape(-1))), "get_matrix")
        return out

    @property
    def D(self):
        return self._matrix(0)

    @property
    def B(self):
        return self._matrix(1)

    @property
    def Rff(self):
        return self._matrix(2)

    # -- checkpoint / restart (test_box.py:53-117 pickles a Box mid-run) ----
    def snapshot(self):
        """Device and noise state as bytes (pbx_sim_snapshot)."""
        import ctypes
        size = int(lib.pbx_sim_snapshot_size(self._sim))
        buf = ctypes.create_string_buffer(size)
        _lib.check_rc(lib.pbx_sim_snapshot(self._sim, buf, size), "snapshot")
        return buf.raw

    def restore(self, data):
        """Resume from snapshot() of a Box with the same beads and keywords."""
        _lib.check_rc(lib.pbx_sim_restore(self._sim, data, len(data)), "restore")
        _lib.check_rc(lib.pbx_sim_get_positions(self._sim, ptr(self._pos.reshape(-1))), "get_positions")
        self._synced[:] = self._pos

    def __getstate__(self):
        return {"beads": self.beads, "input_data": dict(self.inp, seed=self.seed), "init": self._init_args,
                "net_flux": getattr(self, "net_flux", None), "rng": self.pseudorandom_number_generator,
                "snapshot": self.snapshot()}

    def __setstate__(self, st):
        xi, noise, batch = st["init"][:3]
        precision = st["init"][3] if len(st["init"]) > 3 else "fp64"
        graphs = st["init"][4] if len(st["init"]) > 4 else False
        self.__init__(st["beads"], st["input_data"], xi_pool_attempts=xi, noise=noise, attempt_batch=batch,
                      precision=precision, graphs=graphs)
        self.pseudorandom_number_generator = st["rng"]
        self.restore(st["snapshot"])
        if st["net_flux"] is not None:
            self.net_flux = st["net_flux"]

    def close(self):
        if getattr(self, "_sim", None):
            lib.pbx_sim_destroy(self._sim)
            self._sim = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __str__(self):
        return "simulation box"


def _overlap_exemptions(beads):
    """Bead index pairs (flat int32) with a bond between them in the sense of
    Bead.is_there_bond_between (bead.py:187-189): either bead's bonded_with
    holds the other's bead_id."""
    by_id = {}
    for n, b in enumerate(beads):
        if b.bead_id is not None:
            by_id.setdefault(b.bead_id, []).append(n)
    pairs = set()
    for i, b in enumerate(beads):
        for other in b.bonded_with:
            for j in by_id.get(other, ()):
                if j != i:
                    pairs.add((min(i, j), max(i, j)))
    return np.ascontiguousarray(sorted(pairs), dtype=np.int32).reshape(-1)


def ctypes_void_p():
    import ctypes
    return ctypes.c_void_p()
