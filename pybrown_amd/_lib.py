"""ctypes binding of libpybrown_amd.so (C ABI declared in include/pybrown_amd.h).

The library is required: there is no CPU fallback.  Importing this module
raises ImportError when the in-tree build (pybrown_amd/libpybrown_amd.so) is
absent; calls raise RuntimeError on any HIP or argument error reported by
the library.
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpybrown_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("pybrown_amd: %s is not built; run `python -m pybrown_amd.build`" % LIB_PATH)

lib = ctypes.CDLL(LIB_PATH)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_longlong)
_d = ctypes.c_double
_i = ctypes.c_int

NOHI, RPY, RPY_SMITH, RPY_LUB, RPY_SMITH_LUB = range(5)
ERMAK, MIDPOINT = 0, 1
HYDRO_CODES = {"nohi": NOHI, "rpy": RPY, "rpy_smith": RPY_SMITH, "rpy_lub": RPY_LUB, "rpy_smith_lub": RPY_SMITH_LUB}
SCHEME_CODES = {"ermak": ERMAK, "midpoint": MIDPOINT}

ERR_MOVES = -4
ERR_OVERLAP = -5


class Config(ctypes.Structure):
    """Mirror of pbx_config (include/pybrown_amd.h)."""
    _fields_ = [
        ("hydrodynamics", _i), ("scheme", _i), ("ewald_real", _i), ("ewald_imag", _i),
        ("cichocki_correction", _i), ("check_overlaps", _i), ("measure_flux", _i), ("n_labels", _i),
        ("max_move_attempts", ctypes.c_longlong), ("xi_pool_attempts", _i), ("factor_precision", _i),
        ("box_length", _d), ("T", _d), ("viscosity", _d), ("ewald_alpha", _d),
        ("lubrication_cutoff", _d), ("m_midpoint", _d), ("overlap_treshold", _d),
        ("flux_normal", _d * 3), ("flux_plane_point", _d * 3), ("external_force", _d * 3),
        ("divergence_term", _i), ("n_overlap_exempt", _i), ("overlap_exempt", _ip),
        ("n_immobile", _i), ("immobile_xyz", _dp), ("immobile_radii", _dp),
    ]


class Forces(ctypes.Structure):
    """Mirror of pbx_forces (include/pybrown_amd.h)."""
    _fields_ = [
        ("lennard_jones_6", _i), ("lennard_jones_12", _i), ("lennard_jones_alpha", _d),
        ("hard_core_radius", _dp), ("epsilon_LJ", _dp),
        ("n_bonds", _i), ("bond_beads", _ip), ("bond_params", _dp),
        ("n_angles", _i), ("angle_beads", _ip), ("angle_params", _dp),
        ("region_on", _i * 3), ("region", _d * 6), ("unit", _d),
    ]


XI_FILL = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, _dp, ctypes.c_longlong)


def _sig(name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


# reference-compatible layer
_sig("RPY_M_matrix", None, [_dp, _dp, _i, _dp])
_sig("RPY_Smith_M_matrix", None, [_dp, _dp, _d, _d, _i, _i, _i, _dp])
_sig("JO_2B_R_matrix", None, [_d, _d, _d, _d, _d, _dp])
_sig("JO_R_lubrication_correction_matrix", None, [_dp, _dp, _i, _d, _i, _dp])
_sig("pointer_pbc_matrix", None, [_dp, _i, _d, _dp])
_sig("pointer_immobile_pbc_matrix", None, [_dp, _dp, _i, _i, _d, _dp])
_sig("check_overlaps", _i, [_dp, _dp, _i, _d, _d, _ip])
# dense
_sig("pbx_cholesky", _i, [_dp, _i, _dp])
_sig("pbx_cholesky_mixed", _i, [_dp, _i, _dp])
_sig("pbx_spd_inverse_host", _i, [_dp, _i, _d, _dp])
_sig("pbx_spd_inverse_mixed", _i, [_dp, _i, _d, _i, _dp])
_sig("pbx_spd_solve", _i, [_dp, _i, _dp, _i, _dp, _dp, _dp])
_sig("pbx_cholesky_noalloc", _i, [ctypes.c_void_p, _i, _i, ctypes.c_void_p, _i, ctypes.c_void_p])
_sig("pbx_gemm_nt", _i, [_dp, _dp, _dp, _i, _i, _i])
_sig("pbx_set_potrf_variant", _i, [_i])
_sig("pbx_get_potrf_variant", _i, [])
_sig("pbx_set_potrf_chain_jobs", _i, [_i])
_sig("pbx_set_potrf_own_cu_tiles", _i, [_i])
_sig("pbx_get_potrf_own_cu_tiles", _i, [])
_sig("pbx_set_potrf_far_period", _i, [_i])
_sig("pbx_debug_two_col_updates", _i, [_i, _i, _i, _ip, _i])
_sig("pbx_set_potrf_own_cu_roll", _i, [_i])
_sig("pbx_set_potrf_table", _i, [ctypes.POINTER(ctypes.c_int), _i])
_sig("pbx_get_potrf_shape", _i, [_i, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)])
_sig("pbx_time_potrf", _i, [_i, _i, _i, ctypes.POINTER(ctypes.c_double)])
_sig("pbx_device_name", _i, [ctypes.c_char_p, _i])

_sig("pbx_debug_inject_wait_failure", _i, [_i])
_sig("pbx_kernel_launches", ctypes.c_longlong, [])
# simulation
_sig("pbx_sim_create", _i, [ctypes.POINTER(Config), _i, _dp, _dp, _ip, ctypes.POINTER(ctypes.c_void_p)])
_sig("pbx_sim_destroy", _i, [ctypes.c_void_p])
_sig("pbx_sim_set_noise_source", _i, [ctypes.c_void_p, XI_FILL, ctypes.c_void_p])
_sig("pbx_sim_set_seed", _i, [ctypes.c_void_p, ctypes.c_uint])
_sig("pbx_sim_set_attempt_batch", _i, [ctypes.c_void_p, _i])
_sig("pbx_sim_set_graphs", _i, [ctypes.c_void_p, _i])
_sig("pbx_sim_set_inverse_precision", _i, [ctypes.c_void_p, _i])
_sig("pbx_sim_set_lookahead", _i, [ctypes.c_void_p, _i])
_sig("pbx_sim_snapshot_size", ctypes.c_longlong, [ctypes.c_void_p])
_sig("pbx_sim_snapshot", _i, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong])
_sig("pbx_sim_restore", _i, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong])
_sig("pbx_rng_normals", _i, [ctypes.c_uint, ctypes.c_longlong, _dp, ctypes.c_longlong])
_sig("pbx_sim_set_forces", _i, [ctypes.c_void_p, ctypes.POINTER(Forces)])
_sig("pbx_sim_get_energy", _i, [ctypes.c_void_p, _dp])
_sig("pbx_sim_get_force", _i, [ctypes.c_void_p, _dp])
_sig("pbx_sim_propagate", _i, [ctypes.c_void_p, _d, _i, _i, _i])
_sig("pbx_sim_propagate_out", _i, [ctypes.c_void_p, _d, _i, _i, _i, _dp, _lp])
_sig("pbx_sim_run", _i, [ctypes.c_void_p, _d, ctypes.c_longlong, ctypes.c_longlong, _i, _i, _i, _lp, _dp, _i])
_sig("pbx_sim_device_bytes", ctypes.c_longlong, [ctypes.c_void_p])
_sig("pbx_sim_factor_residual", _i, [ctypes.c_void_p, _dp])
_sig("pbx_sim_get_positions", _i, [ctypes.c_void_p, _dp])
_sig("pbx_sim_set_positions", _i, [ctypes.c_void_p, _dp])
_sig("pbx_sim_get_flux", _i, [ctypes.c_void_p, _lp])
_sig("pbx_sim_get_matrix", _i, [ctypes.c_void_p, _i, _dp])
_sig("pbx_sim_get_diag", _i, [ctypes.c_void_p, _i, _dp])
_sig("pbx_sim_draw_count", ctypes.c_longlong, [ctypes.c_void_p])
_sig("pbx_sim_last_attempts", _i, [ctypes.c_void_p])
_sig("pbx_sim_sync", _i, [ctypes.c_void_p])
_sig("pbx_sim_set_timing", _i, [ctypes.c_void_p, _i])
_sig("pbx_sim_get_timing", ctypes.c_longlong, [ctypes.c_void_p, _dp])
# errors
_sig("pbx_last_error", ctypes.c_char_p, [])
_sig("pbx_last_error_code", _i, [])
_sig("pbx_clear_error", None, [])
_sig("pbx_version", ctypes.c_char_p, [])

EXPORTED = [
    "RPY_M_matrix", "RPY_Smith_M_matrix", "JO_2B_R_matrix", "JO_R_lubrication_correction_matrix",
    "pointer_pbc_matrix", "pointer_immobile_pbc_matrix", "check_overlaps", "pbx_cholesky", "pbx_cholesky_mixed",
    "pbx_spd_inverse_host", "pbx_spd_inverse_mixed", "pbx_spd_solve", "pbx_sim_set_inverse_precision", "pbx_gemm_nt", "pbx_cholesky_noalloc",
    "pbx_set_potrf_variant",
    "pbx_get_potrf_variant",
    "pbx_set_potrf_chain_jobs", "pbx_set_potrf_own_cu_tiles", "pbx_get_potrf_own_cu_tiles", "pbx_set_potrf_far_period",
    "pbx_set_potrf_own_cu_roll",
    "pbx_set_potrf_table", "pbx_get_potrf_shape", "pbx_time_potrf", "pbx_device_name",
    "pbx_debug_inject_wait_failure", "pbx_debug_two_col_updates",
    "pbx_kernel_launches",
    "pbx_sim_create", "pbx_sim_destroy", "pbx_sim_set_noise_source", "pbx_sim_set_seed",
    "pbx_sim_set_attempt_batch", "pbx_sim_set_graphs", "pbx_sim_set_lookahead", "pbx_rng_normals", "pbx_sim_snapshot_size", "pbx_sim_snapshot",
    "pbx_sim_restore", "pbx_sim_set_forces", "pbx_sim_get_energy", "pbx_sim_get_force", "pbx_sim_propagate",
    "pbx_sim_propagate_out",
    "pbx_sim_run",
    "pbx_sim_device_bytes",
    "pbx_sim_factor_residual",
    "pbx_sim_get_positions", "pbx_sim_set_positions", "pbx_sim_get_flux", "pbx_sim_get_matrix",
    "pbx_sim_get_diag", "pbx_sim_draw_count", "pbx_sim_last_attempts", "pbx_sim_sync",
    "pbx_sim_set_timing", "pbx_sim_get_timing",
    "pbx_last_error", "pbx_last_error_code", "pbx_clear_error", "pbx_version",
]


def ptr(a):
    """double* of a C-contiguous float64 array."""
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def last_error():
    msg = lib.pbx_last_error()
    return msg.decode() if msg else ""


def check_void(where):
    """Raise if the last void reference-compatible call reported an error."""
    code = lib.pbx_last_error_code()
    if code != 0:
        raise RuntimeError("%s: %s" % (where, last_error()))


def check_rc(rc, where):
    if rc == 0:
        return
    if rc > 0:
        raise np.linalg.LinAlgError("%s: Matrix is not positive definite (%s)" % (where, last_error()))
    if rc == ERR_MOVES:
        raise AssertionError("maximal number of move attepts exceeded")
    if rc == ERR_OVERLAP:
        raise AssertionError("starting configuration contains overlaps")
    raise RuntimeError("%s failed (%d): %s" % (where, rc, last_error()))
