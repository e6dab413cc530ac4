"""Mobility and resistance builders with the call signatures of
pyBrown/diffusion.py, computed by the HIP kernels of libpybrown_amd.so.

Each function takes the same arguments as the reference (beads and the
(N, N, 3) pointer matrix) and returns the same (3N, 3N) row-major array.  The
packed pair list handed to the kernels is built exactly as the reference
builds it (diffusion.py:40), with numpy instead of Python lists.
"""
import numpy as np

from pybrown_amd import _lib
from pybrown_amd._lib import lib, ptr


def _radii(beads):
    return np.ascontiguousarray([b.a for b in beads], dtype=np.float64)


def packed_pointers(pointers):
    """diffusion.py:40: -rij[p][q] for p < q in row-major order."""
    N = pointers.shape[0]
    return np.ascontiguousarray(np.ravel(-pointers[np.triu_indices(n=N, k=1)]), dtype=np.float64)


def RPY_M_matrix(beads, pointers):
    """diffusion.py:30-56 -> clib/diff_tensor.c:73"""
    N = len(beads)
    a, P = _radii(beads), packed_pointers(pointers)
    M = np.zeros(9 * N * N)
    lib.RPY_M_matrix(ptr(a), ptr(P), N, ptr(M))
    _lib.check_void("RPY_M_matrix")
    return M.reshape(3 * N, 3 * N)


def RPY_Smith_M_matrix(beads, pointers, box_length, alpha, m, n):
    """diffusion.py:60-94 -> clib/diff_tensor.c:170"""
    N = len(beads)
    a, P = _radii(beads), packed_pointers(pointers)
    M = np.zeros(9 * N * N)
    lib.RPY_Smith_M_matrix(ptr(a), ptr(P), float(box_length), float(alpha), int(m), int(n), N, ptr(M))
    _lib.check_void("RPY_Smith_M_matrix")
    return M.reshape(3 * N, 3 * N)


def JO_2B_R_matrix(beadi, beadj):
    """diffusion.py:98-141 -> clib/diff_tensor.c:270; returns the dense 6x6."""
    out = np.zeros(36)
    r = np.asarray(beadj.r, dtype=np.float64) - np.asarray(beadi.r, dtype=np.float64)
    lib.JO_2B_R_matrix(float(beadi.a), float(beadj.a), float(r[0]), float(r[1]), float(r[2]), ptr(out))
    _lib.check_void("JO_2B_R_matrix")
    v = out[:18]
    R = np.zeros((6, 6))
    R[0][0], R[1][1], R[2][2] = v[0], v[1], v[2]
    R[0][1] = R[1][0] = v[3]
    R[0][2] = R[2][0] = v[4]
    R[1][2] = R[2][1] = v[5]
    R[3][3], R[4][4], R[5][5] = v[6], v[7], v[8]
    R[3][4] = R[4][3] = v[9]
    R[3][5] = R[5][3] = v[10]
    R[4][5] = R[5][4] = v[11]
    R[3][0] = R[0][3] = v[12]
    R[4][1] = R[1][4] = v[13]
    R[5][2] = R[2][5] = v[14]
    R[4][0] = R[0][4] = R[3][1] = R[1][3] = v[15]
    R[5][0] = R[0][5] = R[3][2] = R[2][3] = v[16]
    R[5][1] = R[1][5] = R[4][2] = R[2][4] = v[17]
    return R


def JO_R_lubrication_correction_matrix(beads, pointers, lubrication_cutoff, cichocki_correction):
    """diffusion.py:145-175 -> clib/diff_tensor.c:346 (into a zeroed matrix)."""
    N = len(beads)
    a, P = _radii(beads), packed_pointers(pointers)
    C = np.zeros(9 * N * N)
    lib.JO_R_lubrication_correction_matrix(ptr(a), ptr(P), N, float(lubrication_cutoff),
                                           1 if cichocki_correction else 0, ptr(C))
    _lib.check_void("JO_R_lubrication_correction_matrix")
    return C.reshape(3 * N, 3 * N)


def cholesky(D, precision="fp64"):
    """np.linalg.cholesky (box.py:554) on the FP64 matrix cores; with
    precision="mixed" the trailing updates run on the FP32 matrix cores
    (operands rounded to FP32) and the diagonal tiles stay FP64."""
    if precision not in ("fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed'")
    D = np.ascontiguousarray(D, dtype=np.float64)
    n = D.shape[0]
    L = np.zeros_like(D)
    f = lib.pbx_cholesky if precision == "fp64" else lib.pbx_cholesky_mixed
    rc = f(ptr(D), n, ptr(L))
    _lib.check_rc(rc, "cholesky")
    return L


def spd_inverse(A, scale=1.0, precision="fp64"):
    """scale * np.linalg.inv(A) for SPD A (box.py:514,524,537).  precision
    "mixed": the factorisation's trailing updates and the tile products of
    the triangular inverse and of inv(L)^T inv(L) on the FP32 matrix cores
    (diagonal tiles and stored values FP64); "mixed_refined": that, then one
    Newton-Schulz step X + X (I - A X) with the residual in FP64."""
    if precision not in ("fp64", "mixed", "mixed_refined"):
        raise ValueError("precision must be 'fp64', 'mixed' or 'mixed_refined'")
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    X = np.zeros_like(A)
    if precision == "fp64":
        rc = lib.pbx_spd_inverse_host(ptr(A), n, float(scale), ptr(X))
    else:
        rc = lib.pbx_spd_inverse_mixed(ptr(A), n, float(scale), 1 if precision == "mixed_refined" else 0, ptr(X))
    _lib.check_rc(rc, "spd_inverse")
    return X
