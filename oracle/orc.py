"""ctypes front-end for the CPU checker libraries.

TEST INFRASTRUCTURE ONLY.  Loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py; never by the engine in pybrown_amd/.

Two libraries are exposed with the same Python signatures:

* ``Oracle(lib='orc')``  -- the C restatement in oracle/orc.c
* ``Oracle(lib='ref')``  -- the reference's own clib sources compiled by
  oracle/Makefile into oracle/_ref/libpyBrown_ref.so (present only where the
  reference tree was available at build time).  Its lubrication entry point
  needs LAPACK dgetri_, which this image lacks, so it is never called.
* ``Oracle(lib='ref_omp')`` -- the same sources built with OpenMP
  (oracle/_ref/libpyBrown_ref_omp.so), for the multi-core CPU timing; the
  thread count is the OpenMP runtime's (OMP_NUM_THREADS).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_PATH = os.path.join(HERE, "liborc.so")
REF_PATH = os.path.join(HERE, "_ref", "libpyBrown_ref.so")
REF_OMP_PATH = os.path.join(HERE, "_ref", "libpyBrown_ref_omp.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _lazy_open(path):
    """dlopen with RTLD_LAZY so unresolved LAPACK symbols of the reference
    build do not block loading (ctypes.CDLL always adds RTLD_NOW)."""
    libdl = ctypes.CDLL("libdl.so.2")
    libdl.dlopen.restype = ctypes.c_void_p
    libdl.dlopen.argtypes = [ctypes.c_char_p, ctypes.c_int]
    h = libdl.dlopen(path.encode(), os.RTLD_LAZY | os.RTLD_LOCAL)
    if not h:
        raise OSError("dlopen failed for %s" % path)
    return ctypes.CDLL(path, handle=h)


def ref_available(omp=False):
    return os.path.exists(REF_OMP_PATH if omp else REF_PATH)


class Oracle:
    """Calls either the restatement ('orc') or the compiled reference ('ref').

    All matrices are returned as numpy float64 arrays, row-major, matching the
    reference's reshape (pyBrown/diffusion.py:54)."""

    def __init__(self, lib="orc"):
        self.kind = lib
        if lib == "orc":
            if not os.path.exists(ORC_PATH):
                raise OSError("oracle/liborc.so not built (run make -C oracle)")
            self.lib = ctypes.CDLL(ORC_PATH)
            p = "orc_"
            self._names = dict(
                pointer=p + "pointer_pbc_matrix", overlaps=p + "check_overlaps",
                immobile=p + "pointer_immobile_pbc_matrix",
                rpy=p + "rpy_m_matrix", smith=p + "rpy_smith_m_matrix",
                jo2b=p + "jo_2b_r_matrix", lub=p + "jo_r_lubrication_correction_matrix")
        elif lib in ("ref", "ref_omp"):
            if not ref_available(lib == "ref_omp"):
                raise OSError("oracle/_ref/libpyBrown_%s.so not built" % lib)
            self.lib = _lazy_open(REF_OMP_PATH if lib == "ref_omp" else REF_PATH)
            self._names = dict(
                pointer="pointer_pbc_matrix", overlaps="check_overlaps",
                immobile="pointer_immobile_pbc_matrix", rpy="RPY_M_matrix", smith="RPY_Smith_M_matrix",
                jo2b="JO_2B_R_matrix", lub=None)
        else:
            raise ValueError(lib)
        L = self.lib
        f = getattr(L, self._names["pointer"])
        f.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _dp]
        f.restype = None
        f = getattr(L, self._names["immobile"])
        f.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
        f.restype = None
        f = getattr(L, self._names["overlaps"])
        f.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _ip]
        f.restype = ctypes.c_int
        f = getattr(L, self._names["rpy"])
        f.argtypes = [_dp, _dp, ctypes.c_int, _dp]
        f.restype = None
        f = getattr(L, self._names["smith"])
        f.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        f.restype = None
        f = getattr(L, self._names["jo2b"])
        f.argtypes = [ctypes.c_double] * 5 + [_dp]
        f.restype = None
        if self._names["lub"]:
            f = getattr(L, self._names["lub"])
            f.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
            f.restype = None
        if lib == "orc":
            L.orc_packed_pointers.argtypes = [_dp, ctypes.c_int, ctypes.c_double, _dp]
            L.orc_rpy_2b_r_matrix.argtypes = [ctypes.c_double] * 5 + [_dp]
            L.orc_smith_self.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, _dp]
            L.orc_smith_pair.argtypes = [ctypes.c_double] * 7 + [ctypes.c_int, ctypes.c_int, _dp]
            L.orc_rpy_pair.argtypes = [ctypes.c_double] * 5 + [_dp]
            L.orc_invert_small.argtypes = [_dp, ctypes.c_int]
            L.orc_invert_small.restype = ctypes.c_int
            L.orc_cichocki.argtypes = [_dp, _dp]
            for nm in ("orc_jo_xa11", "orc_jo_ya11", "orc_jo_xa12", "orc_jo_ya12"):
                g = getattr(L, nm)
                g.argtypes = [ctypes.c_double, ctypes.c_double]
                g.restype = ctypes.c_double

    # -- geometry ---------------------------------------------------------
    def pointer_pbc_matrix(self, xyz, L):
        """bead.c:27-95; returns (N, N, 3) with rij[a][b] = wrap(r_b - r_a)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        N = xyz.size // 3
        out = np.zeros(3 * N * N)
        getattr(self.lib, self._names["pointer"])(_d(xyz), N, float(L), _d(out))
        return out.reshape(N, N, 3)

    def pointer_immobile_pbc_matrix(self, xyz_mobile, xyz_immobile, L):
        """bead.c:99-158; returns (N_mob, N_immob, 3) with rik[i][j] = wrap(r_j - r_i)."""
        xm = np.ascontiguousarray(xyz_mobile, dtype=np.float64).reshape(-1)
        xi = np.ascontiguousarray(xyz_immobile, dtype=np.float64).reshape(-1)
        nm, ni = xm.size // 3, xi.size // 3
        out = np.zeros(3 * nm * ni)
        getattr(self.lib, self._names["immobile"])(_d(xm), _d(xi), nm, ni, float(L), _d(out))
        return out.reshape(nm, ni, 3)

    @staticmethod
    def packed(rij):
        """pyBrown/diffusion.py:40: -rij[p][q] for p<q, row-major."""
        N = rij.shape[0]
        return np.ascontiguousarray(np.ravel(-rij[np.triu_indices(n=N, k=1)]))

    def check_overlaps(self, xyz, a, L, thr, mask=None):
        """bead.c:169-218; mask is the int32 N x N bond matrix (None = zeros)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1)
        a = np.ascontiguousarray(a, dtype=np.float64)
        N = a.size
        if mask is None:
            mask = np.zeros((N, N), dtype=np.int32)
        mask = np.ascontiguousarray(mask, dtype=np.int32).reshape(-1)
        return int(getattr(self.lib, self._names["overlaps"])(
            _d(xyz), _d(a), N, float(L), float(thr), mask.ctypes.data_as(_ip)))

    # -- mobility ---------------------------------------------------------
    def rpy_m_matrix(self, a, P):
        a = np.ascontiguousarray(a, dtype=np.float64)
        P = np.ascontiguousarray(P, dtype=np.float64)
        N = a.size
        M = np.zeros(9 * N * N)
        getattr(self.lib, self._names["rpy"])(_d(a), _d(P), N, _d(M))
        return M.reshape(3 * N, 3 * N)

    def rpy_smith_m_matrix(self, a, P, L, alpha, m, n):
        a = np.ascontiguousarray(a, dtype=np.float64)
        P = np.ascontiguousarray(P, dtype=np.float64)
        N = a.size
        M = np.zeros(9 * N * N)
        getattr(self.lib, self._names["smith"])(_d(a), _d(P), float(L), float(alpha), int(m), int(n), N, _d(M))
        return M.reshape(3 * N, 3 * N)

    def jo_2b_r_matrix(self, ai, aj, r):
        out = np.zeros(36)
        getattr(self.lib, self._names["jo2b"])(float(ai), float(aj), float(r[0]), float(r[1]), float(r[2]), _d(out))
        return out[:18].copy()

    def jo_r_lubrication_correction_matrix(self, a, P, cutoff, cichocki, C=None):
        if self._names["lub"] is None:
            raise NotImplementedError("reference lubrication needs LAPACK dgetri_, absent from this image")
        a = np.ascontiguousarray(a, dtype=np.float64)
        P = np.ascontiguousarray(P, dtype=np.float64)
        N = a.size
        if C is None:
            C = np.zeros(9 * N * N)
        else:
            C = np.ascontiguousarray(C, dtype=np.float64).reshape(-1).copy()
        getattr(self.lib, self._names["lub"])(_d(a), _d(P), N, float(cutoff), int(bool(cichocki)), _d(C))
        return C.reshape(3 * N, 3 * N)

    # -- restatement-only helpers ----------------------------------------
    def rpy_2b_r_matrix(self, ai, aj, r):
        out = np.zeros(18)
        self.lib.orc_rpy_2b_r_matrix(float(ai), float(aj), float(r[0]), float(r[1]), float(r[2]), _d(out))
        return out

    def invert_small(self, A):
        A = np.ascontiguousarray(A, dtype=np.float64).copy()
        n = A.shape[0]
        rc = self.lib.orc_invert_small(_d(A), n)
        if rc:
            raise np.linalg.LinAlgError("singular")
        return A

    def cichocki(self, t18):
        t = np.ascontiguousarray(t18, dtype=np.float64)
        out = np.zeros(18)
        self.lib.orc_cichocki(_d(t), _d(out))
        return out


def block18_to_6x6(R):
    """Expand the packed 18-value two-sphere layout into a dense 6x6
    (pyBrown/diffusion.py:116-133)."""
    M = np.zeros((6, 6))
    full = lambda b: np.array([[b[0], b[3], b[4]], [b[3], b[1], b[5]], [b[4], b[5], b[2]]])
    M[:3, :3] = full(R[0:6])
    M[3:, 3:] = full(R[6:12])
    M[3:, :3] = full(R[12:18])
    M[:3, 3:] = full(R[12:18]).T
    return M
