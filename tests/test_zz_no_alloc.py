"""pbx_cholesky_noalloc (the Cholesky entry point that allocates nothing;
its work area is a static device array) against pbx_cholesky, the entry
point it varies, through the C ABI: the same factor bit for bit, the pad
and upper triangle as documented, on two calls in a row (the second with a
smaller order, so it reuses the array a larger call left behind), a
non-positive pivot reported like pbx_cholesky, and orders beyond the array
rejected before any work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spd(n, seed):
    rng = np.random.RandomState(seed)
    X = rng.normal(size=(n, n))
    return X @ X.T / n + np.eye(n)


def _host_factor(A):
    from pybrown_amd._lib import lib, ptr
    n = A.shape[0]
    L = np.zeros_like(A)
    rc = lib.pbx_cholesky(ptr(np.ascontiguousarray(A)), n, ptr(L))
    return rc, L


def _hip():
    """The HIP runtime the library itself is linked against (already loaded
    with it), for the test's own device buffers."""
    from pybrown_amd._lib import lib  # noqa: F401  (loads libamdhip64.so.7 first)
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


H2D, D2H = 1, 2


def _dev(hip, host):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), host.nbytes) == 0
    assert hip.hipMemcpy(p, host.ctypes.data, host.nbytes, H2D) == 0
    return p


def _noalloc_factor(A, ld_extra):
    from pybrown_amd._lib import lib
    hip = _hip()
    n = A.shape[0]
    np_ = -(-n // 64) * 64
    ld = np_ + ld_extra
    D = _dev(hip, np.ascontiguousarray(A))
    B = _dev(hip, np.full((np_, ld), np.nan))  # the output starts as garbage
    try:
        rc = lib.pbx_cholesky_noalloc(D, n, n, B, ld, None)
        out = np.empty((np_, ld))
        assert hip.hipMemcpy(out.ctypes.data, B, out.nbytes, D2H) == 0
    finally:
        hip.hipFree(D)
        hip.hipFree(B)
    return rc, out


def test_noalloc_cholesky_matches_the_allocating_entry_point():
    from pybrown_amd._lib import lib
    for n, seed, ld_extra in ((1500, 1, 0), (700, 2, 64)):
        A = _spd(n, seed)
        rc_ref, L_ref = _host_factor(A)
        rc, B = _noalloc_factor(A, ld_extra)
        assert rc == 0 and rc_ref == 0, (rc, rc_ref, lib.pbx_last_error())
        np_ = B.shape[0]
        assert np.array_equal(B[:n, :n], L_ref), "n=%d: factor differs from pbx_cholesky" % n
        pad = B[:np_, :np_].copy()
        assert np.array_equal(pad[n:, n:], np.eye(np_ - n))
        assert not pad[:n, n:].any() and not pad[n:, :n].any()
        assert np.array_equal(np.triu(pad, 1), np.zeros_like(pad))


def test_noalloc_cholesky_reports_a_pivot_and_rejects_large_orders():
    from pybrown_amd._lib import lib
    A = _spd(300, 3)
    A[200, 200] = -1.0
    rc_ref, _ = _host_factor(A)
    rc, _ = _noalloc_factor(A, 0)
    assert rc == rc_ref and rc > 0
    # an order beyond the static array: rejected with a clear error, nothing run
    rc = lib.pbx_cholesky_noalloc(ctypes.c_void_p(8), 12289, 12289, ctypes.c_void_p(16), 12352, None)
    assert rc == -1 and b"12288" in lib.pbx_last_error()
