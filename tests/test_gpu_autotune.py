"""The per-order launch-shape table of the factorisation and its tuner on
the device: the tuner's choices install and are reported back, and every
shape factors to numpy's factor within 1e-12."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def clean_table():
    from pybrown_amd import autotune
    autotune.clear()
    yield
    autotune.clear()


def _shape(T):
    from pybrown_amd._lib import lib
    v, o = ctypes.c_int(), ctypes.c_int()
    assert lib.pbx_get_potrf_shape(T, ctypes.byref(v), ctypes.byref(o)) == 0
    return v.value, o.value


def _spd(n, seed):
    rng = np.random.RandomState(seed)
    G = rng.standard_normal((n, n))
    return G @ G.T / n + np.eye(n)


def test_tuner_installs_its_choices(clean_table):
    from pybrown_amd import autotune
    res = autotune.tune(grid=(2, 5, 20), reps=2, log=print)
    assert res["grid"] == [2, 5, 20] and all(c in autotune.CODES for c in res["codes"])
    assert all(v > 0 for t in res["ms"].values() for v in t.values())
    print("device %s, %.2f s" % (res["device"], res["seconds"]))
    codes = autotune.apply(res)
    for T, c in zip(res["grid"], res["codes"]):
        assert codes[T] == c
        v, own = _shape(T)
        assert v == (4 if c == 4 else 6) and (c == 4 or own == (c == 61))


@pytest.mark.parametrize("n", [300, 1000, 1600])
def test_every_shape_factors_correctly(n, clean_table):
    from pybrown_amd import autotune
    from pybrown_amd.diffusion import cholesky
    A = _spd(n, n)
    Ln = np.linalg.cholesky(A)
    T = -(-n // 64)
    out = {}
    for c in autotune.CODES:
        autotune.apply({"grid": [T], "codes": [c]})
        out[c] = cholesky(A)
        rel = np.abs(out[c] - Ln).max() / np.abs(Ln).max()
        print("n=%d shape %d: rel %.2e" % (n, c, rel))
        assert rel <= 1e-12
    for a, b in ((61, 62), (4, 61)):
        print("shapes %d and %d equal to the bit: %s" % (a, b, np.array_equal(out[a], out[b])))


@pytest.mark.parametrize("K", [1, 2, 3, 4])
def test_whole_trailing_update_period(K):
    """Orders above 64 tile columns (two workgroups per CU) update each far
    tile column pair every K launches (a rolling 1/K of them per launch) and
    the next launch's two tile columns in every launch; every period factors
    to numpy's factor (n = 4500, 71 tile columns, 36 launches)."""
    from pybrown_amd._lib import lib
    from pybrown_amd.diffusion import cholesky
    A = _spd(4500, 11)
    Ln = np.linalg.cholesky(A)
    assert lib.pbx_set_potrf_far_period(K) == 0
    try:
        L = cholesky(A)
    finally:
        lib.pbx_set_potrf_far_period(4)
    rel = np.abs(L - Ln).max() / np.abs(Ln).max()
    print("K=%d: rel %.2e" % (K, rel))
    assert rel <= 1e-12
