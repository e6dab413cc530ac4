"""The two-column factorisation's launch plan and tile decoding (variant 6,
pbx_dense.hip: two_col_plan, two_col_trailing), run on the host through
pbx_debug_two_col_updates, against what a right-looking Cholesky requires:
every lower tile (i, j) receives the updates of panels 0 .. c0-1 exactly
once, in panel order, before the launch that factors its column (c0 the
first column of that launch), and no launch applies a panel before the launch
that produced it.  Checked for one and two workgroups per CU, far periods
1..6 and orders up to 70 tile columns plus the rolling/whole-trailing
boundary (128/129) and the served and C4 orders (113, 188)."""
import ctypes

import numpy as np
import pytest


def _updates(T, own_cu, K):
    from pybrown_amd._lib import lib
    n = lib.pbx_debug_two_col_updates(T, own_cu, K, None, 0)
    assert n >= 0
    rec = np.zeros(5 * max(n, 1), dtype=np.int32)
    m = lib.pbx_debug_two_col_updates(T, own_cu, K, rec.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n)
    assert m == n
    return rec[:5 * n].reshape(n, 5)


def _launch_c0(T):
    """c0 of each launch, and the launch that factors each column."""
    c0s, col_launch = [], {}
    c0 = 0
    while c0 < T:
        nc = 2 if T - c0 >= 2 else 1
        for j in range(c0, c0 + nc):
            col_launch[j] = len(c0s)
        c0s.append(c0)
        c0 += nc
    return c0s, col_launch


def _check(T, own_cu, K):
    rec = _updates(T, own_cu, K)
    c0s, col_launch = _launch_c0(T)
    per_tile = {}
    for L, i, j, m0, P in rec:
        assert 0 <= j <= i < T, (T, own_cu, K, L, i, j)
        c0 = c0s[L]
        own = (i == j == c0 + 2 and m0 == c0 and P == 2)  # the rows' own two columns
        assert own or m0 + P <= c0, ("panel used before it is final", T, own_cu, K, L, i, j, m0, P)
        assert L <= col_launch[j], ("tile updated after its column", T, own_cu, K, L, i, j)
        per_tile.setdefault((i, j), []).append((L, m0, P))
    for i in range(T):
        for j in range(i + 1):
            segs = per_tile.get((i, j), [])
            pos = 0
            for L, m0, P in segs:  # records come in launch order
                assert m0 == pos, ("gap or overlap", T, own_cu, K, i, j, segs)
                pos = m0 + P
            assert pos == c0s[col_launch[j]], ("incomplete", T, own_cu, K, i, j, segs)


@pytest.mark.parametrize("own_cu_roll", [4, 0, 2])
def test_one_per_cu_orders(own_cu_roll):
    """The one-workgroup-per-CU schedule: rolling far updates (the default,
    period 4; period 2) and every trailing tile in every launch (0)."""
    from pybrown_amd._lib import lib
    assert lib.pbx_set_potrf_own_cu_roll(own_cu_roll) == 0
    try:
        for T in list(range(1, 71)) + [113, 128, 129]:
            _check(T, 1, 1)
    finally:
        lib.pbx_set_potrf_own_cu_roll(4)


@pytest.mark.parametrize("K", [1, 2, 3, 4, 6])
def test_two_per_cu_orders(K):
    """Two workgroups per CU: far periods 1-6, rolling up to 128 tile
    columns and whole-trailing launches beyond."""
    for T in list(range(1, 71)) + [113, 127, 128, 129, 188, 191]:
        _check(T, 0, K)
