"""The library's native noise stream against numpy.random.RandomState.

box.py:621-629 seeds ``RandomState(seed)`` with an integer and box.py:393-399
draws ``normal(0.0, 1.0, 3N)`` per move attempt; pbx_rng_normals must return
the same values bit for bit, however the stream is split.  Host code only (no
GPU needed)."""
import ctypes

import numpy as np
import pytest

from pybrown_amd._lib import lib


def _native(seed, skip, count):
    out = np.empty(count)
    rc = lib.pbx_rng_normals(seed, skip, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), count)
    assert rc == 0
    return out


@pytest.mark.parametrize("seed", [0, 1, 3, 12345, 2**31, 2**32 - 1])
def test_stream_matches_randomstate(seed):
    ref = np.random.RandomState(seed).normal(0.0, 1.0, 200_001)
    assert np.array_equal(_native(seed, 0, ref.size), ref)


@pytest.mark.parametrize("skip", [1, 2, 3, 624, 625, 9_999])
def test_split_stream(skip):
    # odd offsets start on the cached second value of a polar pair
    ref = np.random.RandomState(7).normal(0.0, 1.0, skip + 5_000)
    assert np.array_equal(_native(7, skip, 5_000), ref[skip:])


def test_chunked_calls_concatenate():
    # RandomState draws in chunks of 3N (one attempt) concatenate to one stream
    rs = np.random.RandomState(42)
    parts = np.concatenate([rs.normal(0.0, 1.0, 3 * 27) for _ in range(40)])
    assert np.array_equal(_native(42, 0, parts.size), parts)


def test_planted_seeding_mistake_is_caught():
    # a generator seeded one off differs immediately
    ref = np.random.RandomState(5).normal(0.0, 1.0, 16)
    assert not np.array_equal(_native(6, 0, 16), ref)


def test_bad_arguments():
    assert lib.pbx_rng_normals(0, -1, None, 0) != 0
    assert lib.pbx_rng_normals(0, 0, None, 4) != 0
