"""The restatement (oracle/liborc.so) against the reference's own C kernels:
live through oracle/_ref where it was built, and through the committed
fixtures written from it (tests/golden/make_ref_vectors.py) everywhere."""
import numpy as np
import pytest

from oracle.orc import Oracle, ref_available


def test_pointers_fixture(orc, ref_vectors):
    v = ref_vectors
    assert np.array_equal(orc.pointer_pbc_matrix(v["ptr_xyz"], float(v["ptr_L"])), v["ptr_P3"])


def test_immobile_pointers_fixture(orc, ref_vectors):
    v = ref_vectors
    got = orc.pointer_immobile_pbc_matrix(v["imm_xm"], v["imm_xi"], float(v["imm_L"]))
    assert np.array_equal(got, v["imm_P3"])


def test_immobile_pointers_closed_form(orc):
    # every component lies in (-L/2, L/2] and differs from the raw
    # difference by a whole number of boxes
    rng = np.random.RandomState(3)
    xm, xi, L = rng.uniform(-40, 40, (9, 3)), rng.uniform(-40, 40, (5, 3)), 13.0
    P = orc.pointer_immobile_pbc_matrix(xm, xi, L)
    raw = xi[None, :, :] - xm[:, None, :]
    k = (raw - P) / L
    assert np.allclose(k, np.round(k), atol=1e-12)
    assert np.all(P <= L / 2) and np.all(P > -L / 2)


def test_overlaps_fixture(orc, ref_vectors):
    v = ref_vectors
    got = [orc.check_overlaps(x, a, float(v["ovl_L"]), float(t)) for x, a, t in zip(v["ovl_xyz"], v["ovl_a"], v["ovl_thr"])]
    assert np.array_equal(np.array(got), v["ovl_out"])


def test_rpy_fixture_bit_exact(orc, ref_vectors):
    v = ref_vectors
    assert np.array_equal(orc.rpy_m_matrix(v["rpy_a"], v["rpy_P"]), v["rpy_M"])


@pytest.mark.parametrize("L", [20, 150])
@pytest.mark.parametrize("mn", [0, 1, 2, 3])
def test_smith_fixture_bit_exact(orc, ref_vectors, L, mn):
    v = ref_vectors
    t = "smith_L%d" % L
    M = orc.rpy_smith_m_matrix(v[t + "_a"], v[t + "_P"], float(L), np.sqrt(np.pi), mn, mn)
    assert np.array_equal(M, v[t + "_M%d" % mn])


def test_jo2b_fixture_bit_exact(orc, ref_vectors):
    v = ref_vectors
    for inp, out in zip(v["jo_in"], v["jo_out"]):
        assert np.array_equal(orc.jo_2b_r_matrix(inp[0], inp[1], inp[2:]), out)


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_live_reference_random_configs(orc):
    ref = Oracle("ref")
    rng = np.random.RandomState(7)
    for trial in range(6):
        N = rng.randint(2, 30)
        L = [20.0, 60.0, 437.6][trial % 3]
        xyz = rng.normal(0, 5, (N, 3)) if L == 20.0 else rng.uniform(-L / 2, L / 2, (N, 3))
        a = rng.choice([0.5, 1.0, 3.0], N) if L < 100 else np.full(N, 10.0)
        P3 = orc.pointer_pbc_matrix(xyz, L)
        assert np.array_equal(P3, ref.pointer_pbc_matrix(xyz, L))
        P = Oracle.packed(P3)
        assert np.array_equal(orc.rpy_m_matrix(a, P), ref.rpy_m_matrix(a, P))
        mn = trial % 3
        assert np.array_equal(orc.rpy_smith_m_matrix(a, P, L, 1.3, mn, mn + 1),
                              ref.rpy_smith_m_matrix(a, P, L, 1.3, mn, mn + 1))
        for thr in (0.0, 0.1, 0.5):
            assert orc.check_overlaps(xyz, a, L, thr) == ref.check_overlaps(xyz, a, L, thr)


@pytest.mark.skipif(not ref_available(omp=True), reason="oracle/_ref OpenMP build absent (reference sources absent)")
def test_openmp_reference_build_gives_the_same_bits():
    """The multi-core CPU baseline uses the reference's own parallel loop over
    bead rows (diff_tensor.c:185); each entry is computed by one thread from
    the same operands, so the matrix equals the serial build's bit for bit."""
    rng = np.random.RandomState(5)
    N, L = 60, 40.0
    xyz = rng.uniform(-L / 2, L / 2, (N, 3))
    a = rng.uniform(0.8, 1.6, N)
    serial, par = Oracle("ref"), Oracle("ref_omp")
    P = Oracle.packed(serial.pointer_pbc_matrix(xyz, L))
    for mn in (0, 2):
        M0 = serial.rpy_smith_m_matrix(a, P, L, np.sqrt(np.pi), mn, mn)
        M1 = par.rpy_smith_m_matrix(a, P, L, np.sqrt(np.pi), mn, mn)
        assert np.array_equal(M0, M1)
