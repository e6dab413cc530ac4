"""HIP kernels through the C ABI against the reference's own outputs
(tests/golden/ref_vectors.npz), the restatement oracle and numpy.linalg.

Tolerances: pointers, overlap flags -- bit-exact; mobility entries --
max|dM| / max|M| per 3x3 block row <= 1e-12; dense factorisations --
normwise relative 1e-12 (Cholesky) / 1e-11 (SPD inverse)."""
import math

import numpy as np
import pytest

from conftest import block_row_rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from pybrown_amd import _lib
    return _lib


def test_mfma_tile_layout(L):
    from pybrown_amd._lib import lib, ptr
    rng = np.random.RandomState(0)
    A = rng.normal(size=(128, 192))
    B = rng.normal(size=(64, 192))
    C = np.zeros((128, 64))
    assert lib.pbx_gemm_nt(ptr(A), ptr(B), ptr(C), 128, 64, 192) == 0
    np.testing.assert_allclose(C, A @ B.T, rtol=0, atol=1e-13 * np.abs(A @ B.T).max())


def test_pointers_bit_exact(ref_vectors):
    from pybrown_amd.bead import Bead, compute_pointer_pbc_matrix
    v = ref_vectors
    beads = [Bead(x, 1.0) for x in v["ptr_xyz"]]
    assert np.array_equal(compute_pointer_pbc_matrix(beads, float(v["ptr_L"])), v["ptr_P3"])


def test_immobile_pointers_bit_exact(ref_vectors):
    from pybrown_amd.bead import Bead, compute_pointer_immobile_pbc_matrix
    v = ref_vectors
    mob = [Bead(x, 1.0) for x in v["imm_xm"]]
    imm = [Bead(x, 1.0, mobile=False) for x in v["imm_xi"]]
    assert np.array_equal(compute_pointer_immobile_pbc_matrix(mob, imm, float(v["imm_L"])), v["imm_P3"])


def test_overlaps_exact(ref_vectors):
    from pybrown_amd.bead import Bead, check_overlaps
    v = ref_vectors
    for x, a, thr, want in zip(v["ovl_xyz"], v["ovl_a"], v["ovl_thr"], v["ovl_out"]):
        beads = [Bead(p, r) for p, r in zip(x, a)]
        got = check_overlaps(beads, float(v["ovl_L"]), float(thr), np.zeros((10, 10), dtype=np.int32))
        assert int(got) == int(want)


def _beads(xyz, a):
    from pybrown_amd.bead import Bead
    return [Bead(p, r) for p, r in zip(xyz, a)]


def test_rpy_vs_reference(ref_vectors):
    from pybrown_amd.bead import compute_pointer_pbc_matrix
    from pybrown_amd.diffusion import RPY_M_matrix
    v = ref_vectors
    beads = _beads(v["rpy_xyz"], v["rpy_a"])
    M = RPY_M_matrix(beads, compute_pointer_pbc_matrix(beads, float(v["rpy_L"])))
    err = block_row_rel_err(M, v["rpy_M"])
    print("RPY max block-row rel err %.3e, bit-exact entries %.4f" % (err, np.mean(M == v["rpy_M"])))
    assert err <= 1e-12


@pytest.mark.parametrize("Lbox", [20, 150])
@pytest.mark.parametrize("mn", [0, 1, 2, 3])
def test_smith_vs_reference(ref_vectors, Lbox, mn):
    from pybrown_amd.bead import compute_pointer_pbc_matrix
    from pybrown_amd.diffusion import RPY_Smith_M_matrix
    v = ref_vectors
    t = "smith_L%d" % Lbox
    beads = _beads(v[t + "_xyz"], v[t + "_a"])
    M = RPY_Smith_M_matrix(beads, compute_pointer_pbc_matrix(beads, float(Lbox)), float(Lbox), math.sqrt(math.pi),
                           mn, mn)
    err = block_row_rel_err(M, v[t + "_M%d" % mn])
    print("Smith L=%d m=n=%d err %.3e" % (Lbox, mn, err))
    assert err <= 1e-12


def test_smith_c2_density_vs_oracle(orc):
    from pybrown_amd.workloads import rsa_positions, box_for
    from pybrown_amd.bead import compute_pointer_pbc_matrix
    from pybrown_amd.diffusion import RPY_Smith_M_matrix
    from oracle.orc import Oracle
    a = np.full(150, 10.0)
    Lb = box_for(a, 0.05)
    xyz = rsa_positions(a, Lb, 5)
    beads = _beads(xyz, a)
    rij = compute_pointer_pbc_matrix(beads, Lb)
    M = RPY_Smith_M_matrix(beads, rij, Lb, math.sqrt(math.pi), 2, 2)
    Mo = orc.rpy_smith_m_matrix(a, Oracle.packed(rij), Lb, math.sqrt(math.pi), 2, 2)
    assert block_row_rel_err(M, Mo) <= 1e-12


def test_jo2b_vs_reference(ref_vectors):
    from pybrown_amd._lib import lib, ptr
    v = ref_vectors
    for inp, want in zip(v["jo_in"], v["jo_out"]):
        out = np.zeros(36)
        lib.JO_2B_R_matrix(*[float(x) for x in inp], ptr(out))
        got = out[:18]
        for b in range(3):
            den = np.abs(want[6 * b:6 * b + 6]).max()
            assert np.abs(got[6 * b:6 * b + 6] - want[6 * b:6 * b + 6]).max() <= 1e-12 * den


@pytest.mark.parametrize("cich", [True, False])
def test_lubrication_vs_oracle(orc, cich):
    from pybrown_amd.bead import compute_pointer_pbc_matrix
    from pybrown_amd.diffusion import JO_R_lubrication_correction_matrix
    from pybrown_amd.workloads import rsa_positions, box_for
    from oracle.orc import Oracle
    rng = np.random.RandomState(4)
    a = rng.choice([10.0, 20.0], 60)
    Lb = box_for(a, 0.3)
    xyz = rsa_positions(a, Lb, 9, min_gap=0.05)
    beads = _beads(xyz, a)
    rij = compute_pointer_pbc_matrix(beads, Lb)
    C = JO_R_lubrication_correction_matrix(beads, rij, 1.0, cich)
    Co = orc.jo_r_lubrication_correction_matrix(a, Oracle.packed(rij), 1.0, cich)
    assert np.count_nonzero(Co) > 0
    err = block_row_rel_err(C, Co + 1e-300)
    print("lubrication (cichocki=%s) err %.3e" % (cich, err))
    assert err <= 1e-11


def test_lubrication_accumulates(orc):
    from pybrown_amd._lib import lib, ptr
    from oracle.orc import Oracle
    xyz = np.array([[0.0, 0, 0], [2.3, 0, 0], [0, 2.2, 0.1]])
    a = np.ones(3)
    P = Oracle.packed(orc.pointer_pbc_matrix(xyz, 100.0))
    C0 = np.random.RandomState(1).normal(size=(9, 9))
    C = C0.copy().reshape(-1)
    lib.JO_R_lubrication_correction_matrix(ptr(a), ptr(P), 3, 5.0, 1, ptr(C))
    Co = orc.jo_r_lubrication_correction_matrix(a, P, 5.0, True, C=C0)
    np.testing.assert_allclose(C.reshape(9, 9), Co, rtol=1e-12, atol=1e-12 * np.abs(Co).max())


@pytest.fixture(params=[1, 2, 4, 6], ids=["potrf_v1", "potrf_v2", "potrf_v4", "potrf_two_col"])
def potrf_variant(request):
    from pybrown_amd._lib import lib
    assert lib.pbx_set_potrf_variant(request.param) == 0
    yield request.param
    lib.pbx_set_potrf_variant(0)


@pytest.mark.parametrize("n", [64, 100, 300, 1000, 3000])
def test_cholesky_vs_numpy(orc, n, potrf_variant):
    from pybrown_amd.diffusion import cholesky
    from pybrown_amd.workloads import rsa_positions, box_for
    from oracle.orc import Oracle
    # free-space RPY with minimum-image vectors is SPD only for small dilute
    # systems, so the large case uses a random SPD matrix
    N = n // 3 if (n % 3 == 0 and n <= 1000) else None
    if N is None:
        rng = np.random.RandomState(n)
        X = rng.normal(size=(n, n))
        D = X @ X.T / n + np.eye(n)
    else:
        a = np.full(N, 10.0)
        Lb = box_for(a, 0.05)
        xyz = rsa_positions(a, Lb, n)
        D = orc.rpy_m_matrix(a, Oracle.packed(orc.pointer_pbc_matrix(xyz, Lb)))
    Lg = cholesky(D)
    Ln = np.linalg.cholesky(D)
    rel = np.linalg.norm(Lg - Ln) / np.linalg.norm(Ln)
    print("cholesky n=%d rel %.3e" % (n, rel))
    assert rel <= 1e-12
    assert np.all(np.triu(Lg, 1) == 0)


@pytest.mark.parametrize("n,precision", [(300, "fp64"), (1000, "fp64"), (3000, "fp64"), (4500, "fp64"),
                                         (1000, "mixed"), (4500, "mixed")])
def test_two_col_chain_jobs_same_bits(n, precision):
    """Two-column Cholesky with the diagonal chain doing both helper jobs of
    every launch (the path taken when the helper workgroups are not yet
    running because other launches fill the compute units): the same factor
    to the bit as with the helpers doing them.  4500 takes the build with two
    workgroups per compute unit (more than 64 tile columns)."""
    from pybrown_amd._lib import lib
    from pybrown_amd.diffusion import cholesky
    rng = np.random.RandomState(n + 7)
    X = rng.normal(size=(n, n))
    D = X @ X.T / n + np.eye(n)
    assert lib.pbx_set_potrf_variant(6) == 0
    try:
        L_help = cholesky(D, precision)
        assert lib.pbx_set_potrf_chain_jobs(1) == 0
        try:
            L_chain = cholesky(D, precision)
        finally:
            lib.pbx_set_potrf_chain_jobs(0)
    finally:
        lib.pbx_set_potrf_variant(0)
    assert np.array_equal(L_help, L_chain)
    Ln = np.linalg.cholesky(D)
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert np.linalg.norm(L_chain - Ln) / np.linalg.norm(Ln) <= tol


def test_cholesky_reports_non_spd(potrf_variant):
    from pybrown_amd.diffusion import cholesky
    A = np.eye(130)
    A[70, 70] = -1.0
    with pytest.raises(np.linalg.LinAlgError):
        cholesky(A)


@pytest.mark.parametrize("n", [64, 200, 700, 3000, 7232])
def test_spd_solve_vs_numpy(n):
    """pbx_spd_solve: y = inv(L) b and x = inv(A) b from the factor and the two
    chained triangular solves, against numpy at 1e-12; repeated solves on the
    same hand-off words (new epochs) give the same bits."""
    from pybrown_amd._lib import lib, ptr
    rng = np.random.RandomState(n)
    X = rng.normal(size=(n, n))
    A = X @ X.T / n + np.eye(n)
    b = rng.normal(size=n)
    y1, x1, y4, x4 = (np.zeros(n) for _ in range(4))
    ms = np.zeros(1)
    assert lib.pbx_spd_solve(ptr(A), n, ptr(b), 1, ptr(y1), ptr(x1), ptr(ms)) == 0, lib.pbx_last_error()
    assert lib.pbx_spd_solve(ptr(A), n, ptr(b), 4, ptr(y4), ptr(x4), ptr(ms)) == 0, lib.pbx_last_error()
    Ln = np.linalg.cholesky(A)
    y_want = np.linalg.solve(Ln, b)
    x_want = np.linalg.solve(A, b)
    ry = np.linalg.norm(y1 - y_want) / np.linalg.norm(y_want)
    rx = np.linalg.norm(x1 - x_want) / np.linalg.norm(x_want)
    print("spd solve n=%d rel y %.3e x %.3e, %.1f us per pair of solves" % (n, ry, rx, 1e3 * ms[0]))
    assert ry <= 1e-12 and rx <= 1e-12
    assert np.array_equal(y1, y4) and np.array_equal(x1, x4)


@pytest.mark.parametrize("n", [64, 200, 700])
def test_spd_inverse_vs_numpy(n, potrf_variant):
    from pybrown_amd.diffusion import spd_inverse
    rng = np.random.RandomState(n)
    X = rng.normal(size=(n, n))
    A = X @ X.T / n + np.eye(n)
    got = spd_inverse(A, 2.5)
    want = 2.5 * np.linalg.inv(A)
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    print("spd inverse n=%d rel %.3e" % (n, rel))
    assert rel <= 1e-11
