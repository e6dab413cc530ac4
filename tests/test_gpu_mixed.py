"""Mixed-precision noise factor: trailing updates on FP32 matrix cores
(operands rounded to FP32, FP32 accumulation added into the FP64 matrix),
diagonal tiles and solves in FP64.  Checked against numpy's FP64 Cholesky and
the FP64 path of this library; the displacement tolerance is the one stated
for the mixed engine (1e-6 relative per step)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _smith_D(N, phi, seed):
    from oracle.orc import Oracle
    from pybrown_amd.workloads import box_for, rsa_positions
    orc = Oracle()
    a = np.full(N, 10.0)
    L = box_for(a, phi)
    xyz = rsa_positions(a, L, seed)
    return orc.rpy_smith_m_matrix(a, Oracle.packed(orc.pointer_pbc_matrix(xyz, L)), L, math.sqrt(math.pi), 2, 2)


@pytest.mark.parametrize("N", [21, 100, 333, 1000])
def test_mixed_cholesky_against_fp64(N):
    from pybrown_amd.diffusion import cholesky
    D = _smith_D(N, 0.1, N)
    Ln = np.linalg.cholesky(D)
    Lm = cholesky(D, precision="mixed")
    L64 = cholesky(D)
    rel = np.linalg.norm(Lm - Ln) / np.linalg.norm(Ln)
    res = np.linalg.norm(Lm @ Lm.T - D) / np.linalg.norm(D)
    res64 = np.linalg.norm(L64 @ L64.T - D) / np.linalg.norm(D)
    print("N=%d mixed: factor rel %.3e residual %.3e (fp64 residual %.3e)" % (N, rel, res, res64))
    assert np.all(np.triu(Lm, 1) == 0)
    T = -(-3 * N // 64)
    if T <= 2:
        # no trailing update: the mixed factor is the FP64 one
        assert np.array_equal(Lm, L64)
    else:
        # the FP32 bulk is in use (an FP64 result here would mean it is not)
        assert rel > 1e-12
    assert rel <= 1e-5
    assert res <= 1e-6


def test_mixed_cholesky_reports_non_spd():
    from pybrown_amd.diffusion import cholesky
    A = np.eye(300)
    A[200, 200] = -1.0
    with pytest.raises(np.linalg.LinAlgError):
        cholesky(A, precision="mixed")


def test_mixed_step_displacement_error():
    """Per-step displacement of the mixed simulation against the FP64 one
    from the same positions and noise: relative error <= 1e-6."""
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    from pybrown_amd.workloads import box_for, rsa_positions
    N = 700
    a = np.full(N, 10.0)
    L = box_for(a, 0.1)
    xyz = rsa_positions(a, L, 5)
    inp = {"box_length": L, "T": 293.15, "viscosity": 0.01005, "seed": 7, "hydrodynamics": "rpy_smith",
           "ewald_alpha": math.sqrt(math.pi), "ewald_real": 2, "ewald_imag": 2, "check_overlaps": False}
    b64 = Box([Bead(p.copy(), r) for p, r in zip(xyz, a)], inp)
    bmx = Box([Bead(p.copy(), r) for p, r in zip(xyz, a)], inp, precision="mixed")
    try:
        errs = []
        for _ in range(4):
            start = b64._pos.copy()
            bmx._pos[:] = start
            b64.propagate(1.0)
            bmx.propagate(1.0)
            d64 = b64._pos - start
            dmx = bmx._pos - start
            errs.append(np.linalg.norm(dmx - d64) / np.linalg.norm(d64))
        r64, rmx = b64.factor_residual(), bmx.factor_residual()
        print("per-step displacement rel error %s; residual fp64 %.3e mixed %.3e" % (errs, r64, rmx))
        assert max(errs) <= 1e-6
        assert rmx <= 1e-6 and r64 <= 1e-14
        assert max(errs) > 0.0
    finally:
        b64.close()
        bmx.close()


def test_factor_residual_matches_numpy():
    from pybrown_amd.bead import Bead
    from pybrown_amd.box import Box
    from pybrown_amd.workloads import box_for, rsa_positions
    N = 150
    a = np.full(N, 10.0)
    L = box_for(a, 0.1)
    xyz = rsa_positions(a, L, 9)
    inp = {"box_length": L, "T": 293.15, "viscosity": 0.01005, "seed": 1, "hydrodynamics": "rpy_smith",
           "ewald_alpha": math.sqrt(math.pi), "ewald_real": 2, "ewald_imag": 2, "check_overlaps": False}
    box = Box([Bead(p.copy(), r) for p, r in zip(xyz, a)], inp, precision="mixed")
    try:
        box.propagate(1.0)
        D, B = box.D, box.B
        want = np.linalg.norm(B @ B.T - D) / np.linalg.norm(D)
        got = box.factor_residual()
        print("residual device %.6e numpy %.6e" % (got, want))
        assert abs(got - want) <= 1e-3 * want + 1e-18
    finally:
        box.close()
